"""Writes tests/golden/equations_reference.npz with the reference's literal
discrete equations: Equations::tendencies, acoustic_tendencies,
max_sound_speed and diagnose of hashed states on an 8 x 6 x 5 grid (with and
without terrain) in three numerics configurations, plus the grid and
reference-state arrays they use.

oracle/ref.mk compiles the reference's src/dyn/equations.cpp (with
src/dyn/advection.cpp in the same translation unit), src/dyn/coriolis.cpp,
src/dyn/damping.cpp, src/dyn/tendency.cpp, src/dyn/state.cpp,
src/dyn/reference_state.cpp and src/grid/geometry.cpp, unmodified, from
where they lie under /root/reference, with oracle/ref_equations.cpp into
oracle/_ref/ref_equations, which prints the arrays as JSON (inputs and cases
are listed in oracle/ref_equations.cpp).
Run:  make -C oracle -f ref.mk && python tests/golden/make_equations_golden.py
"""
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", "ref_equations")


def main():
    assert os.path.exists(EXE), "build oracle/_ref first: make -C oracle -f ref.mk"
    out = subprocess.run([EXE], check=True, capture_output=True, text=True).stdout
    data = {k: np.asarray(v, dtype=np.float64) for k, v in json.loads(out).items()}
    path = os.path.join(HERE, "equations_reference.npz")
    np.savez_compressed(path, **data)
    print(path, len(data), "arrays", sum(v.size for v in data.values()), "values")


if __name__ == "__main__":
    main()
