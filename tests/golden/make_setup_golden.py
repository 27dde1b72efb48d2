"""Writes tests/golden/setup_reference.json with the reference's own grid
generators, sponge profile, isothermal reference state, energy budget,
advective CFL number, stability report and nest operators.

oracle/ref.mk compiles the reference's src/grid/geometry.cpp,
src/dyn/damping.cpp, src/dyn/reference_state.cpp, src/dyn/diagnostics.cpp,
src/dyn/state.cpp, src/dyn/advection.cpp, src/time/timestep_control.cpp and
src/grid/nest.cpp (unmodified, from where
they lie under /root/reference) with oracle/ref_setup.cpp into
oracle/_ref/ref_setup, which prints the JSON (cases listed in
oracle/ref_setup.cpp).
Run:  make -C oracle -f ref.mk && python tests/golden/make_setup_golden.py
"""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", "ref_setup")


def main():
    assert os.path.exists(EXE), "build oracle/_ref first: make -C oracle -f ref.mk"
    out = subprocess.run([EXE], check=True, capture_output=True, text=True).stdout
    data = json.loads(out)
    path = os.path.join(HERE, "setup_reference.json")
    with open(path, "w") as fh:
        json.dump(data, fh)
    print(path, {k: len(v) for k, v in data.items()})


if __name__ == "__main__":
    main()
