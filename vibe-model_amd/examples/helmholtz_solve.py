"""One 3-D Helmholtz solve per solver kind on the GPU, the workload of the
committed Helmholtz kernel profiles:

  python vibe-model_amd/examples/helmholtz_solve.py [nx ny nz]   (default 512 512 64)

a = (dt/2)^2 (Rd/cv) pi0 cp theta0 for dt = 6 s (semi_implicit.cpp:66-72
of the reference) with theta0 in 290..330 K, b = 1, seeded random terrain up
to 800 m, geometric 1.08 stretching, tol 1e-10; prints iterations, relative
residual and wall time of each solve (host copies included)."""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
import vibe_mi  # noqa: E402


def main(argv):
    nx, ny, nz = (int(v) for v in argv[1:4]) if len(argv) >= 4 else (512, 512, 64)
    rng = np.random.default_rng(3)
    d = 1.08 ** np.arange(nz)
    zeta = np.r_[0.0, np.cumsum(d)] / np.sum(d)
    zs = rng.uniform(0, 800, (ny, nx))
    a = 9.0 * 0.4 * 0.9 * 1004.64 * rng.uniform(290, 330, (nz, ny, nx))
    f = rng.normal(size=(nz, ny, nx))
    h = vibe_mi.Helmholtz(nx, ny, nz, 1000.0, 1000.0, 20000.0, zeta, zs)
    h.set_coefficients(a, np.ones((nz, ny, nx)))
    for kind, name in ((vibe_mi.HELM_BICGSTAB, "fp64 bicgstab"), (vibe_mi.HELM_REFINE, "fp32 + refinement")):
        t0 = time.perf_counter()
        h.solve(f, kind=kind, tol=1e-10, max_iter=500)
        ms = (time.perf_counter() - t0) * 1e3
        print(f"helmholtz {nx}x{ny}x{nz} {name}: {h.iterations} iterations, rel residual {h.residual_rel:.2e}, "
              f"{ms:.1f} ms", flush=True)
    h.close()


if __name__ == "__main__":
    main(sys.argv)
