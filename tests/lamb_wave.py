"""A standing Lamb wave on a small periodic box: the closed-form check of the
acoustic coupling (test helper).

Over an isothermal reference state (T0 = theta0 pi0 constant) a vertically
uniform pi' = A cos(k x), with w = theta' = 0, is an exact linear mode of the
perturbation equations: u_t = -cp theta0 dpi'/dx and
pi'_t = kappa (-rho0 du/dx) with kappa = (Rd/cv) pi0/rho0 give
pi'_tt = (cp Rd/cv) T0 pi'_xx at every level, so pi' stays uniform in z and
the vertical pressure gradient, w and theta' stay zero.  On the C-grid the
semi-discrete frequency is w_sd = (2c/dx) sin(k dx/2), c^2 = cp Rd T0/cv;
with nothing but the fast terms acting, a step of the RK3 scheme (its last
stage integrates from s^n over the whole dt) is n_s forward-backward
substeps of dtau = dt/n_s, whose frequency is (2/dtau) asin(w_sd dtau/2).
"""
import math

import numpy as np

RD, CP, CV, G, P0 = 287.05, 1004.64, 717.63, 9.80665, 1e5


def lamb_case(vibe_mi, nx=32, ny=4, nz=6, dx=1000.0, ztop=3000.0, T0=280.0, amp=1e-5, dt=2.0, nsub=4):
    zeta = np.linspace(0.0, 1.0, nz + 1)
    zc = 0.5 * (zeta[1:] + zeta[:-1]) * ztop
    pi0 = np.exp(-G * zc / (CP * T0))
    th0 = T0 / pi0
    rho0 = P0 * pi0 ** (CP / RD) / (RD * T0)
    rep = lambda a: np.repeat(a, nx * ny)
    r0, t0, p0 = rep(rho0), rep(th0), rep(pi0)
    n = nx * ny * nz
    pi = amp * np.cos(2.0 * np.pi * (np.arange(n) % nx) / nx)
    kap = (RD / CV) * p0 / r0
    state = {"u": np.zeros(n), "v": np.zeros(n), "w": np.zeros(nx * ny * (nz + 1)), "rho": pi / kap,
             "theta": np.zeros(n), "pi": pi}
    return vibe_mi.Case("lamb_wave", nx, ny, nz, dx, dx, ztop, zeta, np.zeros(nx * ny), r0, t0, p0, np.zeros(nz),
                        state, dt, nsub, 60, eps_s=0.0)


def discrete_frequency(case, T0=280.0):
    c = math.sqrt(CP * RD / CV * T0)
    w_sd = 2.0 * c / case.dx * math.sin(math.pi / case.nx)
    dtau = case.dt / case.nsub
    return 2.0 / dtau * math.asin(w_sd * dtau / 2.0)


def cos_amplitude(pi, case, level=0):
    """Amplitude of cos(k x) in pi' at one level (mean over y)."""
    p = np.asarray(pi).reshape(case.nz, case.ny, case.nx)[level]
    ck = np.cos(2.0 * np.pi * np.arange(case.nx) / case.nx)
    return float(np.mean(p @ ck)) / (case.nx / 2.0)


def measured_frequency(series, dt):
    """w from a(t) = A cos(w t): cos(w dt) = sum a_n (a_{n+1} + a_{n-1}) / (2 sum a_n^2)."""
    a = np.asarray(series)
    return math.acos(float(np.sum(a[1:-1] * (a[2:] + a[:-2]))) / (2.0 * float(np.sum(a[1:-1] ** 2)))) / dt
