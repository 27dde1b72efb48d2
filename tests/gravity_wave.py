"""An internal gravity wave in a flat, constant-N, resting atmosphere: the
closed-form check of buoyancy and of the theta0-gradient term (test helper).

Linearised about rest, the perturbation equations of the step are (flat
ground, column depth H, periodic x, uniform in y, rigid lid and ground):

    u_t      = -cp theta0 d_x pi'
    w_t      = -cp theta0_w d_z pi' + g theta'_w / theta0_w
    rho'_t   = -d_x(rho0 u) - d_z(rho0_w w)
    theta'_t = -w dtheta0/dz
    pi'_t    = kappa rho'_t + lambda theta'_t        (EOS about the reference)

with kappa = (Rd/cv) pi0/rho0 and lambda = (Rd/cv) pi0/theta0.  On the C-grid a
horizontal Fourier mode exp(i k x) turns d_x into the factor
K = (2/dx) sin(k dx/2) (faces half a cell from the centres), and the vertical
operators are the staggered differences between levels and interfaces.  That
leaves a real linear system per mode in (i u^, w^, theta'^, pi'^) over the
column; rho' is passive (it feeds nothing back at linear order).  Its
eigenvalues are +-i omega: the acoustic pairs (omega >= c pi/H) and the
internal gravity waves (omega < N).  The gravest gravity mode, started from
its eigenvector, must advance at the eigen-frequency: a flipped buoyancy sign
makes the mode grow, a wrong dtheta0/dz coefficient moves its frequency.

The matrix below is assembled from the linear equations above (own code,
numpy eigenvalues); the continuous anelastic value
omega^2 = N^2 k^2 / (k^2 + m^2 + 1/(4 Hrho^2)) is a coarser cross-check.
The step's slow-then-fast stage split is first order in dt, so the measured
frequency is compared after Richardson extrapolation from dt and dt/2.
"""
import math

import numpy as np

RD, CP, CV, G, P0 = 287.05, 1004.64, 717.63, 9.80665, 1e5


def gw_case(vibe_mi, nx=16, ny=4, nz=12, dx=1000.0, ztop=6000.0, bv=0.01, theta_sfc=300.0, amp=1e-3, dt=6.0,
            nsub=6):
    """Flat constant-N reference state (the library's closed-form hydrostatic
    state), started from the gravest gravity eigenmode of amplitude amp K in
    theta'.  Returns (case, omega, mode) with the eigen-frequency omega."""
    from vibe_mi import configs
    zeta = np.linspace(0.0, 1.0, nz + 1)
    zs = np.zeros(nx * ny)
    r0, t0, p0 = configs._reference(1, theta_sfc, bv, nx, ny, nz, ztop, zeta, zs)
    lev = lambda a: a.reshape(nz, ny, nx)[:, 0, 0].copy()
    rho0, th0, pi0 = lev(r0), lev(t0), lev(p0)
    omega, vec = gravest_mode(nx, nz, dx, ztop, zeta, rho0, th0, pi0)
    Ut, W, Th, P = _split(vec, nz)
    x = np.arange(nx) * dx
    k = 2.0 * math.pi / (nx * dx)
    ph_c = np.exp(1j * k * x)               # cell centres
    ph_u = np.exp(1j * k * (x - 0.5 * dx))  # west faces
    scale = amp / np.max(np.abs(Th))
    Ut, W, Th, P = Ut * scale, W * scale, Th * scale, P * scale
    u_hat = -1j * Ut                        # the system carries i u^
    rep = lambda col, ph: np.real(col[:, None, None] * ph[None, None, :]) * np.ones((1, ny, 1))
    Wfull = np.concatenate([[0.0], W, [0.0]])
    th = rep(Th, ph_c).reshape(-1)
    pi = rep(P, ph_c).reshape(-1)
    kap = (RD / CV) * p0 / r0
    lam = (RD / CV) * p0 / t0
    state = {"u": rep(u_hat, ph_u).reshape(-1), "v": np.zeros(nx * ny * nz),
             "w": rep(Wfull, ph_c).reshape(-1), "theta": th, "pi": pi, "rho": (pi - lam * th) / kap}
    case = vibe_mi.Case("gravity_wave", nx, ny, nz, dx, dx, ztop, zeta, zs, r0, t0, p0, np.zeros(nz), state, dt,
                        nsub, 100, eps_s=0.0)
    return case, omega, {"theta": Th, "k": k}


def _split(vec, nz):
    return vec[:nz], vec[nz:2 * nz - 1], vec[2 * nz - 1:3 * nz - 1], vec[3 * nz - 1:]


def linear_operator(nx, nz, dx, ztop, zeta, rho0, th0, pi0):
    """Semi-discrete operator of one x-Fourier mode (k = 2 pi / (nx dx)),
    unknowns [i u^ (nz levels), w^ (interfaces 1..nz-1), theta'^ (nz),
    pi'^ (nz)]."""
    K = (2.0 / dx) * math.sin(math.pi / nx)
    dz = np.diff(zeta) * ztop                     # layer depths
    zc = 0.5 * (zeta[1:] + zeta[:-1]) * ztop
    dzw = np.diff(zc)                             # centre spacing, interfaces 1..nz-1
    th0w = 0.5 * (th0[1:] + th0[:-1])
    r0w = 0.5 * (rho0[1:] + rho0[:-1])
    gam = (th0[1:] - th0[:-1]) / dzw              # dtheta0/dz at interfaces 1..nz-1
    kap = (RD / CV) * pi0 / rho0
    lam = (RD / CV) * pi0 / th0
    nU, nW = nz, nz - 1
    oU, oW, oT, oP = 0, nU, nU + nW, nU + nW + nz
    n = oP + nz
    A = np.zeros((n, n))
    for k in range(nz):                           # (i u^)_t = cp theta0 K pi'^
        A[oU + k, oP + k] = CP * th0[k] * K
    for m in range(nW):                           # interface m+1 between levels m and m+1
        A[oW + m, oP + m + 1] += -CP * th0w[m] / dzw[m]
        A[oW + m, oP + m] += CP * th0w[m] / dzw[m]
        A[oW + m, oT + m] += 0.5 * G / th0w[m]
        A[oW + m, oT + m + 1] += 0.5 * G / th0w[m]
    for k in range(nz):
        # theta'_t = -(w_k gam_k + w_{k+1} gam_{k+1}) / 2 (w = 0 at ground and lid)
        row_t = np.zeros(n)
        if k >= 1:
            row_t[oW + k - 1] += -0.5 * gam[k - 1]
        if k <= nz - 2:
            row_t[oW + k] += -0.5 * gam[k]
        A[oT + k] += row_t
        # rho'_t = -rho0 K (i u^) - (r0w w)_{k+1} - (r0w w)_k) / dz
        row_r = np.zeros(n)
        row_r[oU + k] = -rho0[k] * K
        if k <= nz - 2:
            row_r[oW + k] += -r0w[k] / dz[k]
        if k >= 1:
            row_r[oW + k - 1] += r0w[k - 1] / dz[k]
        A[oP + k] = kap[k] * row_r + lam[k] * row_t
    return A


def gravest_mode(nx, nz, dx, ztop, zeta, rho0, th0, pi0):
    """(omega, eigenvector) of the fastest internal gravity wave (the
    gravest vertical mode) of linear_operator."""
    A = linear_operator(nx, nz, dx, ztop, zeta, rho0, th0, pi0)
    ev, V = np.linalg.eig(A)
    zc = 0.5 * (zeta[1:] + zeta[:-1]) * ztop
    nmax = math.sqrt(G * np.max(np.diff(th0) / np.diff(zc) / (0.5 * (th0[1:] + th0[:-1]))))
    cand = [(ev[i].imag, i) for i in range(len(ev)) if 0.0 < ev[i].imag < 1.5 * nmax]
    om, i = max(cand)
    assert abs(ev[i].real) <= 1e-12 * om, "gravity modes of the linear operator are neutral"
    return om, V[:, i]


def anelastic_frequency(nx, dx, ztop, bv, theta_sfc=300.0):
    """Continuous anelastic estimate N k / sqrt(k^2 + m^2 + 1/(4 Hrho^2)),
    m = pi / H, density scale height Hrho = Rd T / g at the surface."""
    k = 2.0 * math.pi / (nx * dx)
    m = math.pi / ztop
    hr = RD * theta_sfc / G
    return bv * k / math.sqrt(k * k + m * m + 0.25 / (hr * hr))


def mode_phase(theta, case, mode):
    """Complex amplitude of the mode in theta': projection on exp(-i k x) and
    the mode's vertical structure (mean over y)."""
    t = np.asarray(theta).reshape(case.nz, case.ny, case.nx).mean(axis=1)
    x = np.arange(case.nx) * case.dx
    c = t @ np.exp(-1j * mode["k"] * x) * (2.0 / case.nx)
    Th = mode["theta"]
    return complex(np.vdot(Th, c) / np.vdot(Th, Th))


def measured_frequency(phases, dt):
    """omega from the phase advance of the mode amplitude (least-squares
    slope of the unwrapped phase) and the amplitude's log growth rate."""
    ph = np.unwrap(np.angle(np.asarray(phases)))
    t = np.arange(len(ph)) * dt
    om = float(np.polyfit(t, ph, 1)[0])
    grow = float(np.polyfit(t, np.log(np.abs(phases)), 1)[0])
    return om, grow


def run_frequency(stepper, state_of, case, mode, dt, T):
    """Steps a stepper(dt) to time T, recording the mode's complex amplitude."""
    n = int(round(T / dt))
    ph = [mode_phase(state_of()["theta"], case, mode)]
    for _ in range(n):
        stepper(dt)
        ph.append(mode_phase(state_of()["theta"], case, mode))
    return measured_frequency(ph, dt)
