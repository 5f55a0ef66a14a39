"""Synthetic, seeded inputs for the benchmark and test configurations (numpy, host side).

Everything here is set-up: coordinates with ghost points, initial primitive state Q and the inflow
profile.  Metrics and Jacobians are computed from the coordinates on the GPU by
`Flame3D.metrics` (f3d_metrics_f64), except for the Taylor-Green box whose identity metrics are
exact closed forms.

Arrays use the reference's column-major layout as C-order numpy arrays of shape
(nvar, Nz_tot, Ny_tot, Nx_tot) (see the package docstring).
"""
import math

import numpy as np

NG = 4

GAMMA, RG, CS, TS, PR = 1.4, 287.0, 1.458e-6, 110.4, 0.72


def sutherland(T):
    """mu(T) of viscous.jl:22 (SI Sutherland constants of run.jl:11-12)."""
    return CS * T * np.sqrt(T) / (T + TS)


def _prim_from(rho, u, v, w, p):
    T = p / (rho * RG)
    return np.stack([rho, u, v, w, p, T]).astype(np.float64)


def ghost_extrapolate(x, Nx, Ny, Nz):
    """Coordinate ghost points by linear reflection through the first/last interior point
    (parse_mesh.jl:77-136), in the reference's order: x-faces, y-faces, xy-corners, then z."""
    g = NG
    # 0-based: ghost i in [0,NG) <- 2*x[NG] - x[2NG - i]
    for i in range(g):
        x[:, :, i] = 2 * x[:, :, g] - x[:, :, 2 * g - i]
        x[:, :, Nx + 2 * g - 1 - i] = 2 * x[:, :, Nx + g - 1] - x[:, :, Nx + g - 1 - (g - i)]
    for j in range(g):
        x[:, j, :] = 2 * x[:, g, :] - x[:, 2 * g - j, :]
        x[:, Ny + 2 * g - 1 - j, :] = 2 * x[:, Ny + g - 1, :] - x[:, Ny + g - 1 - (g - j), :]
    for k in range(g):
        x[k] = 2 * x[g] - x[2 * g - k]
        x[Nz + 2 * g - 1 - k] = 2 * x[Nz + g - 1] - x[Nz + g - 1 - (g - k)]
    return x


# ----------------------------------------------------------------------------------------- TGV
def tgv(N=64, mach=0.1, Re=1600.0, T0=300.0, noise=0.0, jump=0.0, seed=1):
    """3D Taylor-Green vortex on [0, 2pi)^3 with N^3 cells, identity-direction metrics.

    Dimensional state keeping the Sutherland law unchanged: T0, c0 = sqrt(gamma Rg T0),
    U0 = mach*c0, L = 1 m, rho0 = Re*mu(T0)/(U0 L).  Optional seeded relative noise on every
    primitive and a planted pressure jump (factor 1+jump on the half-box i >= N/2) so that every
    scheme branch of the hybrid reconstruction fires.  Returns a dict with params kwargs, Q (with
    ghosts left zero: fill them with the ghost fill), metrics (9), J and the mesh points x, y, z.
    """
    h = 2 * math.pi / N
    Nt = N + 2 * NG
    c0 = math.sqrt(GAMMA * RG * T0)
    U0 = mach * c0
    rho0 = Re * sutherland(T0) / U0
    p0 = rho0 * RG * T0
    idx = (np.arange(Nt) - NG) * h
    z, y, x = np.meshgrid(idx, idx, idx, indexing="ij")
    u = U0 * np.sin(x) * np.cos(y) * np.cos(z)
    v = -U0 * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    p = p0 + rho0 * U0 ** 2 / 16 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2)
    rho = p / (RG * T0)
    if jump:
        sel = (np.arange(Nt)[None, None, :] >= NG + N // 2) & (np.arange(Nt)[None, None, :] < NG + N)
        sel = np.broadcast_to(sel, p.shape)
        p = np.where(sel, p * (1 + jump), p)
    Q = _prim_from(rho, u, v, w, p)
    if noise:
        rng = np.random.default_rng(seed)
        Q[:5] *= 1 + noise * (2 * rng.random(Q[:5].shape) - 1)
        Q[1:4] += noise * U0 * (2 * rng.random(Q[1:4].shape) - 1)
        Q[5] = Q[4] / (Q[0] * RG)
    Q[:, :NG] = 0
    Q[:, -NG:] = 0
    Q[:, :, :NG] = 0
    Q[:, :, -NG:] = 0
    Q[:, :, :, :NG] = 0
    Q[:, :, :, -NG:] = 0
    metrics = [np.zeros((Nt, Nt, Nt)) for _ in range(9)]
    for q in (0, 4, 8):
        metrics[q][:] = h * h
    J = np.full((Nt, Nt, Nt), 1.0 / h ** 3)
    dt_cfl = None
    return dict(
        params=dict(bc=[0] * 6, viscous=1), dims=(N, N, N), Q=Q, metrics=metrics, J=J, h=h, U0=U0,
        rho0=rho0, p0=p0, T0=T0, dt_hint=dt_cfl,
        x=np.ascontiguousarray(x), y=np.ascontiguousarray(y), z=np.ascontiguousarray(z),
    )


# ----------------------------------------------------------------------------------------- ramp
def tanh_faces(n, H, first):
    """Wall-normal face positions y_f(m), m = 0..n, tanh clustering at the wall with the first
    cell height `first`: y = H * (1 - tanh(b (1 - m/n)) / tanh(b))."""
    def first_of(b):
        return H * (1 - math.tanh(b * (1 - 1.0 / n)) / math.tanh(b))
    lo, hi = 1e-6, 20.0
    if first_of(lo) < first:  # uniform spacing already finer than asked
        return np.linspace(0, H, n + 1)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if first_of(mid) > first:
            lo = mid
        else:
            hi = mid
    b = 0.5 * (lo + hi)
    m = np.arange(n + 1) / n
    return H * (1 - np.tanh(b * (1 - m)) / np.tanh(b))


def ramp_state(mach=2.9, Re_delta=6.0e4, T_inf=108.1, T_wall=307.0, delta=5.0e-3):
    """Freestream of the compression-ramp cases (T_wall/T_inf = 2.84 as SCU-benchmark/flow-inlet.dat)."""
    c = math.sqrt(GAMMA * RG * T_inf)
    u = mach * c
    rho = Re_delta * sutherland(T_inf) / (u * delta)
    p = rho * RG * T_inf
    return dict(rho=rho, u=u, p=p, T=T_inf, Tw=T_wall, mach=mach, delta=delta, c=c)


def bl_profile(yn, st, mach=None, T_wall=None):
    """Laminar boundary-layer profile at wall distance yn: Pohlhausen velocity (2e - 2e^3 + e^4 for
    e = yn/delta < 1) and Crocco-Busemann temperature with an isothermal wall; constant p."""
    mach = st["mach"] if mach is None else mach
    Tw = st["Tw"] if T_wall is None else T_wall
    e = np.clip(yn / st["delta"], 0.0, 1.0)
    f = 2 * e - 2 * e ** 3 + e ** 4
    Tinf = st["T"]
    Tr = Tinf * (1 + math.sqrt(PR) * (GAMMA - 1) / 2 * mach ** 2)
    T = Tw + (Tr - Tw) * f + (Tinf - Tr) * f ** 2
    u = st["u"] * mach / st["mach"] * f
    rho = st["p"] / (RG * T)
    return rho, u, T


def ramp(Nx, Ny, Nz, angle=24.0, sweep=0.0, first=1e-3, seed=2, noise=0.01, st=None,
         Lu=20.0, Ld=20.0, H=8.0, Lz=4.0):
    """Compression ramp (lengths in units of delta) extruded spanwise, optionally swept by a
    spanwise shear of x (x += z*tan(sweep), as swept-compression-ramp.c:368), tanh wall-normal
    clustering with the first cell `first`*delta, laminar BL inflow with seeded uniform +-noise on
    the velocity (frozen in time), isothermal wall, Dirichlet farfield, zero-gradient outlet,
    periodic span.  Returns coordinates (with ghosts), Q, the inflow profile and params kwargs."""
    st = st or ramp_state()
    d = st["delta"]
    Nxt, Nyt, Nzt = Nx + 2 * NG, Ny + 2 * NG, Nz + 2 * NG
    tanA = math.tan(math.radians(angle))
    r = 1.0 * d
    xi = np.linspace(-Lu * d, Ld * d, Nx)
    yw = tanA * 0.5 * (xi + np.sqrt(xi * xi + r * r))  # smoothed corner
    yf = tanh_faces(Ny, H * d, first * d)
    yc = 0.5 * (yf[:-1] + yf[1:])
    zc = (np.arange(Nz) + 0.5) * (Lz * d / Nz)
    x = np.zeros((Nzt, Nyt, Nxt))
    y = np.zeros_like(x)
    z = np.zeros_like(x)
    X = np.broadcast_to(xi[None, None, :], (Nz, Ny, Nx))
    Y = yw[None, None, :] + yc[None, :, None]
    Z = np.broadcast_to(zc[:, None, None], (Nz, Ny, Nx))
    tanS = math.tan(math.radians(sweep))
    x[NG:-NG, NG:-NG, NG:-NG] = X + Z * tanS
    y[NG:-NG, NG:-NG, NG:-NG] = Y
    z[NG:-NG, NG:-NG, NG:-NG] = Z
    for a in (x, y, z):
        ghost_extrapolate(a, Nx, Ny, Nz)
    # initial state: the inflow BL profile at the local wall distance everywhere
    yn = y - (tanA * 0.5 * (x - z * tanS + np.sqrt((x - z * tanS) ** 2 + r * r)))
    rho, u, T = bl_profile(np.maximum(yn, 0.0), st)
    p = np.full_like(u, st["p"])
    Q = _prim_from(rho, u, np.zeros_like(u), np.zeros_like(u), p)
    inflow = inflow_profile(y[:, :, 0:NG].mean(axis=2), st, noise=noise, seed=seed)
    q_inf = [st["rho"], st["u"], 0.0, 0.0, st["p"], st["T"]]
    params = dict(bc=[5, 2, 4, 1, 0, 0], q_inf=q_inf, T_wall=st["Tw"], viscous=1)
    return dict(params=params, dims=(Nx, Ny, Nz), x=x, y=y, z=z, Q=Q, inflow=inflow, state=st)


def inflow_profile(ygh, st, noise=0.01, seed=2, mach=None, T_wall=None, fourier=None):
    """Inflow Dirichlet profile [6, Nz_tot, Ny_tot] from the wall distance of the inflow ghost
    column `ygh` (Nz_tot, Ny_tot): BL profile, seeded uniform +-noise on u and v (relative to u_inf),
    optional v-perturbation sum_m a_m sin(2 pi m z/Lz-ish index) given as amplitudes `fourier`."""
    rho, u, T = bl_profile(np.maximum(ygh, 0.0), st, mach=mach, T_wall=T_wall)
    Nzt, Nyt = ygh.shape
    rng = np.random.default_rng(seed)
    v = np.zeros_like(u)
    if noise:
        u = u * (1 + noise * (2 * rng.random(u.shape) - 1))
        v = v + noise * st["u"] * (2 * rng.random(u.shape) - 1)
    if fourier is not None:
        kz = np.arange(Nzt)[:, None] - NG
        for m, a in enumerate(fourier, start=1):
            v = v + a * st["u"] * np.sin(2 * math.pi * m * kz / max(Nzt - 2 * NG, 1)) * (u / st["u"])
    p = np.full_like(u, st["p"])
    return _prim_from(rho, u, v, np.zeros_like(u), p)


# ---------------------------------------------------------------- inflow-parameter adjoint case
N_FOURIER = 16


def theta0(seed=4, amp=2e-3):
    """Inflow parameters theta = (M_inf, T_wall, a_1..a_16) of the adjoint case: the ramp's
    freestream Mach number and wall temperature, and 16 seeded amplitudes of a spanwise Fourier
    perturbation of the inflow v (relative to u_inf)."""
    st = ramp_state()
    rng = np.random.default_rng(seed)
    return np.concatenate([[st["mach"], st["Tw"]], amp * (2 * rng.random(N_FOURIER) - 1)])


def ramp_theta(ygh, st, theta, noise=0.01, seed=2):
    """Boundary data of the ramp as a function of theta (complex-step safe: complex theta gives
    complex outputs).  The freestream thermodynamic state (rho, p, T) stays fixed; M_inf scales
    the inflow and farfield velocity, T_wall sets the isothermal wall and the Crocco-Busemann
    inflow temperature, a_m add a_m u_inf sin(2 pi m k/Nz) (u/u_inf) to the inflow v.
    Returns (inflow [6, Nz_tot, Ny_tot], q_inf [6], T_wall)."""
    mach, Tw, amps = theta[0], theta[1], theta[2:]
    e = np.clip(np.maximum(ygh, 0.0) / st["delta"], 0.0, 1.0)
    f = 2 * e - 2 * e ** 3 + e ** 4
    Tinf = st["T"]
    Tr = Tinf * (1 + math.sqrt(PR) * (GAMMA - 1) / 2 * mach ** 2)
    T = Tw + (Tr - Tw) * f + (Tinf - Tr) * f ** 2
    uinf = st["u"] * mach / st["mach"]
    u = uinf * f
    rho = st["p"] / (RG * T)
    Nzt, Nyt = ygh.shape
    rng = np.random.default_rng(seed)
    v = np.zeros(u.shape, dtype=u.dtype)
    if noise:
        u = u * (1 + noise * (2 * rng.random(u.shape) - 1))
        v = v + noise * st["u"] * (2 * rng.random(u.shape) - 1)
    kz = np.arange(Nzt)[:, None] - NG
    for m, a in enumerate(amps, start=1):
        v = v + a * uinf * np.sin(2 * math.pi * m * kz / max(Nzt - 2 * NG, 1)) * (u / uinf)
    p = np.full(u.shape, st["p"], dtype=u.dtype)
    inflow = np.stack([rho, u, v, np.zeros(u.shape, dtype=u.dtype), p, p / (rho * RG)])
    q_inf = np.array([st["rho"], uinf, 0.0, 0.0, st["p"], st["T"]], dtype=np.asarray(uinf).dtype)
    return inflow, q_inf, Tw


def ramp_theta_jacobian(ygh, st, theta, noise=0.01, seed=2, h=1e-30):
    """d(inflow, q_inf, T_wall)/d theta_k for every k by complex steps (exact to rounding)."""
    out = []
    for k in range(len(theta)):
        th = np.array(theta, dtype=np.complex128)
        th[k] += 1j * h
        inflow, q_inf, Tw = ramp_theta(ygh, st, th, noise=noise, seed=seed)
        out.append((inflow.imag / h, q_inf.imag / h, float(np.imag(Tw)) / h))
    return out
