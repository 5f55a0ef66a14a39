"""Pins of the oracle's WENO7 / WENO5 reconstructions (finite-difference schemes.jl:813-999 and
finite-volume FVM.jl:116-201) and of its SSP-RK3 stage blend (solver.jl:416-444), from the
mathematics that defines them rather than from the oracle's own formulas.

* The reconstruction on a stencil of cell averages is rebuilt here in exact rational arithmetic from
  its definition (Jiang & Shu 1996, Balsara & Shu 2000): each candidate is the value at the face of
  the polynomial whose cell averages match its sub-stencil; each smoothness indicator is
  sum_l int_cell dx^(2l-1) (d^l p / dx^l)^2 over the cell left of the face; the linear weights are
  the unique ones that turn the candidates into the reconstruction on the whole stencil.  Nothing
  below types a candidate, indicator or weight coefficient; the reference's integer indicators are
  240x (WENO7) and 12x (WENO5) the integral definition, which the epsilon-sensitive cases fix.
* Closed forms: a step between two constant states is reconstructed from the one-sided stencil
  (ENO selection); the observed order on smooth data without critical points is 7 (WENO7) and 5
  (WENO5), for the scalar reconstruction and through the whole face loops (characteristic
  finite-difference advect_c, finite-volume Eigen_/Conser_reconstruct).
* The RK3 time stepper converges at third order in dt on a smooth case (SSP-RK3, Shu & Osher
  1988), in the finite-difference step, the finite-volume step and the forward-mode (JVP) step.
"""
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

from casegen import f3d, tgv_case
from oracle_lib import Oracle, weno_scalar

NG = 4


# ------------------------------------------------------------------ exact rational definition
def _solve(A, b):
    n = len(A)
    M = [list(r) + [bb] for r, bb in zip(A, b)]
    for c in range(n):
        piv = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [M[r][n] / M[r][r] for r in range(n)]


def _poly_from_cell_averages(cells, vals):
    """Coefficients (monomial basis) of the polynomial of degree len(cells)-1 whose averages over the
    unit cells [m - 1/2, m + 1/2] are vals."""
    deg = len(cells)
    A = [[(Fr(2 * m + 1, 2) ** (q + 1) - Fr(2 * m - 1, 2) ** (q + 1)) / (q + 1) for q in range(deg)] for m in cells]
    return _solve(A, vals)


def _ev(c, x):
    return sum(ci * x ** i for i, ci in enumerate(c))


def _der(c):
    return [i * c[i] for i in range(1, len(c))]


def _indicator(c, a, b):
    """sum_{l>=1} int_a^b (d^l p/dx^l)^2 dx (unit spacing)."""
    s, d = Fr(0), c
    for _ in range(1, len(c)):
        d = _der(d)
        sq = [Fr(0)] * (2 * len(d) - 1)
        for i, x in enumerate(d):
            for j, y in enumerate(d):
                sq[i + j] += x * y
        s += sum(ci * (b ** (i + 1) - a ** (i + 1)) / (i + 1) for i, ci in enumerate(sq))
    return s


def _linear_weights(r):
    """Linear weights of the r candidates of a (2r-1)-cell stencil, face at the right edge of cell r."""
    n = 2 * r - 1
    xf = Fr(2 * r + 1, 2)

    def functional(cells):
        out = [Fr(0)] * n
        for m in cells:
            e = [Fr(int(mm == m)) for mm in cells]
            out[m - 1] = _ev(_poly_from_cell_averages(cells, e), xf)
        return out

    cand = [functional(list(range(k + 1, k + r + 1))) for k in range(r)]
    full = functional(list(range(1, n + 1)))
    d = []
    for k in range(r):  # cell k+1 lies in candidates 0..k only: triangular system
        d.append((full[k] - sum(d[j] * cand[j][k] for j in range(k))) / cand[k][k])
    for m in range(n):  # the weights reproduce the whole-stencil reconstruction on every cell
        assert sum(d[k] * cand[k][m] for k in range(r)) == full[m]
    return d


_D7, _D5 = _linear_weights(4), _linear_weights(3)


def weno_definition(V, r, eps, s0, scale):
    """Exact WENO reconstruction (Jiang-Shu weights, power 2) of the face at the right edge of cell r
    of the (2r-1)-cell stencil V; indicators multiplied by `scale` before eps is added."""
    V = [Fr(v) for v in V]
    d = _D7 if r == 4 else _D5
    xf = Fr(2 * r + 1, 2)
    q, beta = [], []
    for k in range(r):
        cells = list(range(k + 1, k + r + 1))
        c = _poly_from_cell_averages(cells, [V[m - 1] for m in cells])
        q.append(_ev(c, xf))
        beta.append(_indicator(c, Fr(2 * r - 1, 2), xf))
    alpha = [dk / (Fr(eps) + scale * b * Fr(s0)) ** 2 for dk, b in zip(d, beta)]
    return sum(a * qq for a, qq in zip(alpha, q)) / sum(alpha), beta


def test_linear_weights_derived():
    """The derived linear weights are the ones the reference's UP7 (run.jl:75-76) is the WENO7 limit of."""
    assert _D7 == [Fr(1, 35), Fr(12, 35), Fr(18, 35), Fr(4, 35)]
    assert _D5 == [Fr(1, 10), Fr(6, 10), Fr(3, 10)]
    up7 = [Fr(c, 420) for c in (-3, 25, -101, 319, 214, -38, 4)]
    # the whole-stencil reconstruction at the face from cells 1..7 (unit data per cell)
    for m in range(1, 8):
        e = [Fr(int(mm == m)) for mm in range(1, 8)]
        assert _ev(_poly_from_cell_averages(list(range(1, 8)), e), Fr(9, 2)) == up7[m - 1]


# (kind, stencil size r, return factor, eps used by the kernel or None = the argument, indicator scale)
KINDS = {
    "fdm_weno7": (0, 4, 12, None, 240),
    "fdm_weno5": (1, 3, 6, None, 12),
    "fvm_weno7": (2, 4, 1, 1e-10, 240),
    "fvm_weno5": (3, 3, 1, 1e-8, 12),
}


def _stencils(rng, r):
    out = []
    for _ in range(6):  # integer stencils
        out.append([int(x) for x in rng.integers(-20, 21, 2 * r - 1)])
    for sc in (1e-3, 1.0, 7.0e2):  # real stencils of several magnitudes
        for _ in range(3):
            out.append(list(sc * (2 * rng.random(2 * r - 1) - 1)))
    return out


def _oracle_args(kind, V):
    k, r = KINDS[kind][:2]
    if kind == "fvm_weno5":
        return [0.0] + list(V) + [0.0]  # the kernel reads V[1..5] of its 7-point stencil
    return list(V)


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("form", ["f64", "ld", "dual"])
def test_weno_equals_definition(kind, form):
    """The oracle's reconstruction equals the exact definition on random stencils, with the epsilon
    negligible (s0 = 1), dominant (the linear scheme) and comparable to the scaled indicators (which
    fixes the 240 / 12 scale of the integer indicators)."""
    k, r, fac, keps, scale = KINDS[kind]
    if form == "dual" and k >= 2:
        pytest.skip("the finite-volume reconstruction has no forward-mode form")
    rng = np.random.default_rng(100 + k)
    worst = 0.0
    for V in _stencils(rng, r):
        Vf = [Fr(v) for v in V]
        _, beta = weno_definition(Vf, r, 0.0, 1.0, scale)
        bmed = float(sorted(beta)[r // 2]) or 1.0
        cases = [(1.0, 1e-20)]
        if keps is None:
            cases += [(1.0, 1e30), (1e-20 / (scale * bmed), 1e-20), (2e-16 / (scale * bmed), 1e-16)]
        else:
            cases += [(keps / (scale * bmed), 0.0), (0.1 * keps / (scale * bmed), 0.0)]
        for s0, eps in cases:
            e_used = eps if keps is None else keps
            exact, _ = weno_definition(Vf, r, e_used, s0, scale)
            got = weno_scalar(k, form, _oracle_args(kind, V), s0, eps) / fac
            mag = max(abs(float(v)) for v in Vf) or 1.0
            err = abs(got - float(exact)) / mag
            worst = max(worst, err)
            assert err < 2e-13, (kind, form, V, s0, eps, got, float(exact))
    print(f"{kind} {form}: max |oracle - definition| / max|V| = {worst:.2e}")


@pytest.mark.parametrize("kind", list(KINDS))
def test_weno_float_form_tracks_definition(kind):
    """The FP32-bulk form (range-safe weights, first-difference indicators) stays within float
    rounding of the exact reconstruction."""
    k, r, fac, keps, scale = KINDS[kind]
    rng = np.random.default_rng(200 + k)
    for V in _stencils(rng, r):
        exact, _ = weno_definition([Fr(v) for v in V], r, keps or 1e-20, 1.0, scale)
        got = weno_scalar(k, "f32", _oracle_args(kind, V), 1.0, 1e-20) / fac
        mag = max(abs(v) for v in V) or 1.0
        assert abs(got - float(exact)) / mag < 2e-5, (kind, V, got, float(exact))


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("form", ["f64", "ld"])
def test_eno_selection_at_a_step(kind, form):
    """A step between two constant states: every candidate whose sub-stencil crosses the step gets a
    negligible weight, so the face value is the one-sided constant (Jiang-Shu's ENO property)."""
    k, r, fac, keps, scale = KINDS[kind]
    n = 2 * r - 1
    for jump_after in range(1, n):  # cells 1..jump_after hold a, the rest b
        for a, b in ((1.0, 2.0), (-3.0, 5.0), (0.25, -0.5)):
            V = [a if m <= jump_after else b for m in range(1, n + 1)]
            # the face is the right edge of cell r: the smooth sub-stencils are those not crossing
            smooth = [kk for kk in range(r) if not (kk + 1 <= jump_after < kk + r)]
            if not smooth:
                continue
            expect = a if all(kk + r <= jump_after for kk in smooth) else b
            got = weno_scalar(k, form, _oracle_args(kind, V), 1.0, 1e-20) / fac
            # weight left on a crossing stencil <= (d_max/d_min) (eps / (scale*beta_jump*s0))^2
            assert abs(got - expect) <= 1e-12 * max(abs(a), abs(b)), (kind, V, got, expect)


def _exp_cell_avg(a, h, m):
    """Average of exp(a x) over the cell of width h centred at m h."""
    return math.exp(a * m * h) * math.sinh(a * h / 2) / (a * h / 2)


@pytest.mark.parametrize("kind", list(KINDS))
def test_weno_order_on_smooth_data(kind):
    """Cell averages of exp(1.3 x) (no critical point): the face value converges at order 7 (WENO7)
    or 5 (WENO5); a wrong linear weight or candidate drops the order to that of one candidate."""
    k, r, fac, keps, scale = KINDS[kind]
    n = 2 * r - 1
    errs = []
    for N in (8, 16, 32):
        h = 1.0 / N
        V = [_exp_cell_avg(1.3, h, m) for m in range(1, n + 1)]
        got = weno_scalar(k, "ld", _oracle_args(kind, V), 1.0 / h ** 2, 1e-20) / fac
        errs.append(abs(got - math.exp(1.3 * (r + 0.5) * h)))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    lo, hi = (6.6, 7.6) if r == 4 else (4.7, 5.6)
    assert lo < orders[-1] < hi, (kind, errs, orders)


# ------------------------------------------------------------------ through the face loops
def _advect_exp_case(d, N, band, a=1.3, b=0.7):
    """Characteristic advect_c along direction d on a unit-spacing-in-index grid of length 1 (h = 1/N)
    with Fp = c_n exp(a x), Fm = e_n exp(-b x), a uniform state (constant frame) and the band forced;
    returns (face fluxes along the line, exact face values)."""
    dims = [4, 4, 4]
    dims[d] = N
    p = f3d.make_params(*dims, bc=[0] * 6, hybrid_phi1=0.0 if band != "up7" else 1e30,
                        hybrid_phi2=1e30 if band == "weno7" else 0.0,
                        hybrid_phi3=1e30 if band in ("weno7", "weno5") else 0.0)
    o = Oracle(p)
    h = 1.0 / N
    nt = N + 2 * NG
    x = (np.arange(nt) - NG) * h
    sh = [1, 1, 1]
    sh[2 - d] = nt
    xs = x.reshape(sh)
    cn = np.array([1.0, 0.4, -0.3, 0.2, 2.0])
    en = np.array([0.5, -0.2, 0.6, -0.1, 1.5])
    Fp = np.ascontiguousarray(np.stack([c * np.exp(a * xs) * np.ones(p.cell_shape()) for c in cn]))
    Fm = np.ascontiguousarray(np.stack([e * np.exp(-b * xs) * np.ones(p.cell_shape()) for e in en]))
    Q = np.zeros(p.cell_shape(6))
    for nq, val in enumerate((1.2, 30.0, -20.0, 10.0, 1.0e5, 1.0e5 / (1.2 * 287.0))):
        Q[nq] = val
    S = np.ones(p.cell_shape())
    A = [np.zeros(p.cell_shape()) for _ in range(3)]
    A[d][:] = 1.0
    phi = np.full(p.cell_shape(), 0.5)  # only its window maximum against the thresholds matters
    F = np.zeros(p.face_shape(d))
    o.advect(d, F, phi, S, Fp, Fm, Q, *A)
    nf = N + 1
    xf = (np.arange(nf) - 0.5) * h  # face f between cells f-1 and f (0-based interior)
    gp = (a * h / 2) / math.sinh(a * h / 2)
    gm = (b * h / 2) / math.sinh(b * h / 2)
    exact = np.stack([c * np.exp(a * xf) * gp + e * np.exp(-b * xf) * gm for c, e in zip(cn, en)])
    idx = [0, 0, 0]
    line = []
    for f in range(nf):
        idx[d] = f
        line.append(F[:, idx[2], idx[1], idx[0]])
    return np.array(line).T, exact


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("band", ["weno7", "weno5"])
def test_advect_c_forced_band_order(d, band):
    """The characteristic face loop with the WENO7 (WENO5) band forced converges at order 7 (5) to the
    exact face flux of smooth split fluxes, in every direction (stencil offsets of Fp and of the
    mirrored Fm, projection and back-projection)."""
    errs = []
    for N in (12, 24):
        F, exact = _advect_exp_case(d, N, band)
        errs.append(np.abs(F - exact).max() / np.abs(exact).max())
    order = math.log2(errs[0] / errs[1])
    lo = 6.5 if band == "weno7" else 4.7
    assert order > lo, (band, d, errs, order)


def _fvm_exp_case(d, N, band, eigen, a=1.1):
    from test_oracle_fvm import face_shape, mesh_points
    dims = [5, 5, 5]
    dims[d] = N
    p = f3d.make_params(*dims, bc=[0] * 6, hybrid_phi1=0.0, hybrid_phi2=1e30 if band == "weno7" else 0.0,
                        hybrid_phi3=1e30)
    h = 1.0 / N
    shp = p.cell_shape()
    idx = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shp), indexing="ij")
    xd = (idx[2 - d] - NG) * h
    avg = np.exp(a * xd) * math.sinh(a * h / 2) / (a * h / 2)
    base = np.array([1.0, 0.3, -0.2, 0.1, 2.5])
    amp = np.array([0.2, 0.05, 0.04, -0.03, 0.3])
    U = np.ascontiguousarray(np.stack([b0 + am * avg for b0, am in zip(base, amp)]))
    rho = U[0]
    u, v, w = U[1] / rho, U[2] / rho, U[3] / rho
    pr = 0.4 * (U[4] - 0.5 * rho * (u * u + v * v + w * w))
    Q = np.ascontiguousarray(np.stack([rho, u, v, w, pr, pr / (rho * p.Rg)]))
    S = np.ones(shp)
    phi = np.full(shp, 0.5)
    X = mesh_points(p, lambda i, j, k: [(i - NG) * h, (j - NG) * h, (k - NG) * h])
    F = np.zeros(face_shape(p, d, 5))
    UL, UR, geo = np.zeros_like(F), np.zeros_like(F), np.zeros(face_shape(p, d, 4))
    Oracle(p).fvm_reconstruct(d, eigen, Q, U, phi, S, F, *X, UL, UR, geo)
    nfd = dims[d] + 1
    xf = (np.arange(nfd) - 0.5) * h
    sh = [1, 1, 1]
    sh[2 - d] = nfd
    exact = np.stack([np.broadcast_to((b0 + am * np.exp(a * xf)).reshape(sh), F.shape[1:])
                      for b0, am in zip(base, amp)])
    return UL, UR, exact


@pytest.mark.parametrize("eigen", [True, False])
@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("band", ["weno7", "weno5"])
def test_fvm_reconstruct_forced_band_order(eigen, d, band):
    """Eigen_/Conser_reconstruct with the WENO7 (WENO5) band forced: both face states of smooth cell
    averages converge at order 7 (5)."""
    errs = []
    for N in (12, 24):
        UL, UR, exact = _fvm_exp_case(d, N, band, eigen)
        sc = np.abs(exact).max()
        errs.append((np.abs(UL - exact).max() / sc, np.abs(UR - exact).max() / sc))
    lo = 6.5 if band == "weno7" else 4.7
    for side in range(2):
        order = math.log2(errs[0][side] / errs[1][side])
        assert order > lo, (band, d, eigen, side, errs, order)


# ------------------------------------------------------------------ SSP-RK3 order in time
def _rk3_case(fvm=False):
    c = tgv_case(12, noise=0.0, jump=0.0, spikes=[], hybrid_phi1=1e30, fvm=f3d.FVM_EIGEN if fvm else 0)
    return c


def _march(c, dt, nsteps, fvm=False, jvp=False):
    p, o = c.params, Oracle(c.params)
    U = c.U()
    Q = c.Q.copy()
    if jvp:
        dU = np.zeros_like(U)
        o.time_step_jvp(nsteps, U, dU, c.metrics, c.J, dt)
        return U
    Un = np.zeros_like(U)
    if fvm:
        o.time_step_fvm(nsteps, U, Un, Q, c.metrics, c.J, c.coords, dt)
    else:
        o.time_step(nsteps, U, Un, Q, c.metrics, c.J, dt)
    return U


@pytest.mark.parametrize("path", ["fdm", "fvm", "jvp"])
def test_rk3_third_order_in_time(path):
    """Over a fixed time T the state error against a dt/64 solution falls 8x per halving of dt
    (third order: the stage blends 1 / (3/4, 1/4) / (1/3, 2/3) of SSP-RK3; any other consistent-looking
    blend changes the order, most lose consistency)."""
    fvm = path == "fvm"
    c = _rk3_case(fvm)
    o = Oracle(c.params)
    dt0 = o.cfl_dt(c.Q, c.metrics, c.J, 0.3)
    n0 = 4
    ref = _march(c, dt0 / 64, 64 * n0, fvm=fvm, jvp=path == "jvp")
    s = (slice(None),) + (slice(NG, -NG),) * 3
    errs = []
    for f in (1, 2, 4):
        U = _march(c, dt0 / f, f * n0, fvm=fvm, jvp=path == "jvp")
        errs.append(np.abs(U[s] - ref[s]).max() / np.abs(ref[s]).max())
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(2.8 < q < 3.2 for q in orders), (path, errs, orders)
