"""CPU pins of the oracle: closed forms, invariants and known answers (no GPU).

Each check is chosen so that a plausible restatement mistake (dropped term, wrong sign, shifted
index, swapped operand) makes it fail.  The reference has no tests or golden vectors of its own.
"""
import math

import numpy as np
import pytest

from casegen import f3d, face_scales, ramp_case, rhs_scales, scaled_err, tgv_case
from oracle_lib import Oracle

NG = 4


def uniform_case(N=(12, 10, 8), h=0.1, q=(1.2, 30.0, -7.0, 3.0, 1.0e5, None), bc=None, viscous=1):
    nx, ny, nz = N
    p = f3d.make_params(nx, ny, nz, bc=bc or [0] * 6, viscous=viscous)
    Q = np.zeros(p.cell_shape(6))
    rho, u, v, w, pr, T = q
    T = pr / (rho * 287.0) if T is None else T
    for n, val in enumerate((rho, u, v, w, pr, T)):
        Q[n] = val
    m = [np.zeros(p.cell_shape()) for _ in range(9)]
    for k in (0, 4, 8):
        m[k][:] = h * h
    J = np.full(p.cell_shape(), 1.0 / h ** 3)
    return p, Q, m, J


# ----------------------------------------------------------------------- conversions
def test_c2prim_prim2c_round_trip_and_floors():
    c = tgv_case(12)
    p, o = c.params, Oracle(c.params)
    U = np.zeros(p.cell_shape(5))
    o.prim2c(U, c.Q)
    Q2 = np.zeros_like(c.Q)
    o.c2prim(U, Q2)
    s = (slice(None), slice(NG, -NG), slice(NG, -NG), slice(NG, -NG))
    assert np.max(np.abs(Q2[s] - c.Q[s]) / np.abs(c.Q[s]).max(axis=(1, 2, 3), keepdims=True)) < 1e-14
    # hand-evaluated cell: rho=2, rho u=(4,-2,6), E=50 -> u=(2,-1,3), e=50-0.5*2*14=36, p=0.4*36
    U[:] = 0
    U[0], U[1], U[2], U[3], U[4] = 2.0, 4.0, -2.0, 6.0, 50.0
    Q3 = np.zeros_like(c.Q)
    o.c2prim(U, Q3)
    cell = (slice(None), NG, NG, NG)
    assert np.allclose(Q3[cell], [2.0, 2.0, -1.0, 3.0, 0.4 * 36.0, 0.4 * 36.0 / (2.0 * 287.0)], rtol=1e-15)
    # eps floors of utils.jl:12,18: negative density and energy clamp to machine epsilon
    U[0] = -1.0
    U[4] = -5.0
    o.c2prim(U, Q3)
    eps = np.finfo(np.float64).eps
    assert Q3[0][NG, NG, NG] == eps and Q3[4][NG, NG, NG] == pytest.approx(0.4 * eps, rel=1e-15)


# ----------------------------------------------------------------------- sensor and mask
def numpy_sensor(P):
    phi = np.zeros_like(P)
    c = P[1:-1, 1:-1, 1:-1]
    tot = 0
    for ax in range(3):
        sl_m = [slice(1, -1)] * 3
        sl_p = [slice(1, -1)] * 3
        sl_m[ax] = slice(0, -2)
        sl_p[ax] = slice(2, None)
        a, b = P[tuple(sl_m)], P[tuple(sl_p)]
        tot = tot + np.abs(-a + 2 * c - b) / (a + 2 * c + b)
    phi[1:-1, 1:-1, 1:-1] = tot
    return phi


def test_sensor_constant_and_numpy():
    p, Q, m, J = uniform_case()
    o = Oracle(p)
    phi = np.ones(p.cell_shape())
    o.shock_sensor(phi, Q)
    inner = (slice(1, -1),) * 3
    assert np.all(phi[inner] == 0.0)
    rng = np.random.default_rng(4)
    Q[4] = 1e5 * (1 + 0.3 * rng.random(Q[4].shape))
    o.shock_sensor(phi, Q)
    assert np.allclose(phi[inner], numpy_sensor(Q[4])[inner], rtol=1e-15, atol=0)
    assert np.all(phi[0] == 1.0), "cells outside [2, N_tot-1] are left untouched (schemes.jl:6)"


def test_mask_band_around_spike():
    """A single pressure spike of factor f at cell c: phi(c) = 3(2f-2)/(2f+2); its 6 neighbours
    get (f-1)/(f+3); every face whose six-cell window [i-3, i+2] contains a nonzero phi takes the
    scheme of the largest one, all other faces are UP7."""
    f = 3.0
    p, Q, m, J = uniform_case(N=(16, 16, 16), q=(1.0, 0.0, 0.0, 0.0, 1.0, None))
    c = (NG + 7, NG + 6, NG + 8)  # (i, j, k) absolute
    Q[4][c[2], c[1], c[0]] *= f
    o = Oracle(p)
    phi = np.zeros(p.cell_shape())
    o.shock_sensor(phi, Q)
    assert phi[c[2], c[1], c[0]] == pytest.approx(3 * (2 * f - 2) / (2 * f + 2), rel=1e-15)
    assert phi[c[2], c[1], c[0] + 1] == pytest.approx((f - 1) / (f + 3), rel=1e-15)
    masks = [np.zeros(p.mask_shape(d), np.uint8) for d in range(3)]
    o.scheme_mask(*masks, phi)
    big = 3 * (2 * f - 2) / (2 * f + 2)   # 1.5 -> WENO5
    small = (f - 1) / (f + 3)             # 0.33 -> WENO7
    for d in range(3):
        exp = np.zeros(p.mask_shape(d), np.uint8)
        nf = exp.shape[::-1]  # (nx_faces, ny, nz) in i, j, k order
        for kk in range(nf[2]):
            for jj in range(nf[1]):
                for ii in range(nf[0]):
                    face = [ii + NG, jj + NG, kk + NG]
                    best = 0.0
                    for s in range(-3, 3):
                        cell = list(face)
                        cell[d] += s
                        delta = [cell[a] - c[a] for a in range(3)]
                        if delta == [0, 0, 0]:
                            best = max(best, big)
                        elif sum(abs(x) for x in delta) == 1:
                            best = max(best, small)
                    exp[kk, jj, ii] = 0 if best < 0.05 else (1 if best < 1 else (2 if best < 10 else 3))
        assert np.array_equal(masks[d], exp), f"direction {d}"
    assert (masks[0] == 2).sum() == 6 and (masks[0] == 1).sum() > 0


# ----------------------------------------------------------------------- SW splitting
def test_split_sums_to_euler_flux_and_supersonic_limit():
    c = ramp_case(16, 12, 8)
    p, o = c.params, Oracle(c.params)
    nc = c.J.size
    for d in range(3):
        A = c.metrics[3 * d:3 * d + 3]
        Fp = np.zeros(p.cell_shape(5))
        Fm = np.zeros_like(Fp)
        o.flux_split(c.Q, Fp, Fm, c.s[d], *A)
        rho, u, v, w, pr = c.Q[0], c.Q[1], c.Q[2], c.Q[3], c.Q[4]
        Uc = A[0] * u + A[1] * v + A[2] * w
        E = pr / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
        F = np.stack([rho * Uc, rho * u * Uc + pr * A[0], rho * v * Uc + pr * A[1], rho * w * Uc + pr * A[2],
                      (E + pr) * Uc])
        scale = np.abs(F).max(axis=(1, 2, 3)) + np.abs(pr * c.s[d]).max()
        err = np.abs(Fp + Fm - F).max(axis=(1, 2, 3)) / scale
        assert err.max() < 1e-13, (d, err)
    # supersonic along +A: all eigenvalues positive -> Fm == 0 exactly
    p2, Q, m, J = uniform_case(q=(1.0, 1000.0, 0.0, 0.0, 1.0e5, None))
    o2 = Oracle(p2)
    S = np.full(J.shape, 0.01)
    Fp = np.zeros(p2.cell_shape(5))
    Fm = np.zeros_like(Fp)
    o2.flux_split(Q, Fp, Fm, S, m[0], m[1], m[2])
    assert np.all(Fm == 0.0) and np.all(Fp[0] > 0)


# ----------------------------------------------------------------------- reconstruction
def _advect_all(p, Q, m, s, phi):
    o = Oracle(p)
    out = []
    for d in range(3):
        A = m[3 * d:3 * d + 3]
        Fp = np.zeros(p.cell_shape(5))
        Fm = np.zeros_like(Fp)
        o.flux_split(Q, Fp, Fm, s[d], *A)
        F = np.zeros(p.face_shape(d))
        o.advect(d, F, phi, s[d], Fp, Fm, Q, *A)
        out.append((F, Fp, Fm))
    return out


@pytest.mark.parametrize("band", [0.0, 0.5, 2.0, 20.0], ids=["UP7", "WENO7", "WENO5", "NND2"])
def test_uniform_state_reconstructs_exactly(band):
    """Uniform state: every scheme returns Fp + Fm of the uniform state (consistency; for the
    characteristic branches this needs L R = I)."""
    p, Q, m, J = uniform_case(q=(1.3, 45.0, -12.0, 8.0, 9.0e4, None))
    s = [np.full(J.shape, 0.01)] * 3
    phi = np.full(J.shape, band)
    for d, (F, Fp, Fm) in enumerate(_advect_all(p, Q, m, s, phi)):
        ref = (Fp + Fm)[:, NG, NG, NG]
        err = np.abs(F - ref[:, None, None, None]).max(axis=(1, 2, 3)) / np.abs(ref).max()
        assert err.max() < 1e-13, (band, d, err)


def test_up7_seventh_order():
    """UP7 flux difference of smooth Fp (Fm = 0) converges at 7th order to h*df/dx."""
    errs = []
    for N in (32, 64):
        p = f3d.make_params(N, 4, 4, bc=[0] * 6)
        o = Oracle(p)
        h = 2 * math.pi / N
        x = (np.arange(N + 2 * NG) - NG) * h
        Fp = np.zeros(p.cell_shape(5))
        Fm = np.zeros_like(Fp)
        Fp[:] = np.sin(x)[None, None, None, :]
        Q = np.ones(p.cell_shape(6))
        S = np.ones(p.cell_shape())
        phi = np.zeros(p.cell_shape())
        F = np.zeros(p.face_shape(0))
        o.advect(0, F, phi, S, Fp, Fm, Q, S, 0 * S, 0 * S)
        D = (F[0, 0, 0, 1:] - F[0, 0, 0, :-1]) / h  # cells 0..N-1
        errs.append(np.abs(D - np.cos(x[NG:NG + N])).max())
    order = math.log2(errs[0] / errs[1])
    assert 6.6 < order < 7.6, (errs, order)


def test_weno_linear_data_reduces_to_linear_schemes():
    """On linear data all smoothness indicators are equal, so WENO7 returns the optimal-weight
    scheme = UP7 and WENO5 the 5th-order upwind (2,-13,47,27,-3)/60 combination."""
    N = 16
    p = f3d.make_params(N, 4, 4, bc=[0] * 6)
    o = Oracle(p)
    x = np.arange(N + 2 * NG, dtype=float)
    Q = np.zeros(p.cell_shape(6))
    Q[0], Q[1], Q[4] = 1.0, 0.0, 1.0e5
    Q[5] = 1.0e5 / 287.0
    S = np.ones(p.cell_shape())
    z = np.zeros(p.cell_shape())
    Fp = np.zeros(p.cell_shape(5))
    Fm = np.zeros_like(Fp)
    for n in range(5):
        Fp[n] = (1.0 + n) + (0.5 + 0.1 * n) * x[None, None, :]
        Fm[n] = -(2.0 - n) + (0.3 - 0.05 * n) * x[None, None, :]
    Fs = {}
    for name, band in (("up7", 0.0), ("weno7", 0.5), ("weno5", 2.0)):
        F = np.zeros(p.face_shape(0))
        o.advect(0, F, np.full(p.cell_shape(), band), S, Fp, Fm, Q, S, z, z)
        Fs[name] = F
    scale = np.abs(Fs["up7"]).max()
    assert np.abs(Fs["weno7"] - Fs["up7"]).max() / scale < 1e-12
    # WENO5 on linear data: linear reconstruction is exact for linear data, as is UP7
    assert np.abs(Fs["weno5"] - Fs["up7"]).max() / scale < 1e-12


def test_nnd_minmod_truth_table():
    N = 12
    p = f3d.make_params(N, 4, 4, bc=[0] * 6)
    o = Oracle(p)
    Fp = np.zeros(p.cell_shape(5))
    Fm = np.zeros_like(Fp)
    # monotone increasing Fp: minmod picks the smaller slope; an extremum gives zero slope
    vals = np.array([0, 1, 3, 4, 8, 9, 9, 7, 7, 6, 2, 1, 0, 0, 0, 0, 0, 0, 0, 0], dtype=float)
    Fp[0] = vals[None, None, :]
    Q = np.ones(p.cell_shape(6))
    S = np.ones(p.cell_shape())
    z = np.zeros(p.cell_shape())
    F = np.zeros(p.face_shape(0))
    o.advect(0, F, np.full(p.cell_shape(), 20.0), S, Fp, Fm, Q, S, z, z)

    def mm(a, b):
        return (b if abs(a) > abs(b) else a) if a * b > 0 else 0.0

    for f in range(N + 1):
        i = NG + f
        exp = vals[i - 1] + 0.5 * mm(vals[i] - vals[i - 1], vals[i - 1] - vals[i - 2])
        assert F[0, 0, 0, f] == pytest.approx(exp, abs=1e-15)


@pytest.mark.parametrize("band", [0.5, 2.0], ids=["WENO7", "WENO5"])
def test_uniform_state_affine_metrics_both_tangent_branches(band):
    """Uniform state on an affine (sheared) grid: constant metric rows whose normals take both
    tangent constructions of schemes.jl:772-782 (|n_z| <= |n_y| and the other one); the
    characteristic reconstruction must still return Fp + Fm exactly (L R = I for any frame)."""
    p, Q, m, J = uniform_case(q=(1.1, 80.0, -25.0, 40.0, 8.0e4, None))
    rows = [(0.9, 0.2, 0.5), (0.1, 0.3, 1.2), (-0.4, 1.1, 0.2)]  # xi: |nz|>|ny|; eta: |nz|>|ny|; zeta: |nz|<=|ny|
    for d, r in enumerate(rows):
        for c in range(3):
            m[3 * d + c][:] = r[c] * 0.01
    s = [np.full(J.shape, 0.01 * math.sqrt(sum(x * x for x in r))) for r in rows]
    phi = np.full(J.shape, band)
    for d, (F, Fp, Fm) in enumerate(_advect_all(p, Q, m, s, phi)):
        ref = (Fp + Fm)[:, NG, NG, NG]
        err = np.abs(F - ref[:, None, None, None]).max(axis=(1, 2, 3)) / np.abs(ref).max()
        assert err.max() < 1e-13, (band, d, err)


# ----------------------------------------------------------------------- viscous and div
def test_viscous_linear_velocity_closed_form():
    """Constant T and linear velocity on a Cartesian grid: tau is constant, so the momentum viscous
    RHS vanishes and the energy RHS equals the dissipation tau:grad(u)."""
    h = 0.01
    p, Q, m, J = uniform_case(N=(10, 10, 10), h=h, q=(1.0, 0.0, 0.0, 0.0, 1.0e5, 300.0))
    nz, ny, nx = p.cell_shape()
    z, y, x = np.meshgrid(np.arange(nz) * h, np.arange(ny) * h, np.arange(nx) * h, indexing="ij")
    G = np.array([[1.0, 2.0, -0.5], [0.3, -1.2, 0.7], [-0.4, 0.9, 0.2]]) * 100.0
    Q[1] = G[0, 0] * x + G[0, 1] * y + G[0, 2] * z
    Q[2] = G[1, 0] * x + G[1, 1] * y + G[1, 2] * z
    Q[3] = G[2, 0] * x + G[2, 1] * y + G[2, 2] * z
    o = Oracle(p)
    Fv = [np.zeros(p.fv_shape()) for _ in range(3)]
    o.viscous(*Fv, Q, m, J)
    U = np.zeros(p.cell_shape(5))
    F0 = [np.zeros(p.face_shape(d)) for d in range(3)]
    o.div(U, *F0, *Fv, 1.0, J)
    mu = 1.458e-6 * 300.0 ** 1.5 / (300.0 + 110.4)
    div = np.trace(G)
    tau = mu * (G + G.T - 2.0 / 3.0 * div * np.eye(3))
    diss = np.sum(tau * G)
    s = (slice(NG, -NG),) * 3
    for n in (1, 2, 3):
        assert np.abs(U[n][s]).max() < 1e-9 * abs(diss)
    assert np.allclose(U[4][s], diss, rtol=1e-9)


def test_periodic_conservation_telescopes():
    """All-periodic box: sum over cells of RHS/J vanishes (flux differences telescope)."""
    c = tgv_case(16)
    p, o = c.params, Oracle(c.params)
    R = np.zeros(p.cell_shape(5))
    o.rhs(R, c.Q, c.metrics, c.J)
    A = np.zeros_like(R)
    o.rhs_absterms(A, c.Q, c.metrics, c.J)
    s = (slice(NG, -NG),) * 3
    for n in range(5):
        tot = np.sum(R[n][s] / c.J[s])
        mag = np.sum(A[n][s] / c.J[s])
        assert abs(tot) < 1e-13 * mag, (n, tot, mag)


def test_free_stream_rhs_is_zero():
    p, Q, m, J = uniform_case(q=(0.9, 120.0, 35.0, -20.0, 7.0e4, None))
    o = Oracle(p)
    R = np.ones(p.cell_shape(5))
    o.rhs(R, Q, m, J)
    s = (slice(None),) + (slice(NG, -NG),) * 3
    assert np.all(R[s] == 0.0)


# ----------------------------------------------------------------------- ghost cells
def test_ghost_fill_index_identities():
    p = f3d.make_params(8, 6, 5, bc=[1, 2, 4, 1, 0, 0], T_wall=300.0,
                        q_inf=[1.1, 2.0, 0.0, 0.0, 1.0e5, 1.0e5 / (1.1 * 287.0)])
    o = Oracle(p)
    nz, ny, nx = p.cell_shape()
    lab = np.arange(nz * ny * nx, dtype=float).reshape(nz, ny, nx) + 1000.0
    Q = np.stack([lab + 1e6 * n for n in range(6)])
    Q[4] = 1.0e5 + lab
    Q[5] = 200.0 + lab * 1e-3
    U = np.zeros(p.cell_shape(5))
    o.fill(Q, U)
    Np = (8, 6, 5)
    # x low: Dirichlet q_inf (fill_y and fill_z later overwrite the edge rows); x high: copy of the
    # last interior layer
    assert np.all(Q[:, NG:-NG, NG:-NG, :NG] == np.array(p.q_inf)[:, None, None, None])
    for g in range(NG):
        assert np.array_equal(Q[:, :, :, NG + Np[0] + g], Q[:, :, :, NG + Np[0] - 1])
    # z periodic: low ghost g <- g + Nzp (applied last, so it holds everywhere)
    for g in range(NG):
        assert np.array_equal(Q[:, g], Q[:, g + Np[2]])
        assert np.array_equal(Q[:, NG + Np[2] + g], Q[:, NG + g])
    # y low isothermal wall: mirror 2NG-1-g, velocity negated, T = 2Tw - T_m, rho = p/(Rg T)
    k, i = NG + 2, NG + 3
    for g in range(NG):
        mi = 2 * NG - 1 - g
        assert Q[1, k, g, i] == -Q[1, k, mi, i] and Q[4, k, g, i] == Q[4, k, mi, i]
        assert Q[5, k, g, i] == pytest.approx(600.0 - Q[5, k, mi, i], rel=1e-15)
        assert Q[0, k, g, i] == pytest.approx(Q[4, k, g, i] / (287.0 * Q[5, k, g, i]), rel=1e-15)
    # y high Dirichlet
    assert np.all(Q[:, NG:-NG, NG + Np[1]:, NG:-NG] == np.array(p.q_inf)[:, None, None, None])
    # U of a Dirichlet ghost is prim2c of q_inf
    qi = p.q_inf
    assert U[4][NG, NG, 0] == pytest.approx(qi[4] / 0.4 + 0.5 * qi[0] * qi[1] ** 2, rel=1e-15)


def test_exchange_periodic_slabs():
    p = f3d.make_params(6, 5, 7, bc=[0] * 6)
    o = Oracle(p)
    Q = np.random.default_rng(9).random(p.cell_shape(6))
    o.exchange(Q, 6)
    for g in range(NG):
        assert np.array_equal(Q[:, :, :, g], Q[:, :, :, g + 6])
        assert np.array_equal(Q[:, :, g, :], Q[:, :, g + 5, :])
        assert np.array_equal(Q[:, g], Q[:, g + 7])
        assert np.array_equal(Q[:, NG + 7 + g], Q[:, NG + g])


# ----------------------------------------------------------------------- dt, metrics
def test_cfl_dt_uniform_flow():
    h = 0.02
    p, Q, m, J = uniform_case(h=h, q=(1.0, 50.0, 0.0, 0.0, 1.0e5, 300.0))
    c = math.sqrt(1.4 * 287.0 * 300.0)
    assert Oracle(p).cfl_dt(Q, m, J, 0.5) == pytest.approx(0.5 * h / (50.0 + c), rel=1e-14)


def test_metrics_affine_grid_exact():
    rng = np.random.default_rng(3)
    M = np.eye(3) * 0.1 + 0.03 * rng.standard_normal((3, 3))
    b = rng.standard_normal(3)
    p = f3d.make_params(6, 5, 4, bc=[0] * 6)
    nz, ny, nx = p.cell_shape()
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X = [np.ascontiguousarray(M[c, 0] * ii + M[c, 1] * jj + M[c, 2] * kk + b[c]) for c in range(3)]
    m = [np.zeros(p.cell_shape()) for _ in range(9)]
    J = np.zeros(p.cell_shape())
    Oracle(p).metrics(*X, m, J)
    cof = np.linalg.inv(M) * np.linalg.det(M)  # J^-1 scaled metrics: d(xi_a)/d(x_c) / J = cof
    for a in range(3):
        for c in range(3):
            assert np.allclose(m[3 * a + c], cof[a, c], rtol=0, atol=1e-14), (a, c)
    assert np.allclose(J, 1.0 / np.linalg.det(M), rtol=1e-13)


# ----------------------------------------------------------------------- rounding floor of the bar
def test_fp64_oracle_rounding_floor_below_parity_bar():
    """The FP64 oracle against the same code in long double, normed by the face-flux scale: its
    own rounding floor sits well below the 1e-12 parity bar used for the GPU."""
    for c in (tgv_case(16, hybrid_phi3=2.5), ramp_case(20, 12, 10, sweep=30.0, jump=0.5)):
        p, o = c.params, Oracle(c.params)
        phi = np.zeros(p.cell_shape())
        o.shock_sensor(phi, c.Q)
        Fd, Fl = [], []
        for d in range(3):
            A = c.metrics[3 * d:3 * d + 3]
            Fp = np.zeros(p.cell_shape(5))
            Fm = np.zeros_like(Fp)
            o.flux_split(c.Q, Fp, Fm, c.s[d], *A)
            a = np.zeros(p.face_shape(d))
            b = np.zeros_like(a)
            o.advect(d, a, phi, c.s[d], Fp, Fm, c.Q, *A)
            o.advect_ld(d, b, phi, c.s[d], Fp, Fm, c.Q, *A)
            Fd.append(a)
            Fl.append(b)
        sc = face_scales(Fl)
        # measured: <= 5.5e-13 on the TGV case (z faces), ~1e-14 on the swept ramp
        assert max(scaled_err(Fd[d], Fl[d], sc) for d in range(3)) < 1e-12
