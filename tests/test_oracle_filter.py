"""Pins of the oracle's filters (utils.jl:102-316, params filtering*) by closed forms on periodic boxes.

  * linearFilter along d multiplies a Fourier mode exp(i k x_d) by 1 - s0 D(k),
    D(k) = d0 + 2 sum_m d_m cos(m k) (D(0) = 0: constants kept; D(pi) = 1: the odd-even mode removed at
    s0 = 1); the x, y, z block multiplies a 3-D mode by the product of the three factors.
  * filter (nonlinear) with a uniform sensor sc = 1 multiplies the mode by
    1 - s0 (c1 (2 cos k - 2) + c2 (2 cos 2k - 2 cos k)); with sc = 0 it changes nothing; in flux form it
    conserves sum U on a periodic box for any sensor field from pre (which needs the ghost-layer sensor).
  * pre along d on an odd-even pressure p0 (1 + eps (-1)^i) gives r = 4 eps^2/(1 + eps (-1)^i)^2 + 1e-16
    and sc = max(0, 1 - rth/r); on a linear pressure r = 1e-16 and sc = 0.
  * The time loop filters after the steps whose index is a multiple of filtering_interval, and leaves Q
    as the last stage made it.
"""
import importlib

import numpy as np
import pytest

from casegen import relerr, tgv_case
from oracle_lib import Oracle

f3d = importlib.import_module("flame3d-mhd_amd")
NG = 4
DLIN = [0.243527493120, -0.204788880640, 0.120007591680, -0.045211119360, 0.008228661760]
C1, C2 = -0.210383, 0.039617
PER = [f3d.BC_PERIODIC] * 6


def _grid_index(p, d):
    """Absolute cell index along d, broadcast to the cell shape (numpy axes (k, j, i))."""
    shp = p.cell_shape()
    idx = np.arange(shp[2 - d], dtype=np.float64)
    sh = [1, 1, 1]
    sh[2 - d] = shp[2 - d]
    return np.broadcast_to(idx.reshape(sh), shp)


def _mode(p, ks, phase=0.3):
    """cos(sum_d k_d (i_d - NG) + phase) on every cell (ghosts included: the periodic images)."""
    arg = phase + sum(ks[d] * (_grid_index(p, d) - NG) for d in range(3))
    return np.cos(arg)


def _D_lin(k):
    return DLIN[0] + 2 * sum(DLIN[m] * np.cos(m * k) for m in range(1, 5))


def _D_nl(k):
    return C1 * (2 * np.cos(k) - 2) + C2 * (2 * np.cos(2 * k) - 2 * np.cos(k))


def _U_from(p, f):
    U = np.zeros(p.cell_shape(5))
    for n in range(5):
        U[n] = (1.0 + n) * f
    return U


INNER = (slice(None),) + (slice(NG, -NG),) * 3


def test_filter_coefficient_identities():
    assert abs(_D_lin(0.0)) < 1e-15 and abs(_D_lin(np.pi) - 1.0) < 1e-11
    assert abs(_D_nl(0.0)) < 1e-15 and abs(_D_nl(np.pi) - 1.0) < 1e-12


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("s0", [1.0, 0.3])
def test_linear_filter_fourier(d, s0):
    p = f3d.make_params(16, 12, 8, bc=PER)
    N = [16, 12, 8][d]
    o = Oracle(p)
    for m in (0, 1, 3, N // 2):
        k = 2 * np.pi * m / N
        ks = [0.0, 0.0, 0.0]
        ks[d] = k
        Un = _U_from(p, _mode(p, ks))
        U = np.full_like(Un, 9.0)
        o.filter_linear(d, U, Un, s0)
        ref = (1 - s0 * _D_lin(k)) * Un
        assert np.abs(U[INNER] - ref[INNER]).max() < 1e-13 * 5, (m, d)
        ghost = np.ones(U.shape, bool)
        ghost[INNER] = False
        assert np.all(U[ghost] == 9.0)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_nonlinear_filter_fourier_and_switch(d):
    p = f3d.make_params(16, 12, 8, bc=PER)
    N = [16, 12, 8][d]
    o = Oracle(p)
    one, zero = np.ones(p.cell_shape()), np.zeros(p.cell_shape())
    for m in (1, 2, N // 2):
        k = 2 * np.pi * m / N
        ks = [0.0, 0.0, 0.0]
        ks[d] = k
        Un = _U_from(p, _mode(p, ks))
        U = np.zeros_like(Un)
        o.filter_nonlinear(d, U, Un, one, 0.7)
        ref = (1 - 0.7 * _D_nl(k)) * Un
        assert np.abs(U[INNER] - ref[INNER]).max() < 5e-14, (m, d)
        o.filter_nonlinear(d, U, Un, zero, 0.7)
        assert np.array_equal(U[INNER], Un[INNER])


@pytest.mark.parametrize("d", [0, 1, 2])
def test_pre_sensor_closed_forms(d):
    p = f3d.make_params(10, 9, 8, bc=PER)
    o = Oracle(p)
    idx = _grid_index(p, d)
    eps, p0, rth = 5e-3, 2.0, 1e-5  # r = 1e-4: sc = 0.9, inside (0, 1)
    Q = np.ones(p.cell_shape(6))
    Q[4] = p0 * (1 + eps * (-1.0) ** idx)
    sc = np.full(p.cell_shape(), -5.0)
    o.pre_sensor(d, sc, np.ascontiguousarray(Q), rth)
    sgn = (-1.0) ** idx
    r = 4 * eps ** 2 / (1 + eps * sgn) ** 2 + 1e-16
    ref = np.maximum(0.0, 1 - rth / r)
    # written: the interior and one ghost layer each side along d
    sl = [slice(NG, -NG)] * 3
    sl[2 - d] = slice(NG - 1, -(NG - 1))
    sl = tuple(sl)
    assert np.abs(sc[sl] - ref[sl]).max() < 1e-12
    mask = np.ones(p.cell_shape(), bool)
    mask[sl] = False
    assert np.all(sc[mask] == -5.0)
    # linear pressure: all second differences vanish, r = 1e-16, sc = 0
    Q[4] = 1.0 + 0.01 * idx
    o.pre_sensor(d, sc, np.ascontiguousarray(Q), rth)
    assert np.all(sc[sl] == 0.0)
    # a pressure jump p0 -> p0 (1 + a) at index i0: r = (5/32) a^2 / (1 + a)^2 at the first high cell
    # and (5/32) a^2 at the last low one (the three second differences worked out by hand)
    a, i0 = 0.02, NG + 4
    Q[4] = np.where(idx >= i0, p0 * (1 + a), p0)
    o.pre_sensor(d, sc, np.ascontiguousarray(Q), rth)
    for at, r in ((i0, 5 / 32 * a * a / (1 + a) ** 2 + 1e-16), (i0 - 1, 5 / 32 * a * a + 1e-16)):
        pos = [NG + 1, NG + 1, NG + 1]
        pos[d] = at
        got = sc[pos[2], pos[1], pos[0]]
        assert abs(got - max(0.0, 1 - rth / r)) < 1e-13, (at, got)


def test_nonlinear_filter_conserves_on_periodic_box():
    """Flux form: with the sensor of a spiky pressure field (nonzero in places), sum_interior U is
    unchanged by each direction's filter; this needs the ghost-layer sensor to be the periodic image."""
    p = f3d.make_params(16, 12, 10, bc=PER)
    o = Oracle(p)
    rng = np.random.default_rng(3)
    Q = np.ones(p.cell_shape(6))
    base = 1.0 + 0.2 * rng.random((10, 12, 16))
    Q[4] = np.pad(base, NG, mode="wrap")  # periodic ghosts
    Un = np.zeros(p.cell_shape(5))
    for n in range(5):
        Un[n] = np.pad(rng.standard_normal((10, 12, 16)), NG, mode="wrap")
    for d in range(3):
        sc = np.zeros(p.cell_shape())
        o.pre_sensor(d, sc, np.ascontiguousarray(Q), 1e-5)
        assert 0.05 < np.mean(sc[NG:-NG, NG:-NG, NG:-NG] > 0) < 1.0
        U = np.zeros_like(Un)
        o.filter_nonlinear(d, U, Un, sc, 1.0)
        for n in range(5):
            s_in, s_out = Un[n][INNER[1:]].sum(), U[n][INNER[1:]].sum()
            assert abs(s_out - s_in) < 1e-12 * np.abs(Un[n]).sum(), (d, n)
        assert relerr(U[INNER], Un[INNER]) > 1e-3  # the filter acted


def test_linear_filter_block_3d_mode():
    p = f3d.make_params(16, 16, 16, bc=PER, filtering=1, filtering_s0=0.8)
    o = Oracle(p)
    ks = [2 * np.pi * 3 / 16, 2 * np.pi * 5 / 16, 2 * np.pi * 4 / 16]
    U = _U_from(p, _mode(p, ks))
    U0 = U.copy()
    Un = np.zeros_like(U)
    Q = np.ones(p.cell_shape(6))
    o.filter_block(U, Un, Q, np.zeros(p.cell_shape()))
    fac = np.prod([1 - 0.8 * _D_lin(k) for k in ks])
    # the x filter's output ghosts keep the unfiltered mode (only the interior is filtered): the y
    # filter then sees unfiltered ghosts, so compare where the y and z stencils stay inside
    sl = (slice(None), slice(NG + 4, -NG - 4), slice(NG + 4, -NG - 4), slice(NG, -NG))
    assert np.abs(U[sl] - fac * U0[sl]).max() < 1e-12


@pytest.mark.parametrize("nonlinear", [0, 1])
def test_time_loop_filters_on_the_interval(nonlinear):
    kw = dict(filtering_nonlinear=nonlinear, filtering_interval=2, filtering_s0=0.5)
    case = tgv_case(12, filtering=1, **kw)
    p = case.params
    plain = type(p).from_buffer_copy(p)
    plain.filtering = 0
    o, op = Oracle(p), Oracle(plain)
    dt = 0.2 * o.cfl_dt(case.Q, case.metrics, case.J, 1.0)
    U = case.U()
    # one step: no filtering yet (1 % 2 != 0)
    U1, Q1 = U.copy(), case.Q.copy()
    o.time_step(1, U1, np.zeros_like(U), Q1, case.metrics, case.J, dt)
    U1p, Q1p = U.copy(), case.Q.copy()
    op.time_step(1, U1p, np.zeros_like(U), Q1p, case.metrics, case.J, dt)
    assert np.array_equal(U1, U1p) and np.array_equal(Q1, Q1p)
    # two steps: the unfiltered two steps, then the filtering block; Q left as the last stage made it
    U2, Q2 = U.copy(), case.Q.copy()
    o.time_step(2, U2, np.zeros_like(U), Q2, case.metrics, case.J, dt)
    U2p, Q2p, Un = U.copy(), case.Q.copy(), np.zeros_like(U)
    op.time_step(2, U2p, Un, Q2p, case.metrics, case.J, dt)
    assert np.array_equal(Q2, Q2p)
    unfiltered = U2p.copy()
    o.filter_block(U2p, Un, Q2p, np.zeros(p.cell_shape()))
    assert np.array_equal(U2, U2p)
    assert relerr(U2[INNER], unfiltered[INNER]) > 1e-8  # the filter acted
