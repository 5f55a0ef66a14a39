"""Pins of the oracle's LES eddy viscosity (viscous.jl:66-106, params les) against the textbook models.

On an affine mesh x = A (i, j, k) with linear velocity u = G x + u0 and temperature T = T0 + t.x the
4th-order central differences are exact, so every cell's viscous flux has a closed form:
  tau = mu_e (G + G^T - 2/3 tr(G) I),  q = lam_e grad T,  Fv_d = M_d [tau; u.tau + q]
with M_d the J^-1-scaled metric row (rows of det(A) A^-1), mu_e = mu + mu_t, lam_e = lam + Cp mu_t / 0.9,
mu_t = min(rho (C Delta)^2 D, 2 mu), Delta = J^(-1/3) = det(A)^(1/3), and
  Smagorinsky (C = 0.1):  D = sqrt(2 S:S),  S = (G + G^T)/2;
  WALE (C = 0.325, Nicoud & Ducros 1999):  D = (Sd:Sd)^(3/2) / ((S:S)^(5/2) + (Sd:Sd)^(5/4)),
  Sd = (G^2 + (G^2)^T)/2 - tr(G^2) I/3 (matrix products here, not the reference's expanded sums).
Special flows fix D further: WALE vanishes in pure shear (G^2 = 0) and Smagorinsky in solid rotation;
both vanish without a velocity gradient (WALE's 0/0 is defined as 0)."""
import importlib

import numpy as np
import pytest

from oracle_lib import Oracle

f3d = importlib.import_module("flame3d-mhd_amd")
NG = 4


def _affine_case(p, A, G, u0, T0, tgrad, rng, rho_rand=True, dtype=np.float64):
    nx, ny, nz = p.tot
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X = np.einsum("ab,bkji->akji", A, np.array([ii, jj, kk], dtype=np.float64))
    Q = np.zeros(p.cell_shape(6))
    Q[0] = 1.0 + 0.5 * rng.random(Q[0].shape) if rho_rand else 1.2
    Q[1:4] = np.einsum("ab,bkji->akji", G, X) + np.asarray(u0)[:, None, None, None]
    Q[5] = T0 + np.einsum("a,akji->kji", np.asarray(tgrad), X)
    Q[4] = Q[0] * p.Rg * Q[5]
    detA = np.linalg.det(A)
    Minv = np.linalg.inv(A) * detA  # rows: xi, eta, zeta metric rows divided by J
    m = [np.ascontiguousarray(np.full(p.cell_shape(), Minv[a, b])).astype(dtype) for a in range(3) for b in range(3)]
    J = np.full(p.cell_shape(), 1.0 / detA)
    return np.ascontiguousarray(Q.astype(dtype)), m, J, X, Minv, detA


def _expected(p, Q, X, Minv, detA, G, tgrad, model):
    """Closed-form Fv (3, 4, cells) of every cell of the viscous domain."""
    T = Q[5].astype(np.float64)
    mu = p.C_s * T * np.sqrt(T) / (T + p.T_s)
    lam = p.Cp * mu / p.Pr
    S = 0.5 * (G + G.T)
    SS = float(np.sum(S * S))
    G2 = G @ G
    Sd = 0.5 * (G2 + G2.T) - np.trace(G2) / 3 * np.eye(3)
    SdSd = float(np.sum(Sd * Sd))
    delta = detA ** (1.0 / 3.0)
    if model == f3d.LES_SMAG:
        D, C = np.sqrt(2 * SS), 0.1
    elif model == f3d.LES_WALE:
        den = SS ** 2.5 + SdSd ** 1.25
        D, C = (SdSd ** 1.5 / den if den > 0 else 0.0), 0.325
    else:
        D, C = 0.0, 0.0
    mut = np.minimum(Q[0].astype(np.float64) * (C * delta) ** 2 * D, 2 * mu)
    mue, lame = mu + mut, lam + p.Cp * mut / 0.9
    tau0 = G + G.T - 2.0 / 3.0 * np.trace(G) * np.eye(3)
    u = Q[1:4].astype(np.float64)
    out = np.zeros((3, 4) + T.shape)
    for d in range(3):
        row = Minv[d]
        for a in range(3):
            out[d, a] = mue * float(row @ tau0[a])
        # energy: row . (u.tau + lam grad T), tau symmetric
        ut = np.einsum("akji,ab->bkji", u, tau0)  # (u.tau0)_b
        out[d, 3] = mue * np.einsum("b,bkji->kji", row, ut) + lame * float(row @ np.asarray(tgrad))
    return out, mut, mu


def _run(p, Q, m, J):
    Fv = [np.zeros(p.fv_shape(), dtype=Q.dtype) for _ in range(3)]
    Oracle(p).viscous(*Fv, Q, m, J)
    return np.array(Fv)


def _cmp(p, Fv, exp):
    """Compare the viscous-domain cells (absolute cells 2 .. N_tot-3, stored at index - 2)."""
    sl = (slice(2, -2),) * 3
    e = exp[(slice(None), slice(None)) + sl]
    scale = np.abs(e).reshape(3, 4, -1).max(axis=(0, 2))
    err = np.abs(Fv - e).reshape(3, 4, -1).max(axis=(0, 2))
    return max(err[n] / (scale[n] if scale[n] > 0 else 1.0) for n in range(4))


A_SHEAR = np.array([[2e-3, 0.3e-3, 0.0], [0.2e-3, 1.5e-3, 0.1e-3], [0.0, 0.25e-3, 1.2e-3]])


@pytest.mark.parametrize("model", [f3d.LES_NONE, f3d.LES_SMAG, f3d.LES_WALE])
@pytest.mark.parametrize("scale,capped", [(2.0, False), (1e6, True)])
def test_les_general_gradient(model, scale, capped):
    p = f3d.make_params(6, 5, 4, les=model)
    rng = np.random.default_rng(7 + model)
    G = scale * rng.standard_normal((3, 3))
    for A in (A_SHEAR, 2.0 * A_SHEAR):
        Q, m, J, X, Minv, detA = _affine_case(p, A, G, [3.0, -0.5, 0.2], 300.0, [400.0, -200.0, 100.0], rng)
        exp, mut, mu = _expected(p, Q, X, Minv, detA, G, [400.0, -200.0, 100.0], model)
        if model != f3d.LES_NONE:
            frac = float(np.mean(mut >= 2 * mu * (1 - 1e-12)))
            assert (frac == 1.0) if capped else (frac == 0.0 and mut.min() > 1e-3 * mu.min()), frac
        e = _cmp(p, _run(p, Q, m, J), exp)
        assert e < 1e-11, (model, scale, e)


@pytest.mark.parametrize("model", [f3d.LES_SMAG, f3d.LES_WALE])
def test_les_special_flows(model):
    p = f3d.make_params(6, 5, 4, les=model)
    lam_p = f3d.make_params(6, 5, 4)
    rng = np.random.default_rng(3)
    gam = 50.0
    cases = {
        "shear": np.array([[0.0, gam, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]),
        "rotation": np.array([[0.0, -gam, 0.0], [gam, 0.0, 0.0], [0.0, 0.0, 0.0]]),
        "uniform": np.zeros((3, 3)),
    }
    A = np.diag([1e-3, 1e-3, 1e-3])
    for name, G in cases.items():
        Q, m, J, X, Minv, detA = _affine_case(p, A, G, [10.0, 1.0, 0.0], 290.0, [0.0, 0.0, 0.0], rng)
        Fv = _run(p, Q, m, J)
        Fl = _run(lam_p, Q, m, J)
        assert np.all(np.isfinite(Fv)), name
        zero = (name == "uniform") or (name == "shear" and model == f3d.LES_WALE) or \
               (name == "rotation" and model == f3d.LES_SMAG)
        if zero:
            # exactly zero model viscosity up to the rounding of the differenced velocity
            assert np.abs(Fv - Fl).max() <= 1e-12 * max(np.abs(Fl).max(), 1e-300), (name, model)
        else:
            # one non-zero closed form per model: Smagorinsky in shear D = gam, WALE in rotation
            # D = sqrt(|Sd|) with |Sd| = gam^2 sqrt(6)/3
            D = gam if model == f3d.LES_SMAG else np.sqrt(gam ** 2 * np.sqrt(6) / 3)
            C = 0.1 if model == f3d.LES_SMAG else 0.325
            T = Q[5]
            mu = p.C_s * T * np.sqrt(T) / (T + p.T_s)
            mut = np.minimum(Q[0] * (C * 1e-3) ** 2 * D, 2 * mu)
            ratio = 1 + mut / mu
            sl = (slice(2, -2),) * 3
            for d in range(3):
                for n in range(3):
                    ref = Fl[d, n] * ratio[sl]
                    s = np.abs(ref).max()
                    if s > 0:
                        assert np.abs(Fv[d, n] - ref).max() < 1e-12 * s, (name, d, n)


def test_les_float_form_range():
    """Near-wall-sized gradients (1e8 1/s): S^5 = 1e40 leaves float's range; the float form's WALE
    viscosity stays finite and matches the FP64 oracle on the same (float-rounded) inputs."""
    p = f3d.make_params(6, 5, 4, les=f3d.LES_WALE)
    rng = np.random.default_rng(11)
    G = 1e8 * rng.standard_normal((3, 3))
    A = np.diag([1e-6, 2e-6, 1e-6])
    Q, m, J, X, Minv, detA = _affine_case(p, A, G, [50.0, 0.0, 0.0], 300.0, [0.0, 0.0, 0.0], rng, dtype=np.float32)
    F32 = _run(p, Q, m, J)
    F64 = _run(p, np.ascontiguousarray(Q.astype(np.float64)), [np.ascontiguousarray(a.astype(np.float64)) for a in m], J)
    assert np.all(np.isfinite(F32))
    s = np.abs(F64).reshape(3, 4, -1).max(axis=(0, 2))
    e = np.abs(F32 - F64).reshape(3, 4, -1).max(axis=(0, 2))
    # float's differences of the linear field lose about u/(G h) ~ 1e-4 of the gradient: a loose bar
    assert max(e[n] / s[n] for n in range(3)) < 1e-2, e / s
