"""The measurement switches of the library (environment variables read once per process: the
kernel-by-kernel chain, the separate viscous kernels, the update tile forms (rows of 64 or two rows of
32 lanes per wave) and march lengths, the prefetch, the march segments, the sensor planes per block, the
block order, one lane or a pair of lanes per listed face) against the oracle.

Each variant runs in a child process (tests/switch_child.py) on a ramp case with a planted pressure
jump and spikes (faces on WENO7, WENO5 and NND), large enough for several update tiles in x and y and
several march segments in y and z; two RK3 steps, FP64 and FP32-bulk.  Bar as in the time-step parity
test: the error against the oracle within 10x the oracle's own response to a last-bit perturbation of
the initial state, plus 1e-12 (FP64) / 1e-6 (FP32-bulk)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from casegen import ramp_case, relerr
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
NSTEPS = 2
VARIANTS = [
    ("default", {}),
    ("chain", {"F3D_RHS_ENGINE": "chain"}),
    ("visc_separate", {"F3D_VISC_FUSED": "0"}),
    ("form8", {"F3D_VISC_FORM": "8"}),
    ("form12", {"F3D_VISC_FORM": "12"}),
    ("form6_kc16", {"F3D_VISC_FORM": "6", "F3D_VISC_KC": "16"}),
    ("w32_form12", {"F3D_VISC_W": "32", "F3D_VISC_FORM": "12"}),
    ("w32_form8_kc8", {"F3D_VISC_W": "32", "F3D_VISC_FORM": "8", "F3D_VISC_KC": "8"}),
    ("w32_form20", {"F3D_VISC_W": "32", "F3D_VISC_FORM": "20"}),
    ("w32_form16", {"F3D_VISC_W": "32", "F3D_VISC_FORM": "16"}),
    ("no_prefetch", {"F3D_VISC_PF": "0"}),
    ("segments", {"F3D_MARCH_SEG_Y": "16", "F3D_MARCH_SEG_Z": "8"}),
    ("sensor_plane", {"F3D_SENSOR_KS": "0"}),
    ("plain_order", {"F3D_XCD_REMAP": "0"}),
    ("heavy_one_lane", {"F3D_HEAVY_PAIR": "0"}),
    ("heavy_lane_pairs", {"F3D_HEAVY_PAIR": "1"}),
]


@pytest.fixture(scope="module")
def setup(tmp_path_factory):
    d = tmp_path_factory.mktemp("switches")
    arrays, ref = {}, {}
    params = None
    for tag, dt_ in (("f64", np.float64), ("f32b", np.float32)):
        c = ramp_case(130, 40, 36, dtype=dt_, jump=1.0, spikes=[(20, 5, 4, 20.0), (70, 30, 20, 3.0)],
                      hybrid_phi3=2.5)
        p, o = c.params, Oracle(c.params)
        params = p
        U = c.U()
        dt = 0.2 * o.cfl_dt(c.Q, c.metrics, c.J, 1.0)
        arrays.update({f"{tag}_U": U, f"{tag}_Q": c.Q, f"{tag}_inflow": c.inflow, f"{tag}_dt": np.float64(dt)})
        arrays.update({f"{tag}_m{i}": a for i, a in enumerate(c.metrics)})
        arrays["J"] = c.J
        Uo, Qo = U.copy(), c.Q.copy()
        o.time_step(NSTEPS, Uo, np.zeros_like(U), Qo, c.metrics, c.J, dt, c.inflow)
        rng = np.random.default_rng(5)
        if dt_ == np.float64:
            Up, Qp = U * (1 + 1e-16 * (2 * rng.random(U.shape) - 1)), c.Q.copy()
        else:
            toward = np.where(rng.random(c.Q.shape) < 0.5, -np.inf, np.inf).astype(np.float32)
            Up, Qp = U.copy(), np.ascontiguousarray(np.nextafter(c.Q, toward))
        o.time_step(NSTEPS, Up, np.zeros_like(U), Qp, c.metrics, c.J, dt, c.inflow)
        inner = (slice(None),) + (slice(4, -4),) * 3
        ref[tag] = dict(U=Uo, Q=Qo, sens_u=relerr(Up[inner], Uo[inner]), sens_q=relerr(Qp, Qo),
                        floor=1e-12 if dt_ == np.float64 else 1e-6)
    arrays["params"] = np.frombuffer(bytes(params), dtype=np.uint8)
    arrays["nsteps"] = np.int64(NSTEPS)
    src = str(d / "in.npz")
    np.savez(src, **arrays)
    return d, src, ref


@pytest.mark.parametrize("name,env", VARIANTS, ids=[v[0] for v in VARIANTS])
def test_switch_variant(setup, name, env):
    d, src, ref = setup
    dst = str(d / f"out_{name}.npz")
    e = dict(os.environ)
    for k in [k for k in e if k.startswith("F3D_")]:
        del e[k]
    e.update(env)
    print(f"switch variant {name}: {env}", flush=True)
    r = subprocess.run([sys.executable, os.path.join(HERE, "switch_child.py"), src, dst], env=e,
                       capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, f"{name}: child exit {r.returncode}\n{r.stderr[-2000:]}"
    out = np.load(dst)
    inner = (slice(None),) + (slice(4, -4),) * 3
    for tag, rf in ref.items():
        eu = relerr(out[f"{tag}_U"][inner], rf["U"][inner])
        eq = relerr(out[f"{tag}_Q"], rf["Q"])
        print(f"{name} {tag}: U err {eu:.2e} (sens {rf['sens_u']:.2e}), Q err {eq:.2e} (sens {rf['sens_q']:.2e})")
        assert eu <= 10 * rf["sens_u"] + rf["floor"], f"{name} {tag} U: {eu}"
        assert eq <= 10 * rf["sens_q"] + rf["floor"], f"{name} {tag} Q: {eq}"
