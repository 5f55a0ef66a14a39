#!/usr/bin/env python3
"""Benchmark of the Flame3D finite-difference RHS + SSP-RK3 path on one MI355X.

A step is one full SSP-RK3 step (three stages: sensor, split/reconstruct in three directions,
viscous flux, divergence, RK blend, c2Prim, ghost fill) over the whole grid.  Workload: the
24-degree compression ramp, 512x192x64 cells, FP64, Mach 2.9, Re_delta 6e4, tanh wall clustering
(first cell 1e-3 delta), laminar boundary-layer inflow with seeded +-1% velocity noise, isothermal
wall, periodic span.  Inputs (state, metrics, Jacobian) are resident in HBM before timing.

Printed: one JSON line.  `value` is Mcell-updates/s of the FP64 form (interior cells x RK3 steps
per second / 1e6).  Also reported: the finite-volume branch (FVM.jl, the reference run's default
scheme) on the same grid and state in both numeric forms, the FP32-bulk form on the same grid and its density and
wall-pressure errors against the FP64 run from the same initial state, the % of HBM peak of the
whole path (algorithmic bytes per cell-update x rate / 8 TB/s), the dominant kernel's roofline
from hipEvent timing of its launches inside the timed region, and the CPU oracle timed on the host
cores on a bounded sample (3 RK3 steps of the same grid).
"""
import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic (compulsory) HBM bytes per interior cell-update of a fully fused RK3 step:
# stage 1 reads U + 9 metrics + J and writes U; stages 2-3 also read Un.
ALG_BYTES_PER_CELL_UPDATE = {"f64": 560.0, "f32b": 452.0}
# The same per RHS + stage (one third of a step): the unit of the dominant kernel's roofline.
ALG_BYTES_PER_CELL_STAGE = {k: v / 3.0 for k, v in ALG_BYTES_PER_CELL_UPDATE.items()}


def kernel_alg_bytes(name, p, es):
    """Compulsory bytes of one launch of a library kernel (es = bytes of the bulk type)."""
    nxt, nyt, nzt = p.tot
    ncell = nxt * nyt * nzt
    nint = p.Nxp * p.Nyp * p.Nzp
    nface = (p.Nxp + 1) * p.Nyp * p.Nzp  # per direction, about the same for y/z
    if name == "flux_split_sw":  # rho,u,v,w,T + 3 metrics in, Fp/Fm out, every cell
        return ncell * (8 * es + 10 * es)
    if name == "advect":  # Fp,Fm (10), u,v,w,T (4), 3 metrics, phi in; 5 fluxes out, per face
        return nface * (10 * es + 4 * es + 3 * es + es + 5 * es)
    if name == "viscous_flux":  # u,v,w,T + 9 metrics in (+J double), 12 fluxes out
        return (nxt - 4) * (nyt - 4) * (nzt - 4) * (13 * es + 8 + 12 * es)
    if name == "div":  # 15 face fluxes + 12 viscous fluxes in, J in, U read+write
        return nint * (27 * es + 8 + 80)
    if name == "lin_comb":
        return nint * 120
    if name == "c2prim":
        return nint * (40 + 6 * es)
    if name == "shock_sensor":
        return ncell * 2 * es
    if name in ("fill_ghost", "exchange_ghost"):
        return (ncell - nint) * 2 * 6 * es
    if name.startswith("fused_"):
        return FUSED_BYTES.get(name, lambda p, es: 0)(p, es)
    return 0


def _nint(p):
    return p.Nxp * p.Nyp * p.Nzp


# Compulsory bytes per launch of the fused kernels (es = bulk element size; U, Un, J are double).
FUSED_BYTES = {
    # p in, phi out on every cell
    "fused_sensor": lambda p, es: p.tot[0] * p.tot[1] * p.tot[2] * 2 * es,
    # rho,u,v,w,T + 3 metrics + phi in, 5 RHS values out, per interior cell
    "fused_sweep_x": lambda p, es: _nint(p) * (9 * es + 5 * es),
    # x and y faces in one march: rho,u,v,w,T + 6 metrics + phi in, 5 RHS values out, per interior cell
    "fused_sweep_xy": lambda p, es: _nint(p) * (12 * es + 5 * es),
    # as x, but the RHS is read and written
    "fused_sweep_y": lambda p, es: _nint(p) * (9 * es + 10 * es),
    "fused_sweep_z": lambda p, es: _nint(p) * (9 * es + 10 * es),
    # viscous divergence fused into the update march: RHS 5 + u,v,w,T 4 + 9 metrics + J in, U and Un in,
    # U and Q out (stage-2/3 form), per interior cell
    "fused_update": lambda p, es: _nint(p) * (5 * es + 4 * es + 9 * es + 8 + 40 + 40 + 40 + 6 * es),
    "fused_heavy": lambda p, es: 0,
    # u,v,w,T + 9 metrics + J in, 12 viscous fluxes out, per cell of the interior+-2 box
    "fused_viscous": lambda p, es: (p.Nxp + 4) * (p.Nyp + 4) * (p.Nzp + 4) * (13 * es + 8 + 12 * es),
    # finite-volume sweeps: U (double) + face geometry (4) + phi in, RHS out (x) / in and out (y, z)
    "fvm_sweep_x": lambda p, es: _nint(p) * (40 + 4 * es + es + 5 * es),
    "fvm_sweep_y": lambda p, es: _nint(p) * (40 + 4 * es + es + 10 * es),
    "fvm_sweep_z": lambda p, es: _nint(p) * (40 + 4 * es + es + 10 * es),
    "fvm_heavy": lambda p, es: 0,
}
# Finite-volume step: per stage U + face geometry (3 directions x 4) + 9 metrics + J read, U written;
# stages 2-3 read Un.
ALG_BYTES_FVM = {"f64": 3 * (40 + 96 + 72 + 8 + 40) + 80.0, "f32b": 3 * (40 + 48 + 36 + 8 + 40) + 80.0}


# HBM bytes per launch actually moved, measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on the
# default workload (ramp 512x192x64, FP64; profiles/ramp512x192x64_fused_stage_fp64_fp32b.md), each scaled
# by what the counter reads for 1 GiB moved by a plain kernel (--read-probe: an FP64 sum, and --copy-probe:
# a device copy, both FETCH_SIZE x 2.000; WRITE_SIZE x 1.000); average over the three stage variants (the
# stage-1 variant's dispatches mix the calls' last steps, which write Un, with the others, which do not).
MEASURED_TRAFFIC = {("fused_update", "512x192x64", "f64"): (1874.4e6 + 690.5e6 + 2166.9e6 + 566.7e6 + 2170.9e6 + 568.6e6) / 3}


def build_case(f3d, cases, torch, nx, ny, nz, seed=2, sweep=0.0):
    import numpy as np
    c = cases.ramp(nx, ny, nz, seed=seed, sweep=sweep)
    p = f3d.make_params(nx, ny, nz, **c["params"])
    ctx = f3d.Flame3D(p)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    m = [torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0") for _ in range(9)]
    J = torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0")
    ctx.metrics(d(c["x"]), d(c["y"]), d(c["z"]), m, J)
    Q = d(c["Q"])
    inflow = d(c["inflow"])
    U = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    ctx.exchange_ghost(Q, 6)
    ctx.fillGhost(Q, U, inflow)
    ctx.prim2c(U, Q)
    dt_t = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    ctx.cfl_dt(dt_t, Q, m, J, 0.5)
    torch.cuda.synchronize()
    return dict(c=c, p=p, ctx=ctx, m=m, J=J, Q=Q, U=U, inflow=inflow, dt=float(dt_t.item()))


def run_steps(torch, ctx, state, steps, warmup, lib=None, table_steps=3, geo=None, capture_after=None):
    """Warm up; run `table_steps` steps with every library launch bracketed by hipEvents (per-kernel
    table); then time `steps` RK3 steps between two synchronisations with events around the
    launches of the dominant kernel only (its live average launch time).  `geo` (face geometry
    arrays): the finite-volume branch.  `capture_after` = n: the timed steps run as two timed parts,
    n steps and steps - n, with a copy of U taken between them outside the timed parts (returned as
    a fifth value).  Returns (ms/step, host wall ms/step, kernel table,
    (dominant name, its total ms, its launches)[, captured U])."""
    U, Un, Q, m, J, inflow, dt = state
    # n steps in one call, as the reference's time loop runs them (solver.jl:410-444)
    if geo is None:
        steps_call = lambda n: ctx.time_step(n, U, Un, Q, m, J, dt, inflow)
    else:
        steps_call = lambda n: ctx.time_step_fvm(n, U, Un, Q, m, J, geo, dt, inflow)
    if warmup:
        steps_call(warmup)
    torch.cuda.synchronize()
    table, dom = {}, None
    if lib is not None and table_steps:
        lib.f3d_timing_only(None)
        lib.f3d_timing_reset()
        lib.f3d_timing_enable(1)
        steps_call(table_steps)
        torch.cuda.synchronize()
        lib.f3d_timing_enable(0)
        table = kernel_table(lib)
        if table:
            dom = max(table.items(), key=lambda kv: kv[1][0])[0]
            lib.f3d_timing_only(dom.encode())
            lib.f3d_timing_reset()
            lib.f3d_timing_enable(1)
    parts = [steps] if capture_after is None else [capture_after, steps - capture_after]
    ms_tot, wall, captured = 0.0, 0.0, None
    for q, n in enumerate(parts):
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record()
        if n:
            steps_call(n)
        ev1.record()
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        ms_tot += ev0.elapsed_time(ev1)
        if q == 0 and capture_after is not None:
            if lib is not None and dom is not None:
                lib.f3d_timing_enable(0)
            captured = U.clone()
            torch.cuda.synchronize()
            if lib is not None and dom is not None:
                lib.f3d_timing_enable(1)
    live = None
    if dom is not None:
        lib.f3d_timing_enable(0)
        live = (dom,) + kernel_table(lib)[dom]
        lib.f3d_timing_only(None)
    if capture_after is not None:
        return ms_tot / steps, wall / steps * 1e3, table, live, captured
    return ms_tot / steps, wall / steps * 1e3, table, live


def progress(msg):
    """One line per finished leg on stderr (long runs stay visibly alive)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_table(lib):
    import ctypes
    buf = ctypes.create_string_buffer(4096)
    lib.f3d_timing_names(buf, 4096)
    out = {}
    for name in filter(None, buf.value.decode().split(",")):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        lib.f3d_timing_query(name.encode(), ctypes.byref(ms), ctypes.byref(n))
        out[name] = (ms.value, n.value)
    return out


def fvm_bench(f3d, torch, lib, st, U0, Q0, steps, warmup, f32b=True):
    """The finite-volume branch (characteristic reconstruction + HLLC, run.jl's default
    finite_volume = eigen_reconstruction = true) on the same grid, initial state and dt: FP64 and
    FP32-bulk rates, the FP32-bulk density error against the FP64 run, and per-kernel tables."""
    c, dt = st["c"], st["dt"]
    nx, ny, nz = st["p"].Nxp, st["p"].Nyp, st["p"].Nzp
    p = f3d.make_params(nx, ny, nz, **c["params"], fvm=f3d.FVM_EIGEN)
    ctx = f3d.Flame3D(p)
    X = [torch.from_numpy(c[a]).to("cuda:0") for a in ("x", "y", "z")]
    geo = lambda dt_: [torch.zeros((4,) + p.face_shape(d)[1:], dtype=dt_, device="cuda:0") for d in range(3)]
    G64, G32 = geo(torch.float64), geo(torch.float32)
    ctx.fvm_geometry(*X, G64)
    ctx.fvm_geometry(*X, G32)
    ncells = nx * ny * nz
    U, Un, Q = U0.clone(), torch.zeros_like(U0), Q0.clone()
    ms64, _, ktab, _ = run_steps(torch, ctx, (U, Un, Q, st["m"], st["J"], st["inflow"], dt), steps, warmup, lib=lib,
                                 geo=G64)
    ctx.nan_check(U)
    v64 = ncells / (ms64 * 1e-3) / 1e6
    progress(f"FVM FP64 {steps} steps: {ms64:.3f} ms/step")
    tab = lambda kt, ms: {k: dict(avg_us=round(v[0] / v[1] * 1e3, 2), launches=v[1], share=round(v[0] / (ms * 3), 4))
                          for k, v in sorted(kt.items())}
    out = dict(scheme="finite volume, characteristic (Roe-frame) WENO7/WENO5/minmod + HLLC (FVM.jl)",
               value=round(v64, 3), unit="Mcell-updates/s", ms_per_step=round(ms64, 4),
               hbm_frac_path=round(ALG_BYTES_FVM["f64"] * v64 * 1e6 / (HBM_PEAK_GBS * 1e9), 4),
               kernels=tab(ktab, ms64))
    if f32b:
        m32 = [a.float().contiguous() for a in st["m"]]
        U32, Un32, Q32 = U0.clone(), torch.zeros_like(U0), Q0.float().contiguous()
        ms32, _, ktab32, _ = run_steps(torch, ctx, (U32, Un32, Q32, m32, st["J"], st["inflow"].float().contiguous(),
                                                    dt), steps, warmup, lib=lib, geo=G32)
        ctx.nan_check(U32)
        v32 = ncells / (ms32 * 1e-3) / 1e6
        progress(f"FVM FP32-bulk {steps} steps: {ms32:.3f} ms/step")
        NG = 4
        inner = (slice(NG, -NG),) * 3
        r64, r32 = U[0][inner], U32[0][inner]
        out["fp32_bulk"] = dict(value=round(v32, 3), ms_per_step=round(ms32, 4), speedup_vs_f64=round(v32 / v64, 3),
                                hbm_frac_path=round(ALG_BYTES_FVM["f32b"] * v32 * 1e6 / (HBM_PEAK_GBS * 1e9), 4),
                                rho_L2_rel=float(torch.linalg.norm(r32 - r64) / torch.linalg.norm(r64)),
                                rho_Linf_rel=float((r32 - r64).abs().max() / r64.abs().max()),
                                steps_compared=steps + warmup + 3, kernels=tab(ktab32, ms32))
    return out


ALTERNATIVES = [
    ("les_smagorinsky", dict(les=1)), ("les_wale", dict(les=2)),
    ("split_lax_friedrichs", dict(split=1)), ("split_van_leer", dict(split=2)), ("split_ausm", dict(split=3)),
    ("component_wise", dict(character=0)), ("local_time_stepping", dict(lts=1)),
    # the filtering block after every step (the reference filters every filtering_interval = 100 steps)
    ("filter_nonlinear_every_step", dict(filtering=1, filtering_nonlinear=1, filtering_interval=1)),
    ("filter_linear_every_step", dict(filtering=1, filtering_nonlinear=0, filtering_interval=1, filtering_s0=0.05)),
]


def alternatives_bench(f3d, torch, lib, st, U0, Q0, steps, warmup):
    """The flag-selected alternatives of the finite-difference step (LES models, other splittings,
    component-wise reconstruction, local time stepping, the filters) on the same grid, initial state and
    dt, FP64:
    ms/step, the
    Mcell-updates/s rate and the cost relative to the default step measured in the same way."""
    base = st["p"]
    out = {}
    for name, kw in [("default", {})] + ALTERNATIVES:
        p = type(base).from_buffer_copy(base)
        for k, v in kw.items():
            setattr(p, k, v)
        ctx = f3d.Flame3D(p)
        U, Un, Q = U0.clone(), torch.zeros_like(U0), Q0.clone()
        ms, _, _, _ = run_steps(torch, ctx, (U, Un, Q, st["m"], st["J"], st["inflow"], st["dt"]), steps, warmup,
                                lib=None)
        try:
            ctx.nan_check(U)
            ok = True
        except f3d.F3DNaNError:
            ok = False
        ncells = p.Nxp * p.Nyp * p.Nzp
        out[name] = dict(ms_per_step=round(ms, 4), value=round(ncells / (ms * 1e-3) / 1e6, 3), finite=ok)
        del ctx
        torch.cuda.empty_cache()
    d = out["default"]["ms_per_step"]
    for v in out.values():
        v["cost_vs_default"] = round(v["ms_per_step"] / d, 3)
    progress("alternatives: " + ", ".join(f"{k} {v['ms_per_step']:.3f}" for k, v in out.items()))
    return out


def shocked_bench(f3d, torch, lib, st, Q0, steps, warmup, amp=0.05):
    """The step on a state with planted pressure noise (relative amplitude `amp`, seeded), so that a few
    per cent of the faces leave the UP7 branch and take the WENO7/WENO5/NND path: the heavy-face lists and
    the characteristic reconstruction of those faces are timed.  Same grid and metrics as the main leg, dt
    scaled by 0.01 so that the noise survives the timed steps; starts from the main leg's initial state.  Reports the share of faces off UP7 in the
    initial state, ms/step FP64 and FP32-bulk and the per-kernel table."""
    p, ctx = st["p"], st["ctx"]
    Q = Q0.clone()
    g = torch.Generator(device="cuda:0").manual_seed(7)
    noise = torch.rand(Q[4].shape, device="cuda:0", generator=g, dtype=torch.float64)
    Q[4] *= 1.0 + amp * (noise - 0.5)
    Q[5] = Q[4] / (Q[0] * p.Rg)
    U = torch.zeros_like(st["U"])
    ctx.fillGhost(Q, U, st["inflow"])
    ctx.prim2c(U, Q)
    phi = torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0")
    ctx.shockSensor(phi, Q)
    masks = [torch.zeros(p.mask_shape(d), dtype=torch.uint8, device="cuda:0") for d in range(3)]
    ctx.schemeMask(*masks, phi)
    torch.cuda.synchronize()
    nf = sum(m.numel() for m in masks)
    nh = sum(int((m != 0).sum()) for m in masks)
    hist = {name: int(sum(int((m == k).sum()) for m in masks)) for k, name in enumerate(SCHEMES)}
    del phi, masks
    out = dict(noise_amplitude=amp, faces=nf, faces_off_up7=nh, frac_off_up7=round(nh / nf, 4),
               mask_histogram_initial=hist)
    ncells = p.Nxp * p.Nyp * p.Nzp
    for name in ("f64", "f32b"):
        if name == "f64":
            state = (U.clone(), torch.zeros_like(U), Q.clone(), st["m"], st["J"], st["inflow"], st["dt"] * 0.01)
        else:
            state = (U.clone(), torch.zeros_like(U), Q.float().contiguous(), [a.float().contiguous() for a in st["m"]],
                     st["J"], st["inflow"].float().contiguous(), st["dt"] * 0.01)
        ms, _, tab, _ = run_steps(torch, ctx, state, steps, warmup, lib=lib, table_steps=2)
        ctx.nan_check(state[0])
        out[name] = dict(ms_per_step=round(ms, 4), value=round(ncells / (ms * 1e-3) / 1e6, 3),
                         kernels={k: dict(avg_us=round(v[0] / v[1] * 1e3, 2), launches=v[1])
                                  for k, v in sorted(tab.items())})
        del state
    torch.cuda.empty_cache()
    progress(f"shocked state ({100 * nh / nf:.1f}% faces off UP7): FP64 {out['f64']['ms_per_step']:.3f}, "
             f"FP32-bulk {out['f32b']['ms_per_step']:.3f} ms/step")
    return out


def io_bench(st, Q0):
    """The reference's on-disk formats on the bench grid (host side, libf3d_h5.so): a chk-*.h5 checkpoint
    of Q as IO.jl:152-179 writes it (Float32, shuffle + deflate level 1, the reference's default) and
    uncompressed, and metrics.h5 (Float32, deflate level 3 as parse_mesh.jl), written and read back."""
    import tempfile
    import numpy as np
    h5 = importlib.import_module("flame3d-mhd_amd.h5io")
    p = st["p"]
    Q = Q0.cpu().numpy()
    m = [a.cpu().numpy() for a in st["m"]]
    J = st["J"].cpu().numpy()
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, compress in (("checkpoint_deflate1", 1), ("checkpoint_raw", 0)):
            path = os.path.join(d, f"chk-{compress}.h5")
            t0 = time.perf_counter()
            h5.write_checkpoint(path, p, Q, compress=compress)
            t1 = time.perf_counter()
            Qr = h5.read_checkpoint(path, p, dtype=np.float64)
            t2 = time.perf_counter()
            mb = Q.size * 4 / 1e6  # Float32 payload
            out[name] = dict(payload_MB=round(mb, 1), file_MB=round(os.path.getsize(path) / 1e6, 1),
                             write_MB_s=round(mb / (t1 - t0), 1), read_MB_s=round(mb / (t2 - t1), 1),
                             exact_float32_round_trip=bool(np.array_equal(Qr, Q.astype(np.float32).astype(np.float64))))
        path = os.path.join(d, "metrics.h5")
        t0 = time.perf_counter()
        h5.write_metrics(path, p, m, J, compress=3)
        t1 = time.perf_counter()
        h5.read_metrics(path, p)
        t2 = time.perf_counter()
        mb = 10 * J.size * 4 / 1e6
        out["metrics_deflate3"] = dict(payload_MB=round(mb, 1), file_MB=round(os.path.getsize(path) / 1e6, 1),
                                       write_MB_s=round(mb / (t1 - t0), 1), read_MB_s=round(mb / (t2 - t1), 1))
    progress("io: " + ", ".join(f"{k} w {v['write_MB_s']} r {v['read_MB_s']} MB/s" for k, v in out.items()))
    return out


ORACLE_FLAGS = "-O3 -std=c++17 -ffp-contract=off -fno-fast-math -fopenmp -fPIC -shared"


def start_native_oracle_build():
    """Start compiling the CPU oracle for this host (-march=native) in the background; the CPU baseline
    waits for it (falls back to the prebuilt -march=x86-64-v3 oracle/liboracle_f3d.so if it fails)."""
    import subprocess
    import tempfile
    d = tempfile.mkdtemp(prefix="f3d_oracle_native_")
    so = os.path.join(d, "liboracle_f3d.so")
    cmd = ["g++", *ORACLE_FLAGS.split(), "-march=native", "-o", so, os.path.join(ROOT, "oracle", "f3d_oracle.cpp")]
    try:
        proc = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    except OSError:
        return None
    return dict(proc=proc, so=so)


def cpu_baseline(f3d, cases, nx, ny, nz, steps=20, native=None):
    """CPU oracle (C++ FP64 restatement, OpenMP) on the same grid for `steps` RK3 steps, built for this
    host's cores (-march=native) when that build succeeded."""
    import numpy as np
    build = "-march=x86-64-v3 (prebuilt oracle/liboracle_f3d.so)"
    if native is not None:
        rc = native["proc"].wait()
        if rc == 0:
            os.environ["F3D_ORACLE_SO"] = native["so"]
            build = "-march=native (compiled on this host)"
        else:
            build += f"; the native build failed ({rc})"
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, max_threads
    c = cases.ramp(nx, ny, nz)
    p = f3d.make_params(nx, ny, nz, **c["params"])
    o = Oracle(p)
    m = [np.zeros(p.cell_shape()) for _ in range(9)]
    J = np.zeros(p.cell_shape())
    o.metrics(c["x"], c["y"], c["z"], m, J)
    Q = c["Q"].copy()
    U = np.zeros(p.cell_shape(5))
    o.exchange(Q, 6)
    o.fill(Q, U, c["inflow"])
    o.prim2c(U, Q)
    Un = np.zeros_like(U)
    dt = o.cfl_dt(Q, m, J, 0.5)
    t0 = time.perf_counter()
    o.time_step(steps, U, Un, Q, m, J, dt, c["inflow"])
    el = time.perf_counter() - t0
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), cpu)
    except OSError:
        pass
    return dict(value=round(nx * ny * nz * steps / el / 1e6, 4), unit="Mcell-updates/s", cores=max_threads(),
                kind="port", sample=f"oracle FP64, {steps} RK3 steps of the {nx}x{ny}x{nz} ramp ({el:.1f} s)",
                omp_threads=max_threads(), nproc=os.cpu_count(), affinity_cpus=len(os.sched_getaffinity(0)),
                cpu_model=cpu, build=build, flags=ORACLE_FLAGS)


def adjoint_bench(f3d, cases, torch, lib, nx, ny, nz, nsteps, fd=True, every=1, sweep=0.0, case=None, U_start=None,
                  dt=None, fd_params=(0, 1, 2)):
    """Inflow-parameter adjoint of the ramp: theta = (M_inf, T_wall, 16 Fourier amplitudes),
    objective = sum over `nsteps` FP64 RK3 steps of dt * sum_wall p * |grad eta|/J (wall pressure
    x wall-face area on the first cell row), checkpoint every `every` RK steps.  Times one forward
    step and the adjoint (forward sweep + reverse with recomputation), chains the boundary-data adjoints
    to d objective / d theta with complex-step derivatives of the boundary data, and checks the
    components in `fd_params` against GPU central finite differences (the amplitudes at two steps).
    `case` (a cases.ramp dict), `U_start` (conserved state to start from) and `dt` reuse a forward leg's."""
    import numpy as np
    NG = 4
    c = case if case is not None else cases.ramp(nx, ny, nz, sweep=sweep)
    st = c["state"]
    ygh = c["y"][:, :, 0:NG].mean(axis=2)
    th = cases.theta0()
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")

    def setup(theta):
        inflow, qinf, Tw = cases.ramp_theta(ygh, st, theta)
        kw = dict(c["params"])
        kw.update(q_inf=list(qinf), T_wall=float(Tw))
        return f3d.make_params(nx, ny, nz, **kw), d(inflow)

    p, inflow = setup(th)
    ctx = f3d.Flame3D(p)
    m = [torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0") for _ in range(9)]
    J = torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0")
    ctx.metrics(d(c["x"]), d(c["y"]), d(c["z"]), m, J)
    if U_start is not None:
        U0 = U_start
        Q0 = torch.zeros(p.cell_shape(6), dtype=torch.float64, device="cuda:0")
        ctx.c2Prim(U0, Q0)
        ctx.exchange_ghost(Q0, 6)
        ctx.fillGhost(Q0, None, inflow)
    else:
        Q0 = d(c["Q"])
        U0 = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
        ctx.exchange_ghost(Q0, 6)
        ctx.fillGhost(Q0, U0, inflow)
        ctx.prim2c(U0, Q0)
    if dt is None:
        dt_t = torch.zeros(1, dtype=torch.float64, device="cuda:0")
        ctx.cfl_dt(dt_t, Q0, m, J, 0.5)
        dt = float(dt_t.item())
    hist = mask_histogram(torch, ctx, p, Q0)
    W = torch.zeros(p.cell_shape(6), dtype=torch.float64, device="cuda:0")
    area = torch.sqrt(m[3] ** 2 + m[4] ** 2 + m[5] ** 2)
    W[4, NG:-NG, NG, NG:-NG] = area[NG:-NG, NG, NG:-NG]
    del area

    def objective(params, inflow_t):
        cx = f3d.Flame3D(params)
        U = U0.clone()
        Un = torch.zeros_like(U)
        Q = torch.zeros(params.cell_shape(6), dtype=torch.float64, device="cuda:0")
        cx.c2Prim(U, Q)
        cx.exchange_ghost(Q, 6)
        cx.fillGhost(Q, None, inflow_t)
        obj = torch.zeros((), dtype=torch.float64, device="cuda:0")
        for _ in range(nsteps):
            cx.time_step(1, U, Un, Q, m, J, dt, inflow_t)
            obj += dt * (W * Q).sum()
        val = float(obj.item())
        del cx, U, Un, Q
        torch.cuda.empty_cache()
        return val

    # forward step time on this grid
    U, Un, Q = U0.clone(), torch.zeros_like(U0), torch.zeros(p.cell_shape(6), dtype=torch.float64, device="cuda:0")
    ctx.c2Prim(U, Q)
    ctx.exchange_ghost(Q, 6)
    ctx.fillGhost(Q, None, inflow)
    ms_fwd, _, _, _ = run_steps(torch, ctx, (U, Un, Q, m, J, inflow, dt), min(nsteps, 50), 2)
    del U, Un, Q
    ctx._ws.clear()  # the adjoint workspace holds its own forward workspace
    torch.cuda.empty_cache()
    # adjoint (one short call first: workspace creation)
    obj = torch.zeros(nsteps, dtype=torch.float64, device="cuda:0")
    lamU0 = torch.zeros_like(U0)
    lam_in = torch.zeros_like(inflow)
    lam_bc = torch.zeros(7, dtype=torch.float64, device="cuda:0")
    ckpt = torch.empty(ctx.adjoint_ckpt_bytes(nsteps, every) // 8, dtype=torch.float64, device="cuda:0")
    ctx.adjoint(1, 1, U0, m, J, dt, W, obj, lamU0, inflow, lam_in, lam_bc, ckpt=ckpt)
    lam_in.zero_()
    lam_bc.zero_()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    ctx.adjoint(nsteps, every, U0, m, J, dt, W, obj, lamU0, inflow, lam_in, lam_bc, ckpt=ckpt)
    ev1.record()
    torch.cuda.synchronize()
    ms_adj = ev0.elapsed_time(ev1)
    # per-kernel time of the reverse sweep (separate run, hipEvent-bracketed launches); the same
    # gradient again (deterministic kernels give the same bits)
    lib.f3d_timing_reset()
    lib.f3d_timing_enable(1)
    obj2 = torch.zeros_like(obj)
    lamU0b, lam_in2, lam_bc2 = torch.zeros_like(U0), torch.zeros_like(inflow), torch.zeros_like(lam_bc)
    ctx.adjoint(nsteps, every, U0, m, J, dt, W, obj2, lamU0b, inflow, lam_in2, lam_bc2, ckpt=ckpt)
    torch.cuda.synchronize()
    lib.f3d_timing_enable(0)
    ktab = {k: v for k, v in kernel_table(lib).items() if k.startswith("adj_")}
    repeat = dict(objective_bits_equal=bool(torch.equal(obj, obj2)), lamU0_bits_equal=bool(torch.equal(lamU0, lamU0b)),
                  lam_inflow_bits_equal=bool(torch.equal(lam_in, lam_in2)),
                  lam_bc_bits_equal=bool(torch.equal(lam_bc, lam_bc2)))
    ckpt_bytes = int(ckpt.numel() * 8)
    del ckpt, lamU0b, lam_in2, lam_bc2
    ctx._aws = None
    torch.cuda.empty_cache()
    # chain to theta
    jac = cases.ramp_theta_jacobian(ygh, st, th)
    li, lb = lam_in.cpu().numpy(), lam_bc.cpu().numpy()
    grad = np.array([float(np.sum(li * dinf) + np.sum(lb[:6] * dq) + lb[6] * dTw) for dinf, dq, dTw in jac])
    J0 = float(obj.sum().item())
    fd_out = []
    if fd:
        j_a, j_b = objective(*setup(th)), objective(*setup(th))  # repeatability of the objective itself
        names = ["M_inf", "T_wall"] + [f"a_{q}" for q in range(1, len(th) - 1)]
        for k in fd_params:
            steps_h = [1e-4 * abs(th[k])] if k < 2 else [1e-2, 4e-2]  # the amplitudes enter almost linearly
            for h in steps_h:
                tp, tm = th.copy(), th.copy()
                tp[k] += h
                tm[k] -= h
                jp = objective(*setup(tp))
                jm = objective(*setup(tm))
                fdv = (jp - jm) / (2 * h)
                fd_out.append(dict(param=names[k], adjoint=grad[k], central_fd=fdv, step=h,
                                   rel_diff=abs(fdv - grad[k]) / max(abs(fdv), 1e-300),
                                   objective_change_rel=abs(jp - jm) / abs(J0),
                                   fd_noise_rel=abs(j_a - j_b) / max(abs(jp - jm), 1e-300)))
        progress("adjoint FD: " + ", ".join(f"{f['param']} (h {f['step']:.1e}) {f['rel_diff']:.1e}" for f in fd_out))
    rev = (ms_adj - nsteps * ms_fwd) / nsteps
    return dict(
        config=f"ramp {nx}x{ny}x{nz}{f' swept {sweep:g} deg' if sweep else ''} FP64, {nsteps} RK3 steps, checkpoint "
               f"every {every} step(s) (states inside a segment recomputed); theta = (M_inf, T_wall, 16 inflow "
               f"Fourier amplitudes); objective = sum_n dt * wall pressure x wall-face area",
        mask_histogram_initial=hist,
        checkpoint_bytes=ckpt_bytes,
        ms_forward_step=round(ms_fwd, 4), ms_adjoint_total=round(ms_adj, 3),
        ms_reverse_step=round(rev, 4), adjoint_over_forward_step=round(rev / ms_fwd, 3),
        adjoint_total_over_forward_step=round(ms_adj / nsteps / ms_fwd, 3),
        ratio_note="adjoint_over_forward_step: reverse sweep (with recomputation) per step / forward step; "
                   "adjoint_total_over_forward_step: the whole call (forward sweep + reverse) per step / forward step",
        objective=J0, objective_check_rel=abs(J0 - float(obj2.sum().item())) / abs(J0), repeat=repeat,
        grad_theta=[float(g) for g in grad], fd_check=fd_out,
        kernels={k: dict(avg_us=round(v[0] / v[1] * 1e3, 2), launches=v[1], total_ms=round(v[0], 3))
                 for k, v in sorted(ktab.items())},
    )


SCHEMES = ("UP7", "WENO7", "WENO5", "NND2")


def mask_histogram(torch, ctx, p, Q):
    """Faces per scheme of the face mask (schemes.jl:724-731, 813, 931) at the state Q, the three
    directions together, and the share of faces off UP7."""
    phi = torch.zeros(p.cell_shape(), dtype=Q.dtype, device="cuda:0")
    ctx.shockSensor(phi, Q)
    masks = [torch.zeros(p.mask_shape(d), dtype=torch.uint8, device="cuda:0") for d in range(3)]
    ctx.schemeMask(*masks, phi)
    cnt = [int(sum(int((m == k).sum()) for m in masks)) for k in range(4)]
    del phi, masks
    out = dict(zip(SCHEMES, cnt))
    out["frac_off_up7"] = round(1.0 - cnt[0] / max(sum(cnt), 1), 6)
    return out


def forward_pair(f3d, cases, torch, lib, nx, ny, nz, steps, warmup, sweep=0.0, tables=True, capture_after=None,
                 f32b=True):
    """The ramp case on an nx x ny x nz grid: FP64 and FP32-bulk RK3 steps from the same initial state and
    dt (warmup + 3 table steps + `steps` timed), with the mask histogram of the initial state and the
    density and wall-pressure errors of FP32-bulk against FP64 after the same steps.  Returns
    (result dict, case dict st, U0 and Q0 of the initial state, extras)."""
    NG = 4
    st = build_case(f3d, cases, torch, nx, ny, nz, sweep=sweep)
    p, ctx, dt = st["p"], st["ctx"], st["dt"]
    ncells = nx * ny * nz
    U0, Q0 = st["U"].clone(), st["Q"].clone()
    hist = mask_histogram(torch, ctx, p, Q0)
    tab = lambda kt, ms: {k: dict(avg_us=round(v[0] / v[1] * 1e3, 2), launches=v[1], share=round(v[0] / (ms * 3), 4))
                          for k, v in sorted(kt.items())}
    U, Un, Q = st["U"], torch.zeros_like(st["U"]), st["Q"]
    ms64, wall64, ktab, live = run_steps(torch, ctx, (U, Un, Q, st["m"], st["J"], st["inflow"], dt), steps, warmup,
                                         lib=lib if tables else None)
    ctx.nan_check(U)
    v64 = ncells / (ms64 * 1e-3) / 1e6
    progress(f"{nx}x{ny}x{nz}{' swept' if sweep else ''} FP64 {steps} steps: {ms64:.3f} ms/step")
    out = dict(value=round(v64, 3), unit="Mcell-updates/s", ms_per_step=round(ms64, 4), steps=steps,
               hbm_frac_path=round(ALG_BYTES_PER_CELL_UPDATE["f64"] * v64 * 1e6 / (HBM_PEAK_GBS * 1e9), 4),
               mask_histogram_initial=hist)
    if tables:
        out["kernels"] = tab(ktab, ms64)
    if not f32b:
        return out, st, U0, Q0, dict(live=live, wall64=wall64, captured=None)
    m32 = [a.float().contiguous() for a in st["m"]]
    U32, Un32, Q32 = U0.clone(), torch.zeros_like(U0), Q0.float().contiguous()
    in32 = st["inflow"].float().contiguous()
    r = run_steps(torch, ctx, (U32, Un32, Q32, m32, st["J"], in32, dt), steps, warmup, lib=lib if tables else None,
                  capture_after=capture_after)
    ms32, ktab32 = r[0], r[2]
    ctx.nan_check(U32)
    v32 = ncells / (ms32 * 1e-3) / 1e6
    progress(f"{nx}x{ny}x{nz}{' swept' if sweep else ''} FP32-bulk {steps} steps: {ms32:.3f} ms/step")
    inner = (slice(NG, -NG),) * 3
    r64, r32 = U[0][inner].double(), U32[0][inner].double()
    pw64 = Q[4][NG:-NG, NG, NG:-NG].mean(dim=0)  # p at the first cell above the wall, spanwise mean
    pw32 = Q32[4][NG:-NG, NG, NG:-NG].double().mean(dim=0)
    f32 = dict(value=round(v32, 3), ms_per_step=round(ms32, 4), speedup_vs_f64=round(v32 / v64, 3),
               hbm_frac_path=round(ALG_BYTES_PER_CELL_UPDATE["f32b"] * v32 * 1e6 / (HBM_PEAK_GBS * 1e9), 4),
               rho_L2_rel=float(torch.linalg.norm(r32 - r64) / torch.linalg.norm(r64)),
               rho_Linf_rel=float((r32 - r64).abs().max() / r64.abs().max()),
               pwall_L2_rel=float(torch.linalg.norm(pw32 - pw64) / torch.linalg.norm(pw64)),
               pwall_Linf_rel=float((pw32 - pw64).abs().max() / pw64.abs().max()),
               steps_compared=steps + warmup + (3 if tables else 0))
    if tables:
        f32["kernels"] = tab(ktab32, ms32)
    out["fp32_bulk"] = f32
    del m32, Un32, Q32, U32
    extras = dict(live=live, wall64=wall64, captured=r[4] if capture_after is not None else None)
    return out, st, U0, Q0, extras


def large_config_leg(f3d, cases, torch, lib, grid, steps, warmup, sweep=0.0):
    """A large configuration of the ramp (FP64 and FP32-bulk, `steps` timed steps each, errors of FP32-bulk
    against FP64), without per-kernel tables."""
    nx, ny, nz = grid
    out, st, U0, Q0, _ = forward_pair(f3d, cases, torch, lib, nx, ny, nz, steps, warmup, sweep=sweep, tables=False)
    out["workload"] = f"compression ramp 24 deg{f', swept {sweep:g} deg' if sweep else ''}, {nx}x{ny}x{nz}"
    del st, U0, Q0
    torch.cuda.empty_cache()
    return out


def swept_leg(f3d, cases, torch, lib, grid, steps, warmup, adj_steps, every, fd=True):
    """The swept-ramp one-GPU subdomain: FP64 and FP32-bulk forward over `steps` timed steps (after warmup
    and table steps), then the FP64 adjoint over the last `adj_steps` of the FP32-bulk trajectory
    (from its state `adj_steps` steps before the end), checkpoint every `every` steps with segment
    recomputation, with central finite-difference checks of d objective / d (M_inf, T_wall)."""
    nx, ny, nz = grid
    out, st, U0, Q0, ex = forward_pair(f3d, cases, torch, lib, nx, ny, nz, steps, warmup, sweep=30.0, tables=False,
                                       capture_after=steps - adj_steps)
    out["workload"] = f"compression ramp 24 deg, swept 30 deg, {nx}x{ny}x{nz}"
    Ustart = ex["captured"]
    dt = st["dt"]
    case = st["c"]
    for k in ("m", "J", "Q", "U", "inflow", "ctx"):
        st.pop(k, None)
    del U0, Q0
    torch.cuda.empty_cache()
    adj = adjoint_bench(f3d, cases, torch, lib, nx, ny, nz, adj_steps, fd=fd, every=every, sweep=30.0, case=case,
                        U_start=Ustart, dt=dt, fd_params=(0, 1))
    adj["start"] = (f"FP32-bulk trajectory after {steps - adj_steps} of its {steps} timed steps "
                    f"({warmup} warmup + 3 table steps before them): the adjoint covers the last {adj_steps}")
    out["adjoint"] = adj
    del Ustart
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", default="512x192x64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f32b", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-grid", default="256x128x64")
    ap.add_argument("--adjoint-steps", type=int, default=50)
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint leg (profiling)")
    ap.add_argument("--no-fd", action="store_true", help="skip the finite-difference check of the adjoint")
    ap.add_argument("--sweep", type=float, default=0.0, help="sweep angle (deg) of the ramp grid")
    ap.add_argument("--adjoint-every", type=int, default=1, help="checkpoint interval of the adjoint (RK3 steps)")
    ap.add_argument("--adjoint-sweep", type=float, default=0.0)
    ap.add_argument("--no-fvm", action="store_true", help="skip the finite-volume leg")
    ap.add_argument("--fvm-only", action="store_true", help="only the finite-volume leg (profiling)")
    ap.add_argument("--no-alternatives", action="store_true", help="skip the flag-selected alternatives leg")
    ap.add_argument("--no-io", action="store_true", help="skip the on-disk formats leg")
    ap.add_argument("--no-shocked", action="store_true",
                    help="skip the leg on a noisy state (faces off UP7 take the WENO/NND path)")
    ap.add_argument("--shocked-only", action="store_true", help="only the noisy-state leg (profiling)")
    ap.add_argument("--no-large", action="store_true", help="skip the 1024x384x256 ramp leg")
    ap.add_argument("--large-grid", default="1024x384x256")
    ap.add_argument("--large-steps", type=int, default=200)
    ap.add_argument("--no-swept", action="store_true", help="skip the swept 1024x256x256 leg")
    ap.add_argument("--swept-grid", default="1024x256x256")
    ap.add_argument("--swept-steps", type=int, default=500)
    ap.add_argument("--swept-adjoint-steps", type=int, default=100)
    ap.add_argument("--swept-adjoint-every", type=int, default=10)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy 5 times (calibrates the HBM write counter)")
    ap.add_argument("--read-probe", action="store_true",
                    help="only run a 1 GiB FP64 sum 5 times (calibrates the HBM read counter: coalesced 8-byte "
                         "loads of 64 lanes, as the marches read)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the ramp's whole subdomain on one MI355X); "
              "--gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)
    native = None if (args.no_cpu or args.copy_probe or args.read_probe or args.adjoint_only or args.fvm_only
                      or args.shocked_only) else start_native_oracle_build()
    import numpy as np
    import torch
    assert torch.cuda.is_available(), "bench needs a GPU"
    if args.copy_probe or args.read_probe:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")  # 1 GiB
        b = torch.empty_like(a)
        for _ in range(5):
            if args.copy_probe:
                b.copy_(a)
            else:
                b[:1] = a.sum()
        torch.cuda.synchronize()
        print(json.dumps({"probe": "copy" if args.copy_probe else "read", "bytes_read": a.numel() * 8, "runs": 5}))
        return
    f3d = importlib.import_module("flame3d-mhd_amd")
    cases = importlib.import_module("flame3d-mhd_amd.cases")
    lib = f3d.lib()
    nx, ny, nz = map(int, args.grid.split("x"))
    ncells = nx * ny * nz
    if args.adjoint_only:
        ax, ay, az = map(int, args.adjoint_grid.split("x"))
        print(json.dumps({"adjoint": adjoint_bench(f3d, cases, torch, lib, ax, ay, az, args.adjoint_steps,
                                                   fd=not args.no_fd, every=args.adjoint_every,
                                                   sweep=args.adjoint_sweep)}))
        return
    if args.fvm_only or args.shocked_only:
        st = build_case(f3d, cases, torch, nx, ny, nz, sweep=args.sweep)
        U0, Q0 = st["U"].clone(), st["Q"].clone()
        if args.fvm_only:
            print(json.dumps({"finite_volume": fvm_bench(f3d, torch, lib, st, U0, Q0, args.steps, args.warmup,
                                                         f32b=not args.no_f32b)}))
        else:
            print(json.dumps({"shocked_state": shocked_bench(f3d, torch, lib, st, Q0, max(3, args.steps // 4), 2)}))
        return
    # ---- headline: FP64 (and FP32-bulk) on the ramp grid
    main_leg, st, U0, Q0, ex = forward_pair(f3d, cases, torch, lib, nx, ny, nz, args.steps, args.warmup,
                                            sweep=args.sweep, f32b=not args.no_f32b)
    p, dt = st["p"], st["dt"]
    value64, ms64 = main_leg["value"], main_leg["ms_per_step"]
    live = ex["live"]
    # ---- roofline of the dominant kernel (hipEvent-timed launches inside the FP64 timed region)
    roof = None
    if live:
        name, tot_ms, n = live
        avg_s = tot_ms / n * 1e-3
        alg = ALG_BYTES_PER_CELL_STAGE["f64"] * ncells  # algorithmic bytes of one cell-stage x interior cells
        kbytes = kernel_alg_bytes(name, p, 8)
        achieved = alg / avg_s / 1e9
        traffic = MEASURED_TRAFFIC.get((name, args.grid, "f64"))
        roof = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel=name,
                    avg_launch_us=round(avg_s * 1e6, 2), alg_bytes_per_launch=alg,
                    alg_bytes_basis="186.7 B per interior cell per RHS + stage (FP64), x interior cells",
                    kernel_bytes_per_launch=kbytes,
                    achieved_kernel_bytes=round(kbytes / avg_s / 1e9, 2),
                    frac_kernel_bytes=round(kbytes / avg_s / 1e9 / HBM_PEAK_GBS, 4),
                    traffic_over_alg=round(traffic / alg, 3) if traffic else None,
                    share_of_step=round(tot_ms / (ms64 * args.steps), 4))
    fvm = None
    if not args.no_fvm:
        fvm = fvm_bench(f3d, torch, lib, st, U0, Q0, args.steps, args.warmup, f32b=not args.no_f32b)
        fvm["mask_histogram_initial"] = main_leg["mask_histogram_initial"]
    alt = None
    if not args.no_alternatives:
        alt = alternatives_bench(f3d, torch, lib, st, U0, Q0, max(3, args.steps // 4), 2)
    shocked = None
    if not args.no_shocked:
        shocked = shocked_bench(f3d, torch, lib, st, Q0, max(3, args.steps // 4), 2)
    io = None
    if not args.no_io:
        try:
            io = io_bench(st, Q0)
        except ImportError as e:  # libf3d_h5.so not built (no HDF5 headers where it was built)
            io = {"skipped": str(e)}
    del st, U0, Q0
    torch.cuda.empty_cache()
    adj = None
    if not args.no_adjoint:
        ax, ay, az = map(int, args.adjoint_grid.split("x"))
        adj = adjoint_bench(f3d, cases, torch, lib, ax, ay, az, args.adjoint_steps, fd=not args.no_fd,
                            every=args.adjoint_every, sweep=args.adjoint_sweep)
        progress(f"adjoint: {adj['ms_reverse_step']:.3f} ms per reverse step, "
                 f"{adj['adjoint_total_over_forward_step']:.3f}x a forward step with the forward sweep")
    large = None
    if not args.no_large:
        large = large_config_leg(f3d, cases, torch, lib, tuple(map(int, args.large_grid.split("x"))),
                                 args.large_steps, args.warmup)
    swept = None
    if not args.no_swept:
        swept = swept_leg(f3d, cases, torch, lib, tuple(map(int, args.swept_grid.split("x"))), args.swept_steps,
                          args.warmup, args.swept_adjoint_steps, args.swept_adjoint_every, fd=not args.no_fd)
        progress(f"swept adjoint: {swept['adjoint']['adjoint_total_over_forward_step']:.3f}x a forward step with "
                 f"the forward sweep")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(f3d, cases, nx, ny, nz, steps=args.cpu_steps, native=native)
    line = {
        "metric": "Mcell-updates/s (FP64 RHS + SSP-RK3 step)",
        "value": round(value64, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms64, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded laminar-BL ramp inflow, generated grid)",
        "config": {"workload": f"compression ramp 24 deg{f', swept {args.sweep:g} deg' if args.sweep else ''}, "
                               f"{nx}x{ny}x{nz}, FP64, M2.9, Re_delta 6e4, isothermal wall, periodic span",
                   "cells": ncells, "dt": dt},
        "hbm_frac_path": main_leg["hbm_frac_path"],
        "mask_histogram_initial": main_leg["mask_histogram_initial"],
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp32_bulk": main_leg.get("fp32_bulk"),
        "large_1024x384x256": large,
        "swept_1024x256x256": swept,
        "finite_volume": fvm,
        "alternatives": alt,
        "shocked_state": shocked,
        "io": io,
        "adjoint": adj,
        "kernels": main_leg.get("kernels"),
        "host_wall_ms_per_step": round(ex["wall64"], 4),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
