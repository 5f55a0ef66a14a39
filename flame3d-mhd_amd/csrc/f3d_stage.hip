// Workspace, RHS and RK3 stage entry points, global reductions (CFL dt, NaN flag) and the FP64
// metric computation.
#include <cmath>
#include <cstring>
#include <mutex>

#include "f3d_internal.hpp"
#include "f3d_fused.hpp"

using namespace f3d;

struct f3d_workspace {
  f3d_params p;
  int prec;
  size_t bytes;
  void* base;
  // bulk-type scratch (element size 8 for F3D_PREC_F64, 4 for F3D_PREC_F32B)
  void *phi, *Fp, *Fm, *Fx, *Fy, *Fz, *Fvx, *Fvy, *Fvz;
  void* R;       // interior RHS accumulator [5][Nzp][Nyp][Nxp] of the bulk type
  double* red;   // reduction partials
  int* flag;     // NaN flag partials
  int nred;      // number of partial slots
  unsigned long long* heavy;  // compacted non-UP7 face list (fused path)
  unsigned int* hcount;
  unsigned long long hcap, hncap;
  void *Qa, *Qb;  // stage outputs of the fused path (primitive state, 6 vars, bulk type)
  double* ldt;    // per-cell dt of local time stepping (cell-shaped; compute_dt at stage 1)
  double* fsc;    // sensor of the nonlinear filter (cell-shaped)
  void* hflux;    // heavy-face corrections per face position (bulk type, 5 per face)
  unsigned long long* hidx;  // per-cell marks of the listed faces (one word per interior cell)
  unsigned int* pflag;       // sensor generation of the last stage with a value outside the UP7 band
  unsigned int pgen;         // sensor generations launched
  unsigned char* cls;        // per-cell sensor class (the scheme its phi alone selects), read by the sweeps
};

namespace {

constexpr int kRedBlocks = 2048;

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

bool same_params(const f3d_params* a, const f3d_params* b) { return std::memcmp(a, b, sizeof(f3d_params)) == 0; }

int check_ws(const f3d_params* p, const f3d_workspace* ws, int prec, const char* who) {
  if (!ws) {
    set_error("%s: workspace is NULL", who);
    return F3D_E_NULL;
  }
  if (ws->prec != prec) {
    set_error("%s: workspace precision %d does not match the entry point", who, ws->prec);
    return F3D_E_WS;
  }
  if (p->Nxp != ws->p.Nxp || p->Nyp != ws->p.Nyp || p->Nzp != ws->p.Nzp || p->NG != ws->p.NG) {
    set_error("%s: workspace was created for another grid", who);
    return F3D_E_WS;
  }
  return F3D_OK;
}

__global__ __launch_bounds__(256) void k_zero(double* __restrict__ a, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    a[t] = 0.0;
}

// Zero the interior of a U-shaped array (5 vars).
__global__ __launch_bounds__(256) void k_zero_interior(KP kp, double* __restrict__ R) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= kp.Nxp) return;
  const long long o = cidx(kp, i + kp.NG, j + kp.NG, k + kp.NG);
#pragma unroll
  for (int n = 0; n < 5; ++n) R[o + n * kp.nc] = 0.0;
}

// ---------------------------------------------------------------- wave64 reductions
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v |= __shfl_xor(v, off, 64);
  return v;
}

template <int NT>
__device__ __forceinline__ double block_min(double v) {
  __shared__ double sm[NT / 64];
  v = wave_min(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sm[wid] = v;
  __syncthreads();
  if (wid == 0) {
    v = lane < NT / 64 ? sm[lane] : INFINITY;
    v = wave_min(v);
  }
  return v;
}

// Global CFL dt, stage 1: per-block minima of 1/(J * max_d(|A_d.u| + c|A_d|)) over interior cells.
template <typename T>
__global__ __launch_bounds__(256) void k_cfl_partial(KP kp, const T* __restrict__ Q, MetricPtrs<T> mp,
                                                     const double* __restrict__ J, double* __restrict__ part) {
  const long long n = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
  double best = INFINITY;
  for (long long t = blockIdx.x * 256LL + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
    const int i = (int)(t % kp.Nxp), j = (int)((t / kp.Nxp) % kp.Nyp), k = (int)(t / ((long long)kp.Nxp * kp.Nyp));
    const long long o = cidx(kp, i + kp.NG, j + kp.NG, k + kp.NG);
    const double u = Q[o + kp.nc], v = Q[o + 2 * kp.nc], w = Q[o + 3 * kp.nc], Tt = Q[o + 5 * kp.nc];
    const double c = sqrt(kp.gamma * kp.Rg * Tt);
    double lam = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double a = mp.m[3 * d][o], b = mp.m[3 * d + 1][o], e = mp.m[3 * d + 2][o];
      lam = fmax(lam, fabs(a * u + b * v + e * w) + c * sqrt(a * a + b * b + e * e));
    }
    best = fmin(best, 1.0 / (J[o] * lam));
  }
  best = block_min<256>(best);
  if (threadIdx.x == 0) part[blockIdx.x] = best;
}

__global__ __launch_bounds__(1024) void k_cfl_final(const double* __restrict__ part, int n, double cfl,
                                                    double* __restrict__ out) {
  double v = INFINITY;
  for (int t = threadIdx.x; t < n; t += 1024) v = fmin(v, part[t]);
  v = block_min<1024>(v);
  if (threadIdx.x == 0) out[0] = cfl * v;
}

__global__ __launch_bounds__(256) void k_nan_partial(const double* __restrict__ U, long long n, int* __restrict__ part) {
  int bad = 0;
  for (long long t = blockIdx.x * 256LL + threadIdx.x; t < n; t += (long long)gridDim.x * 256)
    bad |= !isfinite(U[t]);
  __shared__ int sm[4];
  bad = wave_or(bad);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = bad;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0] | sm[1] | sm[2] | sm[3];
}

__global__ __launch_bounds__(256) void k_nan_final(const int* __restrict__ part, int n, int* __restrict__ out) {
  int bad = 0;
  for (int t = threadIdx.x; t < n; t += 256) bad |= part[t];
  __shared__ int sm[4];
  bad = wave_or(bad);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = bad;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = sm[0] | sm[1] | sm[2] | sm[3];
}

// ---------------------------------------------------------------- metrics  parse_mesh.jl:139-248
__device__ __forceinline__ double dcoord(const double* __restrict__ f, long long o, long long st, int idx, int N) {
  auto F = [&](int s) { return f[o + s * st]; };
  if (idx >= 3 && idx <= N - 4)  // CD6 on 0-based [3, N-4]  (1-based [4, N-3])
    return 1.0 / 60 * (F(3) - F(-3)) - 3.0 / 20 * (F(2) - F(-2)) + 3.0 / 4 * (F(1) - F(-1));
  if (idx <= 2) return 2 * F(1) - 0.5 * F(2) - 1.5 * F(0);  // CD2_L
  return -2 * F(-1) + 0.5 * F(-2) + 1.5 * F(0);             // CD2_R
}

__global__ __launch_bounds__(256) void k_metrics(KP kp, const double* __restrict__ x, const double* __restrict__ y,
                                                 const double* __restrict__ z, MetricPtrs<double> outm,
                                                 double* __restrict__ J, MetricPtrs<double> outs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= kp.Nxt || j >= kp.Nyt || k >= kp.Nzt) return;
  const long long o = cidx(kp, i, j, k);
  const long long st[3] = {1, kp.Nxt, (long long)kp.Nxt * kp.Nyt};
  const int idx[3] = {i, j, k}, N[3] = {kp.Nxt, kp.Nyt, kp.Nzt};
  const double* X[3] = {x, y, z};
  double d[3][3];  // d[c][dir] = d(x_c)/d(xi_dir)
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int e = 0; e < 3; ++e) d[c][e] = dcoord(X[c], o, st[e], idx[e], N[e]);
  const double xa = d[0][0], xb = d[0][1], xc = d[0][2], ya = d[1][0], yb = d[1][1], yc = d[1][2];
  const double za = d[2][0], zb = d[2][1], zc = d[2][2];
  J[o] = 1 / (xa * (yb * zc - yc * zb) - xb * (ya * zc - yc * za) + xc * (ya * zb - yb * za));
  double m[9];
  m[0] = yb * zc - yc * zb;
  m[1] = xc * zb - xb * zc;
  m[2] = xb * yc - xc * yb;
  m[3] = yc * za - ya * zc;
  m[4] = xa * zc - xc * za;
  m[5] = xc * ya - xa * yc;
  m[6] = ya * zb - yb * za;
  m[7] = xb * za - xa * zb;
  m[8] = xa * yb - xb * ya;
#pragma unroll
  for (int q = 0; q < 9; ++q) ((double*)outm.m[q])[o] = m[q];
#pragma unroll
  for (int e = 0; e < 3; ++e)
    if (outs.m[e]) ((double*)outs.m[e])[o] = sqrt(m[3 * e] * m[3 * e] + m[3 * e + 1] * m[3 * e + 1] + m[3 * e + 2] * m[3 * e + 2]);
}

// ---------------------------------------------------------------- RHS and stage
template <typename T>
int rhs_chain(f3d_workspace* ws, const KP& kp, const T* Q, const MetricPtrs<T>& m, const double* J, T* phi,
              hipStream_t s) {
  T* Fp = (T*)ws->Fp;
  T* Fm = (T*)ws->Fm;
  T* F[3] = {(T*)ws->Fx, (T*)ws->Fy, (T*)ws->Fz};
  int rc = launch_shock_sensor<T>(kp, phi, Q, s);
  for (int d = 0; d < 3 && !rc; ++d) {
    rc = launch_flux_split<T>(kp, Q, Fp, Fm, nullptr, m.m[3 * d], m.m[3 * d + 1], m.m[3 * d + 2], s);
    if (!rc)
      rc = launch_advect<T>(kp, d, F[d], phi, nullptr, Fp, Fm, Q, m.m[3 * d], m.m[3 * d + 1], m.m[3 * d + 2], s);
  }
  if (!rc && kp.viscous) rc = launch_viscous<T>(kp, (T*)ws->Fvx, (T*)ws->Fvy, (T*)ws->Fvz, Q, m, J, s);
  return rc;
}

// Kernel-by-kernel finite-volume branch (solver.jl:21-31) into the workspace face arrays.
template <typename T>
int rhs_chain_fvm(f3d_workspace* ws, const KP& kp, const T* Q, const double* U, const MetricPtrs<T>& m,
                  const double* J, const T* const* geo, T* phi, hipStream_t s) {
  T* F[3] = {(T*)ws->Fx, (T*)ws->Fy, (T*)ws->Fz};
  int rc = launch_shock_sensor<T>(kp, phi, Q, s);
  for (int d = 0; d < 3 && !rc; ++d)
    rc = launch_fvm_face<T>(kp, d, Q, U, phi, nullptr, m.m[3 * d], m.m[3 * d + 1], m.m[3 * d + 2], F[d], nullptr,
                            nullptr, nullptr, geo[d], s);
  if (!rc && kp.viscous) rc = launch_viscous<T>(kp, (T*)ws->Fvx, (T*)ws->Fvy, (T*)ws->Fvz, Q, m, J, s);
  return rc;
}

// U, geo: finite-volume branch when geo != nullptr.
template <typename T>
int rhs_impl(const f3d_params* p, f3d_workspace* ws, double* R, const T* Q, const T* const* metrics,
             const double* J, T* phi, hipStream_t s, const double* U = nullptr, const T* const* geo = nullptr) {
  KP kp = make_kp(p);
  MetricPtrs<T> m;
  for (int q = 0; q < 9; ++q) m.m[q] = metrics[q];
  T* ph = phi ? phi : (T*)ws->phi;
  int rc;
  if (use_fused()) {
    if (geo)
      rc = fvm_fused_rhs<T>(kp, ws_view<T>(ws), Q, U, m, J, geo, ph, R, s);
    else
      rc = fused_rhs<T>(kp, ws_view<T>(ws), Q, m, J, ph, R, s);
  } else {
    rc = geo ? rhs_chain_fvm<T>(ws, kp, Q, U, m, J, geo, ph, s) : rhs_chain<T>(ws, kp, Q, m, J, ph, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_zero_interior, dim3((kp.Nxp + 255) / 256, kp.Nyp, kp.Nzp), dim3(256), 0, s, kp, R);
    rc = launch_div<T>(kp, R, (T*)ws->Fx, (T*)ws->Fy, (T*)ws->Fz, (T*)ws->Fvx, (T*)ws->Fvy, (T*)ws->Fvz, 1.0, J, s);
  }
  return rc;
}

}  // namespace

namespace f3d {
template <typename T>
int stage_io(const f3d_params* p, f3d_workspace* ws, int KRK, double* U, double* Un, const T* Qin, T* Qout,
             const T* const* metrics, const double* J, double dt, const T* inflow, hipStream_t s,
             const T* const* geo, double* Uo) {
  KP kp = make_kp(p);
  if (Uo && (geo || !use_fused())) {
    set_error("stage: a separate output state is supported by the fused finite-difference stage only");
    return F3D_E_ARG;
  }
  MetricPtrs<T> m;
  for (int q = 0; q < 9; ++q) m.m[q] = metrics[q];
  int rc = F3D_OK;
  // local time stepping: the per-cell dt of the step from its first stage's state (solver.jl:419-421)
  if (kp.lts && KRK == 1) rc = launch_compute_dt<T>(kp, ws->ldt, Qin, J, nullptr, nullptr, nullptr, &m, s);
  if (rc) return rc;
  if (use_fused()) {
    rc = geo ? fvm_fused_stage<T>(kp, ws_view<T>(ws), KRK, U, Un, Qin, Qout, m, J, geo, dt, s)
                : fused_stage<T>(kp, ws_view<T>(ws), KRK, U, Un, Qin, Qout, m, J, dt, s, Uo);
  } else {
    if (Qout != Qin) F3D_HIP(hipMemcpyAsync(Qout, Qin, sizeof(T) * 6 * kp.nc, hipMemcpyDeviceToDevice, s));
    if (KRK == 1) F3D_HIP(hipMemcpyAsync(Un, U, sizeof(double) * 5 * kp.nc, hipMemcpyDeviceToDevice, s));  // solver.jl:418
    rc = geo ? rhs_chain_fvm<T>(ws, kp, Qout, U, m, J, geo, (T*)ws->phi, s)
                : rhs_chain<T>(ws, kp, Qout, m, J, (T*)ws->phi, s);
    if (!rc)
      rc = launch_div<T>(kp, U, (T*)ws->Fx, (T*)ws->Fy, (T*)ws->Fz, (T*)ws->Fvx, (T*)ws->Fvy, (T*)ws->Fvz, dt, J, s,
                         kp.lts ? ws->ldt : nullptr);
    if (!rc && KRK == 2) rc = launch_lin_comb(kp, U, Un, 5, 0.25, 0.75, s);
    if (!rc && KRK == 3) rc = launch_lin_comb(kp, U, Un, 5, 2.0 / 3.0, 1.0 / 3.0, s);
    if (!rc) rc = launch_c2prim<T>(kp, U, Qout, s);
  }
  // exchange_ghost (mpi.jl:2-103) is not launched here: on one rank it copies the ghost slabs of the
  // periodic dimensions, and fill_d of a periodic dimension performs the same copy over its full slab
  // afterwards.  Every fill reads only cells of its own line (same (j,k) for x, (i,k) for y, (i,j) for
  // z), and each later fill rewrites the whole slab of its dimension, so the ghost values the exchange
  // would have set first are all overwritten with the same final values.
  // U ghosts are read only by the finite-volume reconstruction and the filters; fillGhost refills them
  // from Q's there (boundary.jl fills them always)
  if (!rc) rc = launch_fill<T>(kp, Qout, (geo || p->filtering) ? (Uo ? Uo : U) : nullptr, inflow, s);
  return rc;
}

template <typename T>
int steps_io(const f3d_params* p, f3d_workspace* ws, int nsteps, double* U, double* Un, T* Q,
             const T* const* metrics, const double* J, double dt, const T* inflow, hipStream_t s,
             const T* const* geo) {
  const KP kp = make_kp(p);
  // stage outputs rotate Q -> Qa -> Qb -> Q when the stage cannot write its input: no copy per step
  const bool rot = stage_needs_separate_q(kp);
  T* Qa = rot ? (T*)ws->Qa : Q;
  T* Qb = rot ? (T*)ws->Qb : Q;
  // Fused finite-difference stages, every step but the call's last: U keeps U^n while the stages write
  // U^1 and U^2 into Un, and the third stage blends Un's U^2 with U's U^n into U, so stage 1 copies
  // nothing (solver.jl:418's Un <- U is implicit).  The last step takes the reference's sequence, so
  // that Un holds that step's U^n on return.
  const bool swap = use_fused() && !geo;
  for (int t = 1; t <= nsteps; ++t) {
    if (swap && t < nsteps) {
      int rc = stage_io<T>(p, ws, 1, U, nullptr, Q, Qa, metrics, J, dt, inflow, s, geo, Un);
      if (!rc) rc = stage_io<T>(p, ws, 2, Un, U, Qa, Qb, metrics, J, dt, inflow, s, geo, Un);
      if (!rc) rc = stage_io<T>(p, ws, 3, Un, U, Qb, Q, metrics, J, dt, inflow, s, geo, U);
      if (!rc && p->filtering && t % p->filtering_interval == 0)
        rc = launch_filter_block<T>(kp, U, Un, Q, ws->fsc, p->filtering_nonlinear, p->filtering_rth, p->filtering_s0, s);
      if (rc) return rc;
      continue;
    }
    int rc = stage_io<T>(p, ws, 1, U, Un, Q, Qa, metrics, J, dt, inflow, s, geo);
    if (!rc) rc = stage_io<T>(p, ws, 2, U, Un, Qa, Qb, metrics, J, dt, inflow, s, geo);
    if (!rc) rc = stage_io<T>(p, ws, 3, U, Un, Qb, Q, metrics, J, dt, inflow, s, geo);
    // the filtering block after the steps whose index (counted from this call's first) is a multiple of
    // filtering_interval (solver.jl:446-470); Q stays as the last stage made it
    if (!rc && p->filtering && t % p->filtering_interval == 0)
      rc = launch_filter_block<T>(kp, U, Un, Q, ws->fsc, p->filtering_nonlinear, p->filtering_rth, p->filtering_s0, s);
    if (rc) return rc;
  }
  return F3D_OK;
}

int step_keep_stages(const f3d_params* p, f3d_workspace* ws, double* U, double* Un, double* Q,
                     const double* const* metrics, const double* J, double dt, const double* inflow, double* U1,
                     double* U2, hipStream_t s, bool third, double* Uout) {
  const KP kp = make_kp(p);
  const bool rot = stage_needs_separate_q(kp);
  double* Qa = rot ? (double*)ws->Qa : Q;
  double* Qb = rot ? (double*)ws->Qb : Q;
  const size_t vb = sizeof(double) * 5 * kp.nc;
  if (use_fused()) {  // each stage reads its input state and writes the next one: no copies
    int rc = stage_io<double>(p, ws, 1, U, Un, Q, Qa, metrics, J, dt, inflow, s, nullptr, U1);
    if (!rc) rc = stage_io<double>(p, ws, 2, U1, Un, Qa, Qb, metrics, J, dt, inflow, s, nullptr, U2);
    if (!rc && third) rc = stage_io<double>(p, ws, 3, U2, Un, Qb, Q, metrics, J, dt, inflow, s, nullptr, Uout ? Uout : U);
    return rc;
  }
  if (Uout) {
    set_error("step_keep_stages: a separate output state needs the fused stage");
    return F3D_E_ARG;
  }
  int rc = stage_io<double>(p, ws, 1, U, Un, Q, Qa, metrics, J, dt, inflow, s);
  if (!rc) F3D_HIP(hipMemcpyAsync(U1, U, vb, hipMemcpyDeviceToDevice, s));
  if (!rc) rc = stage_io<double>(p, ws, 2, U, Un, Qa, Qb, metrics, J, dt, inflow, s);
  if (!rc) F3D_HIP(hipMemcpyAsync(U2, U, vb, hipMemcpyDeviceToDevice, s));
  if (!rc && third) rc = stage_io<double>(p, ws, 3, U, Un, Qb, Q, metrics, J, dt, inflow, s);
  return rc;
}

int step_keep_q(const f3d_params* p, f3d_workspace* ws, double* U, double* Un, const double* Qin,
                const double* const* metrics, const double* J, double dt, const double* inflow, double* U1,
                double* U2, double* Q1, double* Q2, double* Qnext, hipStream_t s) {
  if (!use_fused()) {
    set_error("step_keep_q: needs the fused stage");
    return F3D_E_ARG;
  }
  int rc = stage_io<double>(p, ws, 1, U, Un, Qin, Q1, metrics, J, dt, inflow, s, nullptr, U1);
  if (!rc) rc = stage_io<double>(p, ws, 2, U1, Un, Q1, Q2, metrics, J, dt, inflow, s, nullptr, U2);
  if (!rc) rc = stage_io<double>(p, ws, 3, U2, Un, Q2, Qnext, metrics, J, dt, inflow, s, nullptr, U);
  return rc;
}

template <typename T>
int stage_inplace(const f3d_params* p, f3d_workspace* ws, int KRK, double* U, double* Un, T* Q,
                  const T* const* metrics, const double* J, double dt, const T* inflow, hipStream_t s,
                  const T* const* geo = nullptr) {
  const KP kp = make_kp(p);
  if (!stage_needs_separate_q(kp)) return stage_io<T>(p, ws, KRK, U, Un, Q, Q, metrics, J, dt, inflow, s, geo);
  T* Qa = (T*)ws->Qa;
  const int rc = stage_io<T>(p, ws, KRK, U, Un, Q, Qa, metrics, J, dt, inflow, s, geo);
  if (rc) return rc;
  F3D_HIP(hipMemcpyAsync(Q, Qa, sizeof(T) * 6 * kp.nc, hipMemcpyDeviceToDevice, s));
  return F3D_OK;
}

template int stage_io<double>(const f3d_params*, f3d_workspace*, int, double*, double*, const double*, double*,
                              const double* const*, const double*, double, const double*, hipStream_t,
                              const double* const*, double*);
template int stage_io<float>(const f3d_params*, f3d_workspace*, int, double*, double*, const float*, float*,
                             const float* const*, const double*, double, const float*, hipStream_t,
                             const float* const*, double*);
template int steps_io<double>(const f3d_params*, f3d_workspace*, int, double*, double*, double*,
                              const double* const*, const double*, double, const double*, hipStream_t,
                              const double* const*);
}  // namespace f3d

namespace {

template <typename T>
int check_stage_args(const f3d_params* p, const f3d_workspace* ws, int prec, const void* U, const void* Un,
                     const void* Q, const T* const* metrics, const double* J, const T* inflow, const char* who) {
  int rc = f3d_params_check(p);
  if (rc) return rc;
  rc = check_ws(p, ws, prec, who);
  if (rc) return rc;
  if (!U || !Un || !Q || !metrics || !J) {
    set_error("%s: NULL array argument", who);
    return F3D_E_NULL;
  }
  for (int q = 0; q < 9; ++q)
    if (!metrics[q]) {
      set_error("%s: metric %d is NULL", who, q);
      return F3D_E_NULL;
    }
  if (p->bc[0] == F3D_BC_INFLOW && !inflow) {
    set_error("%s: F3D_BC_INFLOW needs an inflow profile", who);
    return F3D_E_NULL;
  }
  return F3D_OK;
}

// The entries without _fvm serve the finite-difference branch, the _fvm entries the finite-volume
// one; each rejects the other's params.
template <typename T>
int check_branch(const f3d_params* p, bool fvm_entry, const double* U, const T* const* geo, const char* who) {
  if (fvm_entry && p->fvm == F3D_FDM) {
    set_error("%s: params select the finite-difference branch (fvm = F3D_FDM); use the entry without _fvm", who);
    return F3D_E_ARG;
  }
  if (!fvm_entry && p->fvm != F3D_FDM) {
    set_error("%s: params select the finite-volume branch (fvm = %d); use the _fvm entry", who, p->fvm);
    return F3D_E_ARG;
  }
  if (fvm_entry) {
    if (!U || !geo || !geo[0] || !geo[1] || !geo[2]) {
      set_error("%s: the finite-volume branch needs U and the face geometry arrays geo[3]", who);
      return F3D_E_NULL;
    }
  }
  return F3D_OK;
}

}  // namespace

// ================================================================ extern "C"
#define F3D_API extern "C" __attribute__((visibility("default")))

F3D_API int f3d_workspace_create(const f3d_params* p, int32_t precision, f3d_workspace** out) {
  int rc = f3d_params_check(p);
  if (rc) return rc;
  if (!out) {
    set_error("workspace_create: out is NULL");
    return F3D_E_NULL;
  }
  if (precision != F3D_PREC_F64 && precision != F3D_PREC_F32B) {
    set_error("workspace_create: unknown precision %d", precision);
    return F3D_E_ARG;
  }
  *out = nullptr;
  const size_t es = precision == F3D_PREC_F64 ? 8 : 4;
  const size_t nc = f3d_cell_count(p);
  const size_t Nx = p->Nxp, Ny = p->Nyp, Nz = p->Nzp, NG = p->NG;
  size_t sz[22];
  sz[0] = nc * es;                               // phi
  sz[1] = nc * 5 * es;                           // Fp
  sz[2] = nc * 5 * es;                           // Fm
  sz[3] = (Nx + 1) * Ny * Nz * 5 * es;           // Fx
  sz[4] = Nx * (Ny + 1) * Nz * 5 * es;           // Fy
  sz[5] = Nx * Ny * (Nz + 1) * 5 * es;           // Fz
  const size_t nv = (Nx + NG) * (Ny + NG) * (Nz + NG) * 4 * es;
  sz[6] = sz[7] = sz[8] = nv;                    // Fv_x, Fv_y, Fv_z
  sz[9] = Nx * Ny * Nz * 5 * es;                 // R
  sz[10] = kRedBlocks * sizeof(double);          // red
  sz[11] = kRedBlocks * sizeof(int);             // flag
  // heavy-face list in per-wave runs: room for the largest layout of the finite-difference sweeps (this
  // precision), the finite-volume sweeps and the adjoint's list kernels
  const KP kp0 = make_kp(p);
  const HList hls[3] = {fdm_heavy_layout(kp0, (int)es), fvm_heavy_layout(kp0, (int)es), adj_heavy_layout(kp0)};
  unsigned long long hslots = 0, hruns = 0;
  for (const HList& h : hls) {
    hslots = h.slots() > hslots ? h.slots() : hslots;
    hruns = h.nruns() > hruns ? h.nruns() : hruns;
  }
  sz[12] = hslots * sizeof(unsigned long long);  // heavy-face list
  sz[13] = hruns * sizeof(unsigned int);         // entries per run
  sz[14] = sz[15] = nc * 6 * es;                 // Qa, Qb
  sz[16] = nc * sizeof(double);                  // ldt
  sz[17] = nc * sizeof(double);                  // fsc
  sz[18] = (size_t)heavy_face_count(kp0) * 5 * es;                 // heavy-face corrections per face
  sz[19] = (size_t)hmark_words(kp0) * sizeof(unsigned long long);  // per-cell mark bytes of the listed faces
  sz[20] = sizeof(unsigned int);                 // sensor band flag
  sz[21] = nc;                                   // sensor class per cell
  size_t total = 0;
  for (size_t b : sz) total += align_up(b);
  void* base = nullptr;
  F3D_HIP(hipMalloc(&base, total));
  f3d_workspace* ws = new f3d_workspace();
  ws->p = *p;
  ws->prec = precision;
  ws->bytes = total;
  ws->base = base;
  char* c = (char*)base;
  void** slots[22] = {&ws->phi, &ws->Fp, &ws->Fm, &ws->Fx, &ws->Fy, &ws->Fz, &ws->Fvx, &ws->Fvy, &ws->Fvz,
                      &ws->R, (void**)&ws->red, (void**)&ws->flag, (void**)&ws->heavy, (void**)&ws->hcount,
                      &ws->Qa, &ws->Qb, (void**)&ws->ldt, (void**)&ws->fsc, &ws->hflux, (void**)&ws->hidx,
                      (void**)&ws->pflag, (void**)&ws->cls};
  for (int i = 0; i < 22; ++i) {
    *slots[i] = c;
    c += align_up(sz[i]);
  }
  ws->nred = kRedBlocks;
  // the marks start cleared (zero fill below); the apply pass leaves them cleared
  ws->pgen = 0;    // the flag starts cleared: generations count from 1
  ws->hcap = hslots;
  ws->hncap = hruns;
  // zeroed once: inviscid runs read zero viscous fluxes in div (div.jl:45-48 adds them always)
  hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, 0, (double*)base, (long long)(total / 8));
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    hipFree(base);
    delete ws;
    set_error("workspace_create: zero fill failed: %s", hipGetErrorString(e));
    return (int)e;
  }
  *out = ws;
  return F3D_OK;
}

F3D_API int f3d_workspace_destroy(f3d_workspace* ws) {
  if (!ws) return F3D_OK;
  hipError_t e = hipFree(ws->base);
  delete ws;
  if (e != hipSuccess) {
    set_error("workspace_destroy: %s", hipGetErrorString(e));
    return (int)e;
  }
  return F3D_OK;
}

F3D_API size_t f3d_workspace_bytes(const f3d_workspace* ws) { return ws ? ws->bytes : 0; }

namespace f3d {
template <typename T>
WsView<T> ws_view(f3d_workspace* ws) {
  WsView<T> v;
  v.R = (T*)ws->R;
  v.Fvx = (T*)ws->Fvx;
  v.Fvy = (T*)ws->Fvy;
  v.Fvz = (T*)ws->Fvz;
  v.phi = (T*)ws->phi;
  v.heavy = ws->heavy;
  v.hcount = ws->hcount;
  v.hcap = ws->hcap;
  v.hncap = ws->hncap;
  v.ldt = ws->ldt;
  v.hflux = (T*)ws->hflux;
  v.hidx = ws->hidx;
  v.pflag = ws->pflag;
  v.pgen = &ws->pgen;
  v.cls = ws->cls;
  return v;
}
template WsView<double> ws_view<double>(f3d_workspace*);
template WsView<float> ws_view<float>(f3d_workspace*);
}  // namespace f3d

F3D_API int f3d_rhs_f64(const f3d_params* p, f3d_workspace* ws, double* R, const double* Q,
                        const double* const* metrics, const double* J, double* phi, f3d_stream_t s) {
  int rc = check_stage_args<double>(p, ws, F3D_PREC_F64, R, R, Q, metrics, J, (const double*)1, "rhs_f64");
  if (rc) return rc;
  rc = check_branch<double>(p, false, nullptr, nullptr, "rhs_f64");
  if (rc) return rc;
  return rhs_impl<double>(p, ws, R, Q, metrics, J, phi, (hipStream_t)s);
}
F3D_API int f3d_rhs_f32b(const f3d_params* p, f3d_workspace* ws, double* R, const float* Q,
                         const float* const* metrics, const double* J, float* phi, f3d_stream_t s) {
  int rc = check_stage_args<float>(p, ws, F3D_PREC_F32B, R, R, Q, metrics, J, (const float*)1, "rhs_f32b");
  if (rc) return rc;
  rc = check_branch<float>(p, false, nullptr, nullptr, "rhs_f32b");
  if (rc) return rc;
  return rhs_impl<float>(p, ws, R, Q, metrics, J, phi, (hipStream_t)s);
}

// ---- FP64 RHS without allocation: the fused RHS of f3d_rhs_f64 with its scratch (the interior RHS
// accumulator, the heavy-face runs and their counts, and phi when the caller passes none) carved from
// one file-scope device array instead of a workspace.  Each GPU the library runs on holds its own copy
// of the array (kNoAllocScratchBytes) until the process ends; its address is looked up once per device.
// Only the per-cell marks of the heavy faces are cleared on every call (by a fill kernel, since the
// previous call may have laid the array out for another grid); everything else is written before it is
// read: the x+y march (or the x sweep) writes every accumulator entry before the z march adds to it, each
// sweep wave writes its run's count, the heavy faces' forming pass writes the corrections of the faces it
// marks (the update reads only those), and the sensor writes phi.
namespace {
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct ScratchSlot {
  std::once_flag once;
  void* ptr = nullptr;
  hipError_t err = hipSuccess;
};
ScratchSlot g_scratch[kMaxDevices];

// Device address of `scratch` on the current device (looked up with that device current, once).
int scratch_address(void** out, const char* who) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("%s: hipGetDevice: %s", who, hipGetErrorString(e));
    return (int)e;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("%s: device %d beyond the %d the library tracks", who, dev, kMaxDevices);
    return F3D_E_ARG;
  }
  ScratchSlot& sl = g_scratch[dev];
  std::call_once(sl.once, [&] { sl.err = hipGetSymbolAddress(&sl.ptr, HIP_SYMBOL(scratch)); });
  if (sl.err != hipSuccess || !sl.ptr) {
    set_error("%s: hipGetSymbolAddress(scratch): %s", who, hipGetErrorString(sl.err));
    return sl.err != hipSuccess ? (int)sl.err : F3D_E_ARG;
  }
  *out = sl.ptr;
  return F3D_OK;
}

// Scratch layout of the no-allocation RHS: offsets of R, the heavy-face runs, their counts and phi.
struct NoAllocLayout {
  size_t R, heavy, hcount, hflux, hidx, cls, phi, total;
  unsigned long long hcap, hncap;
  long long nidx;  // 8-byte words of the listed faces' mark bytes (one byte per interior cell)
};
NoAllocLayout noalloc_layout(const f3d_params* p, bool with_phi) {
  const KP kp = make_kp(p);
  const HList h = fdm_heavy_layout(kp, 8);
  NoAllocLayout L;
  L.hcap = h.slots();
  L.hncap = h.nruns();
  size_t off = 0;
  L.R = off;
  off += align_up((size_t)p->Nxp * p->Nyp * p->Nzp * 5 * sizeof(double));
  L.heavy = off;
  off += align_up(L.hcap * sizeof(unsigned long long));
  L.hcount = off;
  off += align_up(L.hncap * sizeof(unsigned int));
  L.hflux = off;
  off += align_up((size_t)heavy_face_count(kp) * 5 * sizeof(double));
  L.hidx = off;
  L.nidx = hmark_words(kp);
  off += align_up((size_t)(L.nidx + 1) * sizeof(unsigned long long));  // the marks, then the sensor band flag
  L.cls = off;
  off += align_up(f3d_cell_count(p));
  L.phi = off;
  if (with_phi) off += align_up(f3d_cell_count(p) * sizeof(double));
  L.total = off;
  return L;
}
}  // namespace

F3D_API int f3d_rhs_noalloc_f64(const f3d_params* p, double* R, const double* Q, const double* const* metrics,
                                const double* J, double* phi, f3d_stream_t s) {
  const char* who = "rhs_noalloc_f64";
  int rc = f3d_params_check(p);
  if (rc) return rc;
  if (!R || !Q || !metrics || !J) {
    set_error("%s: NULL array argument", who);
    return F3D_E_NULL;
  }
  for (int q = 0; q < 9; ++q)
    if (!metrics[q]) {
      set_error("%s: metric %d is NULL", who, q);
      return F3D_E_NULL;
    }
  rc = check_branch<double>(p, false, nullptr, nullptr, who);
  if (rc) return rc;
  const KP kp = make_kp(p);
  if (!use_fused() || fused_needs_fv(kp)) {
    set_error("%s: runs the fused engine with the viscous fluxes inside the update march only (no LES model, "
              "F3D_RHS_ENGINE and F3D_VISC_FUSED unset); use f3d_rhs_f64 with a workspace", who);
    return F3D_E_ARG;
  }
  const NoAllocLayout L = noalloc_layout(p, phi == nullptr);
  if (L.total > kNoAllocScratchBytes) {
    set_error("%s: a %dx%dx%d grid needs %zu bytes of scratch, the no-allocation entry holds %zu "
              "(use f3d_rhs_f64 with a workspace)", who, p->Nxp, p->Nyp, p->Nzp, L.total, kNoAllocScratchBytes);
    return F3D_E_DIMS;
  }
  void* base = nullptr;
  rc = scratch_address(&base, who);
  if (rc) return rc;
  char* c = (char*)base;
  WsView<double> v{};
  v.R = (double*)(c + L.R);
  v.heavy = (unsigned long long*)(c + L.heavy);
  v.hcount = (unsigned int*)(c + L.hcount);
  v.hcap = L.hcap;
  v.hncap = L.hncap;
  v.hflux = (double*)(c + L.hflux);
  v.hidx = (unsigned long long*)(c + L.hidx);
  unsigned int gen = 0;  // the flag word after the marks, cleared with them: this call's sensor is generation 1
  v.pflag = (unsigned int*)(v.hidx + L.nidx);
  v.pgen = &gen;
  v.phi = phi ? phi : (double*)(c + L.phi);
  v.cls = (unsigned char*)(c + L.cls);
  clear_hidx(v.hidx, L.nidx + 1, (hipStream_t)s);
  MetricPtrs<double> m;
  for (int q = 0; q < 9; ++q) m.m[q] = metrics[q];
  return fused_rhs<double>(kp, v, Q, m, J, v.phi, R, (hipStream_t)s);
}

F3D_API int f3d_rk3_stage_f64(const f3d_params* p, f3d_workspace* ws, int32_t KRK, double* U, double* Un,
                              double* Q, const double* const* metrics, const double* J, double dt,
                              const double* inflow, f3d_stream_t s) {
  int rc = check_stage_args<double>(p, ws, F3D_PREC_F64, U, Un, Q, metrics, J, inflow, "rk3_stage_f64");
  if (rc) return rc;
  rc = check_branch<double>(p, false, nullptr, nullptr, "rk3_stage_f64");
  if (rc) return rc;
  if (KRK < 1 || KRK > 3) {
    set_error("rk3_stage_f64: KRK=%d not in 1..3", KRK);
    return F3D_E_ARG;
  }
  return stage_inplace<double>(p, ws, KRK, U, Un, Q, metrics, J, dt, inflow, (hipStream_t)s);
}
F3D_API int f3d_rk3_stage_f32b(const f3d_params* p, f3d_workspace* ws, int32_t KRK, double* U, double* Un,
                               float* Q, const float* const* metrics, const double* J, double dt,
                               const float* inflow, f3d_stream_t s) {
  int rc = check_stage_args<float>(p, ws, F3D_PREC_F32B, U, Un, Q, metrics, J, inflow, "rk3_stage_f32b");
  if (rc) return rc;
  rc = check_branch<float>(p, false, nullptr, nullptr, "rk3_stage_f32b");
  if (rc) return rc;
  if (KRK < 1 || KRK > 3) {
    set_error("rk3_stage_f32b: KRK=%d not in 1..3", KRK);
    return F3D_E_ARG;
  }
  return stage_inplace<float>(p, ws, KRK, U, Un, Q, metrics, J, dt, inflow, (hipStream_t)s);
}

F3D_API int f3d_time_step_f64(const f3d_params* p, f3d_workspace* ws, int32_t nsteps, double* U, double* Un,
                              double* Q, const double* const* metrics, const double* J, double dt,
                              const double* inflow, f3d_stream_t s) {
  int rc = check_stage_args<double>(p, ws, F3D_PREC_F64, U, Un, Q, metrics, J, inflow, "time_step_f64");
  if (rc) return rc;
  rc = check_branch<double>(p, false, nullptr, nullptr, "time_step_f64");
  if (rc) return rc;
  if (nsteps < 0) {
    set_error("time_step_f64: nsteps < 0");
    return F3D_E_ARG;
  }
  return steps_io<double>(p, ws, nsteps, U, Un, Q, metrics, J, dt, inflow, (hipStream_t)s);
}
F3D_API int f3d_time_step_f32b(const f3d_params* p, f3d_workspace* ws, int32_t nsteps, double* U, double* Un,
                               float* Q, const float* const* metrics, const double* J, double dt,
                               const float* inflow, f3d_stream_t s) {
  int rc = check_stage_args<float>(p, ws, F3D_PREC_F32B, U, Un, Q, metrics, J, inflow, "time_step_f32b");
  if (rc) return rc;
  rc = check_branch<float>(p, false, nullptr, nullptr, "time_step_f32b");
  if (rc) return rc;
  if (nsteps < 0) {
    set_error("time_step_f32b: nsteps < 0");
    return F3D_E_ARG;
  }
  return steps_io<float>(p, ws, nsteps, U, Un, Q, metrics, J, dt, inflow, (hipStream_t)s);
}

F3D_API int f3d_rhs_fvm_f64(const f3d_params* p, f3d_workspace* ws, double* R, const double* Q, const double* U,
                            const double* const* metrics, const double* J, const double* const* geo, double* phi,
                            f3d_stream_t s) {
  int rc = check_stage_args<double>(p, ws, F3D_PREC_F64, R, R, Q, metrics, J, (const double*)1, "rhs_fvm_f64");
  if (!rc) rc = check_branch<double>(p, true, U, geo, "rhs_fvm_f64");
  if (rc) return rc;
  return rhs_impl<double>(p, ws, R, Q, metrics, J, phi, (hipStream_t)s, U, geo);
}
F3D_API int f3d_rhs_fvm_f32b(const f3d_params* p, f3d_workspace* ws, double* R, const float* Q, const double* U,
                             const float* const* metrics, const double* J, const float* const* geo, float* phi,
                             f3d_stream_t s) {
  int rc = check_stage_args<float>(p, ws, F3D_PREC_F32B, R, R, Q, metrics, J, (const float*)1, "rhs_fvm_f32b");
  if (!rc) rc = check_branch<float>(p, true, U, geo, "rhs_fvm_f32b");
  if (rc) return rc;
  return rhs_impl<float>(p, ws, R, Q, metrics, J, phi, (hipStream_t)s, U, geo);
}

#define F3D_STAGE_FVM(SUF, T, PREC)                                                                              \
  F3D_API int f3d_rk3_stage_fvm_##SUF(const f3d_params* p, f3d_workspace* ws, int32_t KRK, double* U, double* Un, \
                                      T* Q, const T* const* metrics, const double* J, const T* const* geo,        \
                                      double dt, const T* inflow, f3d_stream_t s) {                              \
    int rc = check_stage_args<T>(p, ws, PREC, U, Un, Q, metrics, J, inflow, "rk3_stage_fvm_" #SUF);              \
    if (!rc) rc = check_branch<T>(p, true, U, geo, "rk3_stage_fvm_" #SUF);                                       \
    if (rc) return rc;                                                                                           \
    if (KRK < 1 || KRK > 3) {                                                                                    \
      set_error("rk3_stage_fvm_" #SUF ": KRK=%d not in 1..3", KRK);                                              \
      return F3D_E_ARG;                                                                                          \
    }                                                                                                            \
    return stage_inplace<T>(p, ws, KRK, U, Un, Q, metrics, J, dt, inflow, (hipStream_t)s, geo);                  \
  }                                                                                                              \
  F3D_API int f3d_time_step_fvm_##SUF(const f3d_params* p, f3d_workspace* ws, int32_t nsteps, double* U,          \
                                      double* Un, T* Q, const T* const* metrics, const double* J,                 \
                                      const T* const* geo, double dt, const T* inflow, f3d_stream_t s) {         \
    int rc = check_stage_args<T>(p, ws, PREC, U, Un, Q, metrics, J, inflow, "time_step_fvm_" #SUF);              \
    if (!rc) rc = check_branch<T>(p, true, U, geo, "time_step_fvm_" #SUF);                                       \
    if (rc) return rc;                                                                                           \
    if (nsteps < 0) {                                                                                            \
      set_error("time_step_fvm_" #SUF ": nsteps < 0");                                                           \
      return F3D_E_ARG;                                                                                          \
    }                                                                                                            \
    return steps_io<T>(p, ws, nsteps, U, Un, Q, metrics, J, dt, inflow, (hipStream_t)s, geo);                    \
  }
F3D_STAGE_FVM(f64, double, F3D_PREC_F64)
F3D_STAGE_FVM(f32b, float, F3D_PREC_F32B)

template <typename T>
static int cfl_impl(const f3d_params* p, f3d_workspace* ws, int prec, double* dt_out, const T* Q,
                    const T* const* metrics, const double* J, double cfl, hipStream_t s, const char* who) {
  int rc = check_stage_args<T>(p, ws, prec, dt_out, dt_out, Q, metrics, J, (const T*)1, who);
  if (rc) return rc;
  if (!(cfl > 0)) {
    set_error("%s: cfl must be > 0", who);
    return F3D_E_ARG;
  }
  KP kp = make_kp(p);
  MetricPtrs<T> m;
  for (int q = 0; q < 9; ++q) m.m[q] = metrics[q];
  const long long n = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
  const int nb = (int)std::min<long long>(ws->nred, (n + 255) / 256);
  hipLaunchKernelGGL(k_cfl_partial<T>, dim3(nb), dim3(256), 0, s, kp, Q, m, J, ws->red);
  hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(1024), 0, s, ws->red, nb, cfl, dt_out);
  return check_launch(who);
}

F3D_API int f3d_cfl_dt_f64(const f3d_params* p, f3d_workspace* ws, double* dt_out, const double* Q,
                           const double* const* metrics, const double* J, double cfl, f3d_stream_t s) {
  return cfl_impl<double>(p, ws, F3D_PREC_F64, dt_out, Q, metrics, J, cfl, (hipStream_t)s, "cfl_dt_f64");
}
F3D_API int f3d_cfl_dt_f32b(const f3d_params* p, f3d_workspace* ws, double* dt_out, const float* Q,
                            const float* const* metrics, const double* J, double cfl, f3d_stream_t s) {
  return cfl_impl<float>(p, ws, F3D_PREC_F32B, dt_out, Q, metrics, J, cfl, (hipStream_t)s, "cfl_dt_f32b");
}

// The filtering block of the time loop (solver.jl:446-470) with the params' filter settings.
template <typename T>
static int filter_step_impl(const f3d_params* p, f3d_workspace* ws, int prec, double* U, double* Un, const T* Q,
                            hipStream_t s, const char* who) {
  int rc = f3d_params_check(p);
  if (!rc) rc = check_ws(p, ws, prec, who);
  if (rc) return rc;
  if (!U || !Un || !Q) {
    set_error("%s: NULL array argument", who);
    return F3D_E_NULL;
  }
  if (p->filtering_interval < 1 || !(p->filtering_s0 >= 0)) {
    set_error("%s: needs filtering_interval >= 1 and filtering_s0 >= 0", who);
    return F3D_E_ARG;
  }
  return launch_filter_block<T>(make_kp(p), U, Un, Q, ws->fsc, p->filtering_nonlinear, p->filtering_rth,
                                p->filtering_s0, s);
}
F3D_API int f3d_filter_step_f64(const f3d_params* p, f3d_workspace* ws, double* U, double* Un, const double* Q,
                                f3d_stream_t s) {
  return filter_step_impl<double>(p, ws, F3D_PREC_F64, U, Un, Q, (hipStream_t)s, "filter_step_f64");
}
F3D_API int f3d_filter_step_f32b(const f3d_params* p, f3d_workspace* ws, double* U, double* Un, const float* Q,
                                 f3d_stream_t s) {
  return filter_step_impl<float>(p, ws, F3D_PREC_F32B, U, Un, Q, (hipStream_t)s, "filter_step_f32b");
}

F3D_API int f3d_nan_check(const f3d_params* p, f3d_workspace* ws, int32_t* flag_out, const double* U, int32_t NV,
                          f3d_stream_t s) {
  int rc = f3d_params_check(p);
  if (rc) return rc;
  if (!ws || !flag_out || !U) {
    set_error("nan_check: NULL argument");
    return F3D_E_NULL;
  }
  if (NV < 1 || NV > 6) {
    set_error("nan_check: NV=%d out of range", NV);
    return F3D_E_ARG;
  }
  const long long n = (long long)f3d_cell_count(p) * NV;
  const int nb = (int)std::min<long long>(ws->nred, (n + 255) / 256);
  hipLaunchKernelGGL(k_nan_partial, dim3(nb), dim3(256), 0, (hipStream_t)s, U, n, ws->flag);
  hipLaunchKernelGGL(k_nan_final, dim3(1), dim3(256), 0, (hipStream_t)s, ws->flag, nb, (int*)flag_out);
  return check_launch("nan_check");
}

F3D_API int f3d_nan_check_sync(const f3d_params* p, f3d_workspace* ws, const double* U, int32_t NV,
                               f3d_stream_t s) {
  if (!ws) {
    set_error("nan_check_sync: workspace is NULL");
    return F3D_E_NULL;
  }
  // the flag lands in the last partial slot; the partials of this call use at most nred-1 slots
  int32_t* dflag = ws->flag + ws->nred - 1;
  const long long n = (long long)f3d_cell_count(p) * (NV > 0 ? NV : 1);
  int rc = f3d_params_check(p);
  if (rc) return rc;
  if (!U) {
    set_error("nan_check_sync: U is NULL");
    return F3D_E_NULL;
  }
  if (NV < 1 || NV > 6) {
    set_error("nan_check_sync: NV=%d out of range", NV);
    return F3D_E_ARG;
  }
  const int nb = (int)std::min<long long>(ws->nred - 1, (n + 255) / 256);
  hipLaunchKernelGGL(k_nan_partial, dim3(nb), dim3(256), 0, (hipStream_t)s, U, n, ws->flag);
  hipLaunchKernelGGL(k_nan_final, dim3(1), dim3(256), 0, (hipStream_t)s, ws->flag, nb, dflag);
  rc = check_launch("nan_check_sync");
  if (rc) return rc;
  int32_t h = 0;
  F3D_HIP(hipMemcpyAsync(&h, dflag, sizeof(int32_t), hipMemcpyDeviceToHost, (hipStream_t)s));
  F3D_HIP(hipStreamSynchronize((hipStream_t)s));
  if (h) {
    set_error("NaN or Inf detected in U (solver.jl:481-486)");
    return F3D_E_NAN;
  }
  return F3D_OK;
}

F3D_API int f3d_metrics_f64(const f3d_params* p, const double* x, const double* y, const double* z,
                            double* const* metrics, double* J, double* const* s_out, f3d_stream_t s) {
  int rc = f3d_params_check(p);
  if (rc) return rc;
  if (!x || !y || !z || !metrics || !J) {
    set_error("metrics_f64: NULL argument");
    return F3D_E_NULL;
  }
  MetricPtrs<double> m, so;
  for (int q = 0; q < 9; ++q) {
    if (!metrics[q]) {
      set_error("metrics_f64: metric %d is NULL", q);
      return F3D_E_NULL;
    }
    m.m[q] = metrics[q];
  }
  for (int q = 0; q < 9; ++q) so.m[q] = nullptr;
  if (s_out)
    for (int e = 0; e < 3; ++e) so.m[e] = s_out[e];
  KP kp = make_kp(p);
  hipLaunchKernelGGL(k_metrics, dim3((kp.Nxt + 63) / 64, (kp.Nyt + 3) / 4, kp.Nzt), dim3(64, 4, 1), 0,
                     (hipStream_t)s, kp, x, y, z, m, J, so);
  return check_launch("metrics_f64");
}
