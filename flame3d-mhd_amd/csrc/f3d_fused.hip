// Fused RHS / RK3 stage for gfx950.
//
// Per stage (solver.jl:416-444 with flowAdvance's finite-difference branch, solver.jl:133-188):
//   k_sensor_f    shock sensor phi (schemes.jl:1-24); clears the heavy-face counter
//   k_sweep_x     x faces: SW split of a 128-cell row segment into LDS, UP7 face fluxes, R = dF_x
//   k_sweep_march y and z faces: one lane per (i, line), marching along the direction; the split
//                 fluxes of each new cell are pushed into an 8-slot FIR of UP7 face sums, so no
//                 split flux ever goes to HBM; R += dF_y, dF_z
//   k_heavy       faces whose sensor maximum selects WENO7/WENO5/NND (appended to a compacted list
//                 by the sweeps): split fluxes of the 8-cell stencil recomputed, characteristic
//                 reconstruction (schemes.jl:756-1009), R += (F_scheme - F_UP7) on both cells
//   k_viscous     viscous fluxes (viscous.jl:1-134), reference layout, into the workspace
//   k_update      viscous divergence + conservative update + RK3 blend + c2Prim (div.jl:1-50,
//                 utils.jl:2-70)
// Splitting every face as UP7 first and correcting the rare non-UP7 faces from a compacted list
// keeps the sweep waves free of the characteristic branch's divergence and registers.
#include <cstdlib>
#include <cstring>

#include "f3d_fused.hpp"

namespace f3d {

bool use_fused() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("F3D_RHS_ENGINE");
    v = (e && std::strcmp(e, "chain") == 0) ? 0 : 1;
  }
  return v == 1;
}

template <typename T>
struct FArgs {
  const T* Q;
  MetricPtrs<T> m;
  const double* J;
  const T* phi;
  T* R;
  unsigned long long* heavy;
  unsigned int* hcount;
};

__device__ __forceinline__ long long ridx(const KP& kp, int i, int j, int k) {
  // interior RHS layout [5][Nzp][Nyp][Nxp], 0-based interior indices
  return (long long)i + (long long)kp.Nxp * ((long long)j + (long long)kp.Nyp * k);
}

// Wave-aggregated append of heavy faces.
__device__ __forceinline__ void append_heavy(unsigned long long* list, unsigned int* count, bool pred,
                                             unsigned long long entry) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0) return;
  const int lane = __lane_id();
  unsigned int base = 0;
  const int leader = __ffsll((long long)mask) - 1;
  if (lane == leader) base = atomicAdd(count, (unsigned int)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (pred) {
    const unsigned long long below = lane == 0 ? 0ULL : (mask & ((~0ULL) >> (64 - lane)));
    list[base + __popcll(below)] = entry;
  }
}

template <typename T>
__device__ __forceinline__ T up7c(const KP& kp, int m) {
  return (T)kp.lin[m];
}

// ------------------------------------------------------------------ sensor (+ counter reset)
template <typename T>
__global__ __launch_bounds__(256) void k_sensor_f(KP kp, T* __restrict__ phi, const T* __restrict__ Q,
                                                  unsigned int* __restrict__ hcount) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y, k = 1 + blockIdx.z;
  if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0) *hcount = 0;
  if (i > kp.Nxt - 2 || j > kp.Nyt - 2 || k > kp.Nzt - 2) return;
  const T* P = Q + 4 * kp.nc;
  const long long o = cidx(kp, i, j, k), sy = kp.Nxt, sz = (long long)kp.Nxt * kp.Nyt;
  phi[o] = sensor_cell<T>(P[o - 1], P[o], P[o + 1], P[o - sy], P[o + sy], P[o - sz], P[o + sz]);
}

// ------------------------------------------------------------------ x sweep (LDS row segments)
constexpr int XW = 128;  // lanes per row segment (cells incl. the 4+4 halo)

template <typename T>
__global__ __launch_bounds__(256) void k_sweep_x(KP kp, FArgs<T> a, int L) {
  __shared__ T sFp[5][2][XW];
  __shared__ T sFm[5][2][XW];
  __shared__ T sF[5][2][XW];
  __shared__ T sphi[2][XW];
  const int l = threadIdx.x, r = threadIdx.y;
  const int jj = blockIdx.y * 2 + r, kk = blockIdx.z;
  const bool rowok = jj < kp.Nyp;
  const int j = kp.NG + (rowok ? jj : kp.Nyp - 1), k = kp.NG + kk;
  const int f0 = blockIdx.x * L;
  const int ncell = min(L, kp.Nxp - f0);
  const int nface = ncell + 1;
  const int x = kp.NG + f0 - 4 + l;  // cell this lane splits
  const long long nc = kp.nc;
  if (l < nface + 7) {
    const long long o = cidx(kp, x, j, k);
    const T A1 = a.m.m[0][o], A2 = a.m.m[1][o], A3 = a.m.m[2][o];
    const T S = sqrt(A1 * A1 + A2 * A2 + A3 * A3);
    T fp[5], fm[5];
    sw_split<T>(kp, a.Q[o], a.Q[o + nc], a.Q[o + 2 * nc], a.Q[o + 3 * nc], a.Q[o + 5 * nc], A1, A2, A3, S, fp, fm);
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      sFp[n][r][l] = fp[n];
      sFm[n][r][l] = fm[n];
    }
    sphi[r][l] = a.phi[o];
  }
  __syncthreads();
  const bool fact = l >= 4 && l < 4 + nface;
  const int f = f0 + l - 4;
  int sch = F3D_SCHEME_UP7;
  if (fact) {
    T m = sphi[r][l - 3];
#pragma unroll
    for (int s = -2; s <= 2; ++s) m = jmax(m, sphi[r][l + s]);
    sch = scheme_of<T>(kp, m);
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      T acc = 0;
#pragma unroll
      for (int q = 0; q < 7; ++q) acc += up7c<T>(kp, q) * sFp[n][r][l - 4 + q];
#pragma unroll
      for (int q = 0; q < 7; ++q) acc += up7c<T>(kp, q) * sFm[n][r][l + 3 - q];
      sF[n][r][l] = acc;
    }
  }
  const bool owned = rowok && fact && (f < f0 + L || f == kp.Nxp);
  append_heavy(a.heavy, a.hcount, owned && sch != F3D_SCHEME_UP7,
               ((unsigned long long)cidx(kp, kp.NG + f, j, k) << 4) | ((unsigned long long)sch << 2) | 0ULL);
  __syncthreads();
  if (rowok && l >= 4 && l < 4 + ncell) {
    const long long ro = ridx(kp, f, jj, kk), rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
#pragma unroll
    for (int n = 0; n < 5; ++n) a.R[ro + n * rn] = sF[n][r][l] - sF[n][r][l + 1];
  }
}

// ------------------------------------------------------------------ y / z sweeps (register march)
// Lane = interior i (64 consecutive), threadIdx.y = transverse line; each block marches a segment
// of L cells along D.  Face f (between cells NG+f-1 and NG+f) collects
//   sum_q c_q Fp[NG+f-4+q] + sum_q c_q Fm[NG+f+3-q]
// in FIR slot (f - f0) & 7 as the cells stream in; it is complete once cell NG+f+3 has arrived.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_sweep_march(KP kp, FArgs<T> a, int L) {
  const int ii = blockIdx.x * 64 + threadIdx.x;
  const int tt = blockIdx.y * 4 + threadIdx.y;
  const int Nt2 = D == 1 ? kp.Nzp : kp.Nyp;  // extent of the transverse line index
  const int Np = D == 1 ? kp.Nyp : kp.Nzp;
  const bool act = ii < kp.Nxp && tt < Nt2;
  const int i = kp.NG + min(ii, kp.Nxp - 1), t = kp.NG + min(tt, Nt2 - 1);
  const int f0 = blockIdx.z * L;
  const int ncell = min(L, Np - f0);
  const int nstep = ncell + 8;
  const long long st = dstride(kp, D), nc = kp.nc;
  const long long obase = D == 1 ? cidx(kp, i, 0, t) : cidx(kp, i, t, 0);
  const T* A1p = a.m.m[3 * D];
  const T* A2p = a.m.m[3 * D + 1];
  const T* A3p = a.m.m[3 * D + 2];
  const long long rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
  T c[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) c[q] = up7c<T>(kp, q);
  T acc[8][5];
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int n = 0; n < 5; ++n) acc[s][n] = 0;
  T ring[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) ring[s] = 0;
  T Fprev[5] = {0, 0, 0, 0, 0};
  for (int g = 0; g < nstep; g += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int step = g + u;
      if (step < nstep) {
        const long long o = obase + (long long)(kp.NG + f0 - 4 + step) * st;
        const T A1 = A1p[o], A2 = A2p[o], A3 = A3p[o];
        const T S = sqrt(A1 * A1 + A2 * A2 + A3 * A3);
        T fp[5], fm[5];
        sw_split<T>(kp, a.Q[o], a.Q[o + nc], a.Q[o + 2 * nc], a.Q[o + 3 * nc], a.Q[o + 5 * nc], A1, A2, A3, S, fp,
                    fm);
        ring[u] = a.phi[o];
#pragma unroll
        for (int q = 0; q < 7; ++q)
#pragma unroll
          for (int n = 0; n < 5; ++n) {
            acc[(u - q) & 7][n] += c[q] * fp[n];
            acc[(u + 1 + q) & 7][n] += c[q] * fm[n];
          }
        const int fi = step - 7;  // face completed by this cell, relative to f0
        const int cs = (u + 1) & 7;
        if (fi >= 0) {
          T F[5];
#pragma unroll
          for (int n = 0; n < 5; ++n) {
            F[n] = acc[cs][n];
            acc[cs][n] = 0;
          }
          T m = ring[(u + 2) & 7];
#pragma unroll
          for (int s = 3; s <= 7; ++s) m = jmax(m, ring[(u + s) & 7]);
          const int sch = scheme_of<T>(kp, m);
          const int f = f0 + fi;
          const bool owned = act && (fi < L || f == Np);
          const long long oface = obase + (long long)(kp.NG + f) * st;
          append_heavy(a.heavy, a.hcount, owned && sch != F3D_SCHEME_UP7,
                       ((unsigned long long)oface << 4) | ((unsigned long long)sch << 2) | (unsigned long long)D);
          if (fi >= 1 && act) {
            const long long ro = D == 1 ? ridx(kp, ii, f - 1, tt) : ridx(kp, ii, tt, f - 1);
#pragma unroll
            for (int n = 0; n < 5; ++n) a.R[ro + n * rn] += Fprev[n] - F[n];
          }
#pragma unroll
          for (int n = 0; n < 5; ++n) Fprev[n] = F[n];
        } else {
#pragma unroll
          for (int n = 0; n < 5; ++n) acc[cs][n] = 0;
        }
      }
    }
  }
}

// ------------------------------------------------------------------ heavy faces
template <typename T>
__global__ __launch_bounds__(64) void k_heavy(KP kp, FArgs<T> a) {
  const unsigned int n = *a.hcount;
  const long long rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp, nc = kp.nc;
  for (unsigned int e = blockIdx.x * 64 + threadIdx.x; e < n; e += gridDim.x * 64) {
    const unsigned long long v = a.heavy[e];
    const int D = (int)(v & 3), sch = (int)((v >> 2) & 3);
    const long long o = (long long)(v >> 4);
    const long long st = dstride(kp, D);
    const T* A1p = a.m.m[3 * D];
    const T* A2p = a.m.m[3 * D + 1];
    const T* A3p = a.m.m[3 * D + 2];
    T fp[8][5], fm[8][5];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const long long oc = o + (s - 4) * st;
      const T A1 = A1p[oc], A2 = A2p[oc], A3 = A3p[oc];
      sw_split<T>(kp, a.Q[oc], a.Q[oc + nc], a.Q[oc + 2 * nc], a.Q[oc + 3 * nc], a.Q[oc + 5 * nc], A1, A2, A3,
                  sqrt(A1 * A1 + A2 * A2 + A3 * A3), fp[s], fm[s]);
    }
    auto FP = [&](int s, int q) { return fp[s + 4][q]; };
    auto FM = [&](int s, int q) { return fm[s + 4][q]; };
    const long long oL = o - st;
    const T qL[4] = {a.Q[oL + nc], a.Q[oL + 2 * nc], a.Q[oL + 3 * nc], a.Q[oL + 5 * nc]};
    const T qR[4] = {a.Q[o + nc], a.Q[o + 2 * nc], a.Q[o + 3 * nc], a.Q[o + 5 * nc]};
    const T AL[3] = {A1p[oL], A2p[oL], A3p[oL]};
    const T AR[3] = {A1p[o], A2p[o], A3p[o]};
    const T SL = sqrt(AL[0] * AL[0] + AL[1] * AL[1] + AL[2] * AL[2]);
    const T SR = sqrt(AR[0] * AR[0] + AR[1] * AR[1] + AR[2] * AR[2]);
    T Fs[5], Fu[5];
    face_flux<T>(kp, sch, FP, FM, qL, qR, AL, AR, SL, SR, Fs);
#pragma unroll
    for (int q = 0; q < 5; ++q) {  // the UP7 flux the sweep accumulated, same summation order
      T acc = 0;
#pragma unroll
      for (int s = 0; s < 7; ++s) acc += up7c<T>(kp, s) * fp[s][q];
#pragma unroll
      for (int s = 0; s < 7; ++s) acc += up7c<T>(kp, s) * fm[7 - s][q];
      Fu[q] = acc;
    }
    // decode the right cell's interior index along D
    const int ci = (int)(o % kp.Nxt), cj = (int)((o / kp.Nxt) % kp.Nyt), ck = (int)(o / ((long long)kp.Nxt * kp.Nyt));
    const int I = ci - kp.NG, Jj = cj - kp.NG, K = ck - kp.NG;
    const int f = D == 0 ? I : (D == 1 ? Jj : K);
    const int Np = D == 0 ? kp.Nxp : (D == 1 ? kp.Nyp : kp.Nzp);
    const long long rr = ridx(kp, I, Jj, K);
    const long long rl = rr - (D == 0 ? 1 : (D == 1 ? (long long)kp.Nxp : (long long)kp.Nxp * kp.Nyp));
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const T dF = Fs[q] - Fu[q];
      if (f <= Np - 1) atomicAdd(&a.R[rr + q * rn], dF);   // face is the left face of cell f
      if (f >= 1) atomicAdd(&a.R[rl + q * rn], -dF);       // and the right face of cell f-1
    }
  }
}

// ------------------------------------------------------------------ update
// MODE 0: RHS out (U layout, double) = J * (R + dFv).  MODE 1..3: RK3 stage KRK = MODE.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_update(KP kp, FArgs<T> a, const T* __restrict__ Fvx,
                                                const T* __restrict__ Fvy, const T* __restrict__ Fvz,
                                                double* __restrict__ U, double* __restrict__ Un, T* __restrict__ Q,
                                                double* __restrict__ Rout, double dt) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= kp.Nxp || j >= kp.Nyp) return;
  const int NG = kp.NG;
  const long long o = cidx(kp, i + NG, j + NG, k + NG), nc = kp.nc;
  const long long rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp, ro = ridx(kp, i, j, k);
  T Rt[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) Rt[n] = a.R[ro + n * rn];
  if (kp.viscous) {
    const T c1 = T(1) / T(12), c2 = T(-2) / T(3);
    const int Nv1 = kp.Nxp + NG, Nv2 = kp.Nyp + NG;
    const long long vn = (long long)Nv1 * Nv2 * (kp.Nzp + NG);
    const long long vo = (i + 2) + (long long)Nv1 * ((j + 2) + (long long)Nv2 * (k + 2));
    const long long vs[3] = {1, Nv1, (long long)Nv1 * Nv2};
    const T* Fv[3] = {Fvx, Fvy, Fvz};
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      T s = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T* f = Fv[d] + vo + n * vn;
        s += c1 * (f[-2 * vs[d]] - f[2 * vs[d]]) + c2 * (f[-vs[d]] - f[vs[d]]);
      }
      Rt[n + 1] += s;
    }
  }
  const double Jc = a.J[o];
  if (MODE == 0) {
#pragma unroll
    for (int n = 0; n < 5; ++n) Rout[o + n * nc] = (double)Rt[n] * Jc;
    return;
  }
  const double Jact = Jc * dt;
  double un[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const double u0 = U[o + n * nc];
    const double u1 = u0 + (double)Rt[n] * Jact;
    double r;
    if (MODE == 1) {
      Un[o + n * nc] = u0;
      r = u1;
    } else if (MODE == 2) {
      r = u1 * 0.25 + Un[o + n * nc] * 0.75;
    } else {
      r = u1 * (2.0 / 3.0) + Un[o + n * nc] * (1.0 / 3.0);
    }
    un[n] = r;
    U[o + n * nc] = r;
  }
  // c2Prim  utils.jl:2-30
  const double eps = 2.220446049250313e-16;
  const double rho = jmax(un[0], eps);
  const double ri = 1.0 / rho;
  const double u = un[1] * ri, v = un[2] * ri, w = un[3] * ri;
  const double ei = jmax(un[4] - 0.5 * rho * (u * u + v * v + w * w), eps);
  const double p = (kp.gamma - 1) * ei;
  Q[o] = (T)rho;
  Q[o + nc] = (T)u;
  Q[o + 2 * nc] = (T)v;
  Q[o + 3 * nc] = (T)w;
  Q[o + 4 * nc] = (T)p;
  Q[o + 5 * nc] = (T)(p / (rho * kp.Rg));
}

// ------------------------------------------------------------------ orchestration
static int seg_len(int Np, int max_len) {
  const int nseg = (Np + max_len - 1) / max_len;
  return (Np + nseg - 1) / nseg;
}

template <typename T>
static int fused_flux(const KP& kp, const WsView<T>& ws, const T* Q, const MetricPtrs<T>& m, const double* J,
                      T* phi, hipStream_t s) {
  FArgs<T> a{Q, m, J, phi, ws.R, ws.heavy, ws.hcount};
  {
    TimedScope ts("fused_sensor", s);
    hipLaunchKernelGGL(k_sensor_f<T>, dim3((kp.Nxt - 2 + 63) / 64, (kp.Nyt - 2 + 3) / 4, kp.Nzt - 2), dim3(64, 4, 1),
                       0, s, kp, phi, Q, ws.hcount);
  }
  {
    TimedScope ts("fused_sweep_x", s);
    const int L = seg_len(kp.Nxp, XW - 8);
    hipLaunchKernelGGL(k_sweep_x<T>, dim3((kp.Nxp + L - 1) / L, (kp.Nyp + 1) / 2, kp.Nzp), dim3(XW, 2, 1), 0, s, kp,
                       a, L);
  }
  {
    TimedScope ts("fused_sweep_y", s);
    const int L = seg_len(kp.Nyp, 32);
    hipLaunchKernelGGL((k_sweep_march<T, 1>), dim3((kp.Nxp + 63) / 64, (kp.Nzp + 3) / 4, (kp.Nyp + L - 1) / L),
                       dim3(64, 4, 1), 0, s, kp, a, L);
  }
  {
    TimedScope ts("fused_sweep_z", s);
    const int L = seg_len(kp.Nzp, 32);
    hipLaunchKernelGGL((k_sweep_march<T, 2>), dim3((kp.Nxp + 63) / 64, (kp.Nyp + 3) / 4, (kp.Nzp + L - 1) / L),
                       dim3(64, 4, 1), 0, s, kp, a, L);
  }
  {
    TimedScope ts("fused_heavy", s);
    hipLaunchKernelGGL(k_heavy<T>, dim3(2048), dim3(64), 0, s, kp, a);
  }
  int rc = check_launch("fused sweeps");
  if (rc) return rc;
  if (kp.viscous) rc = launch_viscous<T>(kp, ws.Fvx, ws.Fvy, ws.Fvz, Q, m, J, s);
  return rc;
}

template <typename T>
int fused_rhs(const KP& kp, const WsView<T>& ws, const T* Q, const MetricPtrs<T>& m, const double* J, T* phi,
              double* R, hipStream_t s) {
  int rc = fused_flux<T>(kp, ws, Q, m, J, phi, s);
  if (rc) return rc;
  FArgs<T> a{Q, m, J, phi, ws.R, ws.heavy, ws.hcount};
  TimedScope ts("fused_update", s);
  hipLaunchKernelGGL((k_update<T, 0>), dim3((kp.Nxp + 63) / 64, (kp.Nyp + 3) / 4, kp.Nzp), dim3(64, 4, 1), 0, s, kp,
                     a, ws.Fvx, ws.Fvy, ws.Fvz, nullptr, nullptr, nullptr, R, 0.0);
  return check_launch("fused rhs update");
}

template <typename T>
int fused_stage(const KP& kp, const WsView<T>& ws, int KRK, double* U, double* Un, T* Q, const MetricPtrs<T>& m,
                const double* J, double dt, hipStream_t s) {
  int rc = fused_flux<T>(kp, ws, Q, m, J, ws.phi, s);
  if (rc) return rc;
  FArgs<T> a{Q, m, J, ws.phi, ws.R, ws.heavy, ws.hcount};
  const dim3 g((kp.Nxp + 63) / 64, (kp.Nyp + 3) / 4, kp.Nzp), b(64, 4, 1);
  TimedScope ts("fused_update", s);
  if (KRK == 1)
    hipLaunchKernelGGL((k_update<T, 1>), g, b, 0, s, kp, a, ws.Fvx, ws.Fvy, ws.Fvz, U, Un, Q, nullptr, dt);
  else if (KRK == 2)
    hipLaunchKernelGGL((k_update<T, 2>), g, b, 0, s, kp, a, ws.Fvx, ws.Fvy, ws.Fvz, U, Un, Q, nullptr, dt);
  else
    hipLaunchKernelGGL((k_update<T, 3>), g, b, 0, s, kp, a, ws.Fvx, ws.Fvy, ws.Fvz, U, Un, Q, nullptr, dt);
  return check_launch("fused stage update");
}

template int fused_rhs<double>(const KP&, const WsView<double>&, const double*, const MetricPtrs<double>&,
                               const double*, double*, double*, hipStream_t);
template int fused_rhs<float>(const KP&, const WsView<float>&, const float*, const MetricPtrs<float>&, const double*,
                              float*, double*, hipStream_t);
template int fused_stage<double>(const KP&, const WsView<double>&, int, double*, double*, double*,
                                 const MetricPtrs<double>&, const double*, double, hipStream_t);
template int fused_stage<float>(const KP&, const WsView<float>&, int, double*, double*, float*,
                                const MetricPtrs<float>&, const double*, double, hipStream_t);
}  // namespace f3d
