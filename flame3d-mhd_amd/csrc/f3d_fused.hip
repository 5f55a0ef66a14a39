// Fused RHS / RK3 stage for gfx950.
//
// Per stage (solver.jl:416-444 with flowAdvance's finite-difference branch, solver.jl:133-188):
//   k_sensor_march shock sensor phi (schemes.jl:1-24), marching k; clears the heavy-face counter
//   k_sweep_x     x faces: SW split of a 128-cell row segment into LDS, UP7 face fluxes, R = dF_x
//   k_sweep_march y and z faces: one lane per (i, line), marching along the direction; the split
//                 fluxes of each new cell are pushed into an 8-face shift register of UP7 sums, so
//                 no split flux ever goes to HBM; R += dF_y, dF_z (the R row loaded a step ahead)
//   k_heavy       faces whose sensor maximum selects WENO7/WENO5/NND (appended to a compacted list
//                 by the sweeps): split fluxes of the 8-cell stencil recomputed, characteristic
//                 reconstruction (schemes.jl:756-1009), F_scheme - F_UP7 kept per face; k_heavy_apply
//                 adds it to both cells (R += on the right cell, -= on the left, fixed order per cell)
//   k_visc_update_qt viscous fluxes (viscous.jl:1-134), their divergence (div.jl:14-48), the
//                 conservative update, RK3 blend and c2Prim (utils.jl:2-70) in one k-march; the
//                 viscous fluxes stay in LDS and registers
//   k_viscous_march + k_update  the same as two kernels (viscous fluxes through the workspace), for
//                 the LES models and the inviscid case
// Splitting every face as UP7 first and correcting the rare non-UP7 faces from a compacted list
// keeps the sweep waves free of the characteristic branch's divergence and registers.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "f3d_fused.hpp"

namespace f3d {

int xcd_remap() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("F3D_XCD_REMAP");
    v = (e && std::strcmp(e, "0") == 0) ? 0 : 1;
  }
  return v;
}

// 1 unless F3D_VISC_FUSED=0: the viscous divergence computed inside the update march (k_visc_update_qt)
// instead of k_viscous_march + k_update (A/B measurement).
static int visc_fused() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("F3D_VISC_FUSED");
    v = (e && std::strcmp(e, "0") == 0) ? 0 : 1;
  }
  return v;
}

__global__ __launch_bounds__(256) void k_clear_hidx(unsigned long long* __restrict__ a, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    a[t] = 0ULL;
}

void clear_hidx(unsigned long long* idx, long long n, hipStream_t s) {
  const long long nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_clear_hidx, dim3((unsigned int)(nb < 4096 ? (nb ? nb : 1) : 4096)), dim3(256), 0, s, idx, n);
}

bool fused_needs_fv(const KP& kp) { return kp.viscous && !(visc_fused() && kp.les == F3D_LES_NONE); }

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
  T* R;
  unsigned long long* heavy;
  unsigned int* hcount;
  const double* ldt = nullptr;  // per-cell dt of local time stepping (div_LTS), else the scalar dt
  double* Uo = nullptr;         // conserved state written by the update (nullptr: U itself, in place)
  const unsigned int* pflag = nullptr;  // sensor band flag of this stage (heavy_possible), and its generation
  unsigned int pgen = 0;
  const unsigned char* cls = nullptr;   // per-cell sensor class (scheme_of of the cell's phi), read by the sweeps
};

__device__ __forceinline__ long long ridx(const KP& kp, int i, int j, int k) {
  // interior RHS layout [5][Nzp][Nyp][Nxp], 0-based interior indices
  return (long long)i + (long long)kp.Nxp * ((long long)j + (long long)kp.Nyp * k);
}

template <typename T>
__device__ __forceinline__ T up7c(const KP& kp, int m) {
  return Coef<T>::lin(kp, m);
}

// ------------------------------------------------------------------ sensor (+ counter reset)
// A wave with a sensor value not below phi1 (NaN included) records the stage's generation in the flag
// (every such wave stores the same value).
template <typename T>
__device__ __forceinline__ void mark_band(const KP& kp, bool out, unsigned int* pflag, unsigned int gen) {
  if (pflag && __ballot(out) != 0ULL && __lane_id() == __ffsll((unsigned long long)__ballot(1)) - 1) *pflag = gen;
}
template <typename T>
__device__ __forceinline__ bool out_of_band(const KP& kp, T ph) {
  return !(ph < (T)kp.phi1);
}

template <typename T>
__global__ __launch_bounds__(256) void k_sensor_f(KP kp, T* __restrict__ phi, unsigned char* __restrict__ cls,
                                                  const T* __restrict__ Q, unsigned int* pflag, unsigned int gen) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y, k = 1 + blockIdx.z;
  if (i > kp.Nxt - 2 || j > kp.Nyt - 2 || k > kp.Nzt - 2) return;
  const T* P = Q + 4 * kp.nc;
  const long long o = cidx(kp, i, j, k), sy = kp.Nxt, sz = (long long)kp.Nxt * kp.Nyt;
  const T ph = sensor_cell<T>(P[o - 1], P[o], P[o + 1], P[o - sy], P[o + sy], P[o - sz], P[o + sz]);
  if (phi) phi[o] = ph;
  cls[o] = (unsigned char)scheme_of<T>(kp, ph);
  mark_band<T>(kp, out_of_band(kp, ph), pflag, gen);
}

// Same sensor, one lane per (i, j) column of [1, N_tot-2] marching a segment of KS planes: p of the
// column at k-1, k, k+1 stays in registers, so each cell costs one new plane load plus the four xi/eta
// neighbours (L1/L2) instead of six neighbour loads.  Identical arithmetic (sensor_cell).
template <typename T>
__global__ __launch_bounds__(256) void k_sensor_march(KP kp, T* __restrict__ phi, unsigned char* __restrict__ cls,
                                                      const T* __restrict__ Q, int KS, unsigned int* pflag,
                                                      unsigned int gen) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k0 = 1 + blockIdx.z * KS;
  if (i > kp.Nxt - 2 || j > kp.Nyt - 2) return;
  const int k1 = min(k0 + KS, kp.Nzt - 1);  // planes k0 .. k1-1
  const T* P = Q + 4 * kp.nc;
  const long long sy = kp.Nxt, sz = (long long)kp.Nxt * kp.Nyt;
  long long o = cidx(kp, i, j, k0);
  T pm = P[o - sz], pc = P[o];
  bool out = false;
#pragma unroll 1
  for (int k = k0; k < k1; ++k, o += sz) {
    const T pp = P[o + sz];
    const T ph = sensor_cell<T>(P[o - 1], pc, P[o + 1], P[o - sy], P[o + sy], pm, pp);
    if (phi) phi[o] = ph;
    cls[o] = (unsigned char)scheme_of<T>(kp, ph);
    out |= out_of_band(kp, ph);
    pm = pc;
    pc = pp;
  }
  mark_band<T>(kp, out, pflag, gen);
}

// ------------------------------------------------------------------ x sweep (LDS row segments)
constexpr int XW = 128;  // lanes per row segment (cells incl. the 4+4 halo)

template <typename T, int SPLIT>
__global__ __launch_bounds__(256) void k_sweep_x(KP kp, FArgs<T> a, int L, HSeg hs) {
  __shared__ T sFp[5][2][XW];
  __shared__ T sFm[5][2][XW];
  __shared__ T sF[5][2][XW];
  __shared__ unsigned char scl[2][XW];
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
    split_cell<T, SPLIT>(kp, a.Q[o], a.Q[o + nc], a.Q[o + 2 * nc], a.Q[o + 3 * nc], a.Q[o + split_qvar<SPLIT>() * nc],
                         A1, A2, A3, S, fp, fm);
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      sFp[n][r][l] = fp[n];
      sFm[n][r][l] = fm[n];
    }
    scl[r][l] = a.cls[o];
  }
  __syncthreads();
  const bool fact = l >= 4 && l < 4 + nface;
  const int f = f0 + l - 4;
  int sch = F3D_SCHEME_UP7;
  if (fact) {
    // the face's scheme: the largest sensor class of its six cells (the class of the cells' largest phi)
    int m = scl[r][l - 3];
#pragma unroll
    for (int s = -2; s <= 2; ++s) m = max(m, (int)scl[r][l + s]);
    sch = m;
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
  {  // the wave's run of the heavy list (2 waves per row, 4 per block; at most one face per lane)
    const unsigned int w = launch_wave(4, r * 2 + (l >> 6));
    unsigned int cnt = 0;
    run_append(a.heavy, hs.base + (unsigned long long)w * hs.cap, cnt, owned && sch != F3D_SCHEME_UP7,
               ((unsigned long long)cidx(kp, kp.NG + f, j, k) << 4) | ((unsigned long long)sch << 2) | 0ULL);
    if ((l & 63) == 0) a.hcount[hs.first + w] = cnt;
  }
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
// in a shift register acc[0..7] (acc[m] holds face f0+step-7+m); acc[0] is complete once cell
// NG+f+3 has streamed in.  The next cell's loads are issued before the current cell is split.
template <typename T, int D, int SPLIT>
__global__ __launch_bounds__(256) void k_sweep_march(KP kp, FArgs<T> a, int L, HSeg hs) {
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
  const long long obase = (D == 1 ? cidx(kp, i, 0, t) : cidx(kp, i, t, 0)) + (long long)(kp.NG + f0 - 4) * st;
  const T* __restrict__ A1p = a.m.m[3 * D];
  const T* __restrict__ A2p = a.m.m[3 * D + 1];
  const T* __restrict__ A3p = a.m.m[3 * D + 2];
  const long long rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
  T acc[8][5];
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int n = 0; n < 5; ++n) acc[s][n] = 0;
  int ring[6];  // sensor classes of the last six cells
#pragma unroll
  for (int s = 0; s < 6; ++s) ring[s] = 0;
  T Fprev[5] = {0, 0, 0, 0, 0};
  T Rn[5] = {0, 0, 0, 0, 0};  // RHS of the row the next step closes, loaded a step ahead
  // the wave's run of the heavy list: at most one face per lane and completed face (<= 64 (L + 1))
  const unsigned int hw = launch_wave(4, threadIdx.y);
  const unsigned long long hrun = hs.base + (unsigned long long)hw * hs.cap;
  unsigned int hcnt = 0;
  // no face can leave UP7 when no sensor value of the stage left the band: no sensor reads, no masks
  const bool hp = heavy_possible(a.pflag, a.pgen);
  // prefetched inputs of the next cell
  T nq0, nq1, nq2, nq3, nq5, na1, na2, na3;
  int ncl;
  {
    const long long o = obase;
    nq0 = a.Q[o]; nq1 = a.Q[o + nc]; nq2 = a.Q[o + 2 * nc]; nq3 = a.Q[o + 3 * nc];
    nq5 = a.Q[o + split_qvar<SPLIT>() * nc];
    na1 = A1p[o]; na2 = A2p[o]; na3 = A3p[o]; ncl = hp ? (int)a.cls[o] : 0;
  }
#pragma unroll 1
  for (int step = 0; step < nstep; ++step) {
    const T q0 = nq0, q1 = nq1, q2 = nq2, q3 = nq3, q5 = nq5, A1 = na1, A2 = na2, A3 = na3;
    const int cl = ncl;
    if (step + 1 < nstep) {
      const long long o = obase + (long long)(step + 1) * st;
      nq0 = a.Q[o]; nq1 = a.Q[o + nc]; nq2 = a.Q[o + 2 * nc]; nq3 = a.Q[o + 3 * nc];
      nq5 = a.Q[o + split_qvar<SPLIT>() * nc];
      na1 = A1p[o]; na2 = A2p[o]; na3 = A3p[o];
      if (hp) ncl = a.cls[o];
    }
    T Rc[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) Rc[n] = Rn[n];
    if (step >= 7 && step + 1 < nstep && act) {  // row f0 + step - 7, closed at step + 1
      const long long ro = D == 1 ? ridx(kp, ii, f0 + step - 7, tt) : ridx(kp, ii, tt, f0 + step - 7);
#pragma unroll
      for (int n = 0; n < 5; ++n) Rn[n] = a.R[ro + n * rn];
    }
    T fp[5], fm[5];
    split_cell<T, SPLIT>(kp, q0, q1, q2, q3, q5, A1, A2, A3, sqrt(A1 * A1 + A2 * A2 + A3 * A3), fp, fm);
#pragma unroll
    for (int q = 0; q < 7; ++q)
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        acc[7 - q][n] += up7c<T>(kp, q) * fp[n];
        acc[q][n] += up7c<T>(kp, q) * fm[n];
      }
    if (step >= 7) {  // face f0 + step - 7 complete
      const int fi = step - 7;
      const int f = f0 + fi;
      if (hp) {
        int sch = ring[0];  // the largest class of the face's six cells
#pragma unroll
        for (int s = 1; s < 6; ++s) sch = max(sch, ring[s]);
        run_append(a.heavy, hrun, hcnt, act && (fi < L || f == Np) && sch != F3D_SCHEME_UP7,
                   ((unsigned long long)(obase + (long long)(fi + 4) * st) << 4) | ((unsigned long long)sch << 2) |
                       (unsigned long long)D);
      }
      if (fi >= 1 && act) {
        const long long ro = D == 1 ? ridx(kp, ii, f - 1, tt) : ridx(kp, ii, tt, f - 1);
#pragma unroll
        for (int n = 0; n < 5; ++n) a.R[ro + n * rn] = Rc[n] + (Fprev[n] - acc[0][n]);
      }
#pragma unroll
      for (int n = 0; n < 5; ++n) Fprev[n] = acc[0][n];
    }
#pragma unroll
    for (int s = 0; s < 7; ++s)
#pragma unroll
      for (int n = 0; n < 5; ++n) acc[s][n] = acc[s + 1][n];
#pragma unroll
    for (int n = 0; n < 5; ++n) acc[7][n] = 0;
#pragma unroll
    for (int s = 0; s < 5; ++s) ring[s] = ring[s + 1];
    ring[5] = cl;
  }
  if (threadIdx.x == 0) a.hcount[hs.first + hw] = hcnt;
}

// ------------------------------------------------------------------ x and y faces in one march
// Lanes = a window of XYW cells along x (4 halo cells each side, XYW - 8 interior columns), threadIdx.y =
// z line; the block marches a segment of L rows along y.  Per step the new row's cells are split along x
// into LDS and the UP7 x-face fluxes formed (k_sweep_x's sums, same order), kept for four steps in an LDS
// ring; the y part is k_sweep_march's shift register.  When a y face closes row r,
//   R[r] = (F_x,left - F_x,right) + (F_y,prev - F_y,new),
// the same sums and roundings as the x sweep's store followed by the y march's update (the compiler
// contracts some products into FMAs differently once both splits of a cell share a kernel, so the two
// forms agree to rounding), while R is written once instead of written, read and written again, and Q and
// phi are read once.  The x faces go to one run of the heavy list per wave, the y faces to another.
constexpr int XYW = 64;        // lanes of a row window (one wave)
constexpr int XYI = XYW - 8;   // interior columns per window
constexpr int XYL = 2;         // z lines (waves) per block

template <typename T, int SPLIT>
__global__ __launch_bounds__(XYW * XYL) void k_sweep_xy(KP kp, FArgs<T> a, int L, HSeg hx, HSeg hy) {
  __shared__ T sFp[5][XYL][XYW];
  __shared__ T sFm[5][XYL][XYW];
  __shared__ unsigned char scl[XYL][XYW];
  __shared__ T sF[4][5][XYL][XYW];  // x-face fluxes of the last four rows (slot = step & 3)
  const int l = threadIdx.x, r = threadIdx.y;
  const int fx0 = blockIdx.x * XYI;
  const int ncx = min(XYI, kp.Nxp - fx0);  // interior columns of the window
  const int nfx = ncx + 1;                 // x faces formed, fx0 .. fx0 + ncx (lanes 4 .. 4 + ncx)
  const int tt = blockIdx.y * XYL + r;
  const bool lineok = tt < kp.Nzp;
  const int k = kp.NG + min(tt, kp.Nzp - 1);
  const int x = min(kp.NG + fx0 - 4 + l, kp.Nxt - 1);  // the lane's cell (lanes past the stencil clamped)
  const bool yin = l >= 4 && l < 4 + ncx;               // interior column: its y line is updated
  const bool act = lineok && yin;
  const int ii = fx0 + l - 4;
  const int f0 = blockIdx.z * L;
  const int ncell = min(L, kp.Nyp - f0);
  const int nstep = ncell + 8;
  const long long st = kp.Nxt, nc = kp.nc;
  const long long obase = cidx(kp, x, kp.NG + f0 - 4, k);
  const T* __restrict__ A1p = a.m.m[3];
  const T* __restrict__ A2p = a.m.m[4];
  const T* __restrict__ A3p = a.m.m[5];
  const T* __restrict__ X1p = a.m.m[0];
  const T* __restrict__ X2p = a.m.m[1];
  const T* __restrict__ X3p = a.m.m[2];
  const long long rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
  T acc[8][5];
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int n = 0; n < 5; ++n) acc[s][n] = 0;
  int ring[6];  // sensor classes of the last six cells
#pragma unroll
  for (int s = 0; s < 6; ++s) ring[s] = 0;
  T Fprev[5] = {0, 0, 0, 0, 0};
  const unsigned int hw = launch_wave(XYL, r);
  const unsigned long long xrun = hx.base + (unsigned long long)hw * hx.cap;
  const unsigned long long yrun = hy.base + (unsigned long long)hw * hy.cap;
  unsigned int xcnt = 0, ycnt = 0;
  // no face can leave UP7 when no sensor value of the stage left the band: no sensor reads, no masks
  const bool hp = heavy_possible(a.pflag, a.pgen);
  // prefetched inputs of the next cell (x metrics only on the steps that form x faces: rows f0 ..)
  T nq0, nq1, nq2, nq3, nq5, na1, na2, na3, nb1 = 0, nb2 = 0, nb3 = 0;
  int ncl;
  {
    const long long o = obase;
    nq0 = a.Q[o]; nq1 = a.Q[o + nc]; nq2 = a.Q[o + 2 * nc]; nq3 = a.Q[o + 3 * nc];
    nq5 = a.Q[o + split_qvar<SPLIT>() * nc];
    na1 = A1p[o]; na2 = A2p[o]; na3 = A3p[o]; ncl = hp ? (int)a.cls[o] : 0;
  }
#pragma unroll 1
  for (int step = 0; step < nstep; ++step) {
    const T q0 = nq0, q1 = nq1, q2 = nq2, q3 = nq3, q5 = nq5, A1 = na1, A2 = na2, A3 = na3;
    const int cl = ncl;
    const T B1 = nb1, B2 = nb2, B3 = nb3;
    const bool xstep = step >= 4 && step < 4 + ncell;  // this step's row is interior: form its x faces
    if (step + 1 < nstep) {
      const long long o = obase + (long long)(step + 1) * st;
      nq0 = a.Q[o]; nq1 = a.Q[o + nc]; nq2 = a.Q[o + 2 * nc]; nq3 = a.Q[o + 3 * nc];
      nq5 = a.Q[o + split_qvar<SPLIT>() * nc];
      na1 = A1p[o]; na2 = A2p[o]; na3 = A3p[o];
      if (hp) ncl = a.cls[o];
      if (step + 1 >= 4 && step + 1 < 4 + ncell) {
        nb1 = X1p[o]; nb2 = X2p[o]; nb3 = X3p[o];
      }
    }
    if (xstep) {
      T fp[5], fm[5];
      split_cell<T, SPLIT>(kp, q0, q1, q2, q3, q5, B1, B2, B3, sqrt(B1 * B1 + B2 * B2 + B3 * B3), fp, fm);
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        sFp[n][r][l] = fp[n];
        sFm[n][r][l] = fm[n];
      }
      scl[r][l] = (unsigned char)cl;
    }
    {
      T fp[5], fm[5];
      split_cell<T, SPLIT>(kp, q0, q1, q2, q3, q5, A1, A2, A3, sqrt(A1 * A1 + A2 * A2 + A3 * A3), fp, fm);
#pragma unroll
      for (int q = 0; q < 7; ++q)
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          acc[7 - q][n] += up7c<T>(kp, q) * fp[n];
          acc[q][n] += up7c<T>(kp, q) * fm[n];
        }
    }
    if (step >= 7) {  // y face f0 + step - 7 complete
      const int fi = step - 7;
      const int f = f0 + fi;
      if (hp) {
        int sch = ring[0];  // the largest class of the face's six cells
#pragma unroll
        for (int s = 1; s < 6; ++s) sch = max(sch, ring[s]);
        run_append(a.heavy, yrun, ycnt, act && (fi < L || f == kp.Nyp) && sch != F3D_SCHEME_UP7,
                   ((unsigned long long)(obase + (long long)(fi + 4) * st) << 4) | ((unsigned long long)sch << 2) | 1ULL);
      }
      if (fi >= 1 && act) {  // row f - 1: its x faces were formed four steps ago (slot step & 3)
        const long long ro = ridx(kp, ii, f - 1, tt);
        const int sl = step & 3;
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          const T Rc = sF[sl][n][r][l] - sF[sl][n][r][l + 1];
          a.R[ro + n * rn] = Rc + (Fprev[n] - acc[0][n]);
        }
      }
#pragma unroll
      for (int n = 0; n < 5; ++n) Fprev[n] = acc[0][n];
    }
#pragma unroll
    for (int s = 0; s < 7; ++s)
#pragma unroll
      for (int n = 0; n < 5; ++n) acc[s][n] = acc[s + 1][n];
#pragma unroll
    for (int n = 0; n < 5; ++n) acc[7][n] = 0;
#pragma unroll
    for (int s = 0; s < 5; ++s) ring[s] = ring[s + 1];
    ring[5] = cl;
    __syncthreads();  // split fluxes of this row written; the ring slot read above is free
    if (xstep) {
      const bool fact = l >= 4 && l < 4 + nfx;
      const int f = fx0 + l - 4;
      const int sl = step & 3;
      int sch = F3D_SCHEME_UP7;
      if (fact) {
        if (hp) {
          sch = scl[r][l - 3];
#pragma unroll
          for (int s = -2; s <= 2; ++s) sch = max(sch, (int)scl[r][l + s]);
        }
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          T acx = 0;
#pragma unroll
          for (int q = 0; q < 7; ++q) acx += up7c<T>(kp, q) * sFp[n][r][l - 4 + q];
#pragma unroll
          for (int q = 0; q < 7; ++q) acx += up7c<T>(kp, q) * sFm[n][r][l + 3 - q];
          sF[sl][n][r][l] = acx;
        }
      }
      const int j = kp.NG + f0 + step - 4;
      if (hp)
        run_append(a.heavy, xrun, xcnt, lineok && fact && (f < fx0 + XYI || f == kp.Nxp) && sch != F3D_SCHEME_UP7,
                   ((unsigned long long)cidx(kp, kp.NG + f, j, k) << 4) | ((unsigned long long)sch << 2) | 0ULL);
    }
    __syncthreads();  // the split fluxes are read before the next row overwrites them
  }
  if (l == 0) {
    a.hcount[hx.first + hw] = xcnt;
    a.hcount[hy.first + hw] = ycnt;
  }
}

// ------------------------------------------------------------------ heavy faces
// Every listed face (all three directions, one launch): its characteristic WENO7/WENO5/NND flux
// (schemes.jl:756-1009) minus the UP7 flux the sweep accumulated, kept at the face's position and marked in
// its two cells (heavy_keep); k_heavy_apply then adds each cell's corrections in a fixed order, so no
// floating-point atomics: the corrections are summed the same way on every evaluation.
template <typename T, int SPLIT, bool CHAR>
__global__ __launch_bounds__(64) void k_heavy(KP kp, FArgs<T> a, HList hl, T* __restrict__ hcorr, HMark hm) {
  if (!heavy_possible(a.pflag, a.pgen)) return;  // nothing listed this stage
  const long long nc = kp.nc;
  for_each_heavy(hl, a.hcount, a.heavy, [&](const unsigned long long v) {
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
      split_cell<T, SPLIT>(kp, a.Q[oc], a.Q[oc + nc], a.Q[oc + 2 * nc], a.Q[oc + 3 * nc],
                           a.Q[oc + split_qvar<SPLIT>() * nc], A1, A2, A3,
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
    face_flux<T, CHAR>(kp, sch, FP, FM, qL, qR, AL, AR, SL, SR, Fs);
#pragma unroll
    for (int q = 0; q < 5; ++q) {  // the UP7 flux the sweep accumulated, same summation order
      T acc = 0;
#pragma unroll
      for (int s = 0; s < 7; ++s) acc += up7c<T>(kp, s) * fp[s][q];
#pragma unroll
      for (int s = 0; s < 7; ++s) acc += up7c<T>(kp, s) * fm[7 - s][q];
      Fu[q] = acc;
    }
    // the face is the left face of its right cell (interior index I, Jj, K) and the right face of the cell
    // before it along D
    const int ci = (int)(o % kp.Nxt), cj = (int)((o / kp.Nxt) % kp.Nyt), ck = (int)(o / ((long long)kp.Nxt * kp.Nyt));
    const int I = ci - kp.NG, Jj = cj - kp.NG, K = ck - kp.NG;
    const int f = D == 0 ? I : (D == 1 ? Jj : K);
    const int Np = D == 0 ? kp.Nxp : (D == 1 ? kp.Nyp : kp.Nzp);
    T dF[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) dF[q] = Fs[q] - Fu[q];
    heavy_keep(kp, hm, hcorr, D, I, Jj, K, f, Np, dF);
  });
}

// The value `v` of the other lane of this lane's pair (lanes 2p, 2p + 1): a DPP quad permutation (1, 0, 3, 2),
// no LDS.  Both lanes of a pair must be active.
__device__ __forceinline__ float pair_swap(float v) {
  const int b = __float_as_int(v);
  return __int_as_float(__builtin_amdgcn_update_dpp(b, b, 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double pair_swap(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = (int)(unsigned int)(unsigned long long)b, hi = (int)(b >> 32);
  const int slo = __builtin_amdgcn_update_dpp(lo, lo, 0xB1, 0xF, 0xF, false);
  const int shi = __builtin_amdgcn_update_dpp(hi, hi, 0xB1, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned int)shi << 32) | (unsigned int)slo));
}

// k_heavy with a pair of lanes per listed face: the even lane splits stencil cells 0..3 and reconstructs
// the F+ half (cells 0..6), the odd lane splits cells 4..7 and reconstructs the F- half (cells 7..1); the
// split fluxes each half needs from the other lane's cells (three cells, five components), the frame's
// two cells and the two halves' results cross with DPP moves.  The same loads and splits as one lane per
// face (each cell once), half the registers per lane (two or more waves per SIMD instead of one in FP64),
// the same arithmetic per field (F+ part + F- part); the UP7 flux subtracted is the F+ sum plus the F- sum.
template <typename T, int SPLIT, bool CHAR>
__global__ __launch_bounds__(64) void k_heavy_pair(KP kp, FArgs<T> a, HList hl, T* __restrict__ hcorr, HMark hm) {
  if (!heavy_possible(a.pflag, a.pgen)) return;  // nothing listed this stage
  const long long nc = kp.nc;
  const bool odd = (threadIdx.x & 1) != 0;
  for_each_heavy_pairs(hl, a.hcount, a.heavy, [&](const unsigned long long v, unsigned long long) {
    const int D = (int)(v & 3), sch = (int)((v >> 2) & 3);
    const long long o = (long long)(v >> 4);
    const long long st = dstride(kp, D);
    const T* A1p = a.m.m[3 * D];
    const T* A2p = a.m.m[3 * D + 1];
    const T* A3p = a.m.m[3 * D + 2];
    // this lane's cells (even: 0..3, odd: 4..7) split; its frame cell (even: 3 = the face's left cell, odd: 4)
    T fp[4][5], fm[4][5], cq[4], ca[3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long long oc = o + ((odd ? 4 : 0) + q - 4) * st;
      const T A1 = A1p[oc], A2 = A2p[oc], A3 = A3p[oc];
      const T r = a.Q[oc], u = a.Q[oc + nc], w2 = a.Q[oc + 2 * nc], w3 = a.Q[oc + 3 * nc];
      split_cell<T, SPLIT>(kp, r, u, w2, w3, a.Q[oc + split_qvar<SPLIT>() * nc], A1, A2, A3,
                           sqrt(A1 * A1 + A2 * A2 + A3 * A3), fp[q], fm[q]);
      if (q == (odd ? 0 : 3)) {
        cq[0] = u; cq[1] = w2; cq[2] = w3; cq[3] = a.Q[oc + 5 * nc];
        ca[0] = A1; ca[1] = A2; ca[2] = A3;
      }
    }
    T qL[4], qR[4], AL[3], AR[3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const T x = pair_swap(cq[q]);
      qL[q] = odd ? x : cq[q];
      qR[q] = odd ? cq[q] : x;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T x = pair_swap(ca[q]);
      AL[q] = odd ? x : ca[q];
      AR[q] = odd ? ca[q] : x;
    }
    const T SL = sqrt(AL[0] * AL[0] + AL[1] * AL[1] + AL[2] * AL[2]);
    const T SR = sqrt(AR[0] * AR[0] + AR[1] * AR[1] + AR[2] * AR[2]);
    // V[s]: even lane F+ of cell s, odd lane F- of cell 7 - s (s = 0..6); three of them come from the
    // other lane (even: F+ of cells 4..6, odd: F- of cells 3..1)
    T V[7][5];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        const T x = pair_swap(odd ? fp[q][n] : fm[q + 1][n]);
        if (odd) V[6 - q][n] = x; else V[4 + q][n] = x;
      }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int n = 0; n < 5; ++n) V[q][n] = odd ? fm[3 - q][n] : fp[q][n];
    T Fu[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) {  // the UP7 flux: F+ sum (even lane) + F- sum (odd lane)
      T acc = 0;
#pragma unroll
      for (int q = 0; q < 7; ++q) acc += up7c<T>(kp, q) * V[q][n];
      const T other = pair_swap(acc);
      Fu[n] = odd ? other + acc : acc + other;
    }
    T Fs[5];
    if (sch == F3D_SCHEME_NND2) {  // schemes.jl:1000-1009, component-wise in both forms
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        const T half = odd ? V[3][n] - T(0.5) * minmod(V[2][n] - V[3][n], V[3][n] - V[4][n])
                           : V[3][n] + T(0.5) * minmod(V[4][n] - V[3][n], V[3][n] - V[2][n]);
        const T other = pair_swap(half);
        Fs[n] = odd ? other + half : half + other;
      }
    } else {
      const T s0 = T(2) / (SL + SR);
      T flux[5];
      const CharFrame<T> fr(kp, qL, qR, AL, AR, SL, SR);
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        T L[5] = {0, 0, 0, 0, 0};
        if (CHAR) fr.lrow(n, L);
        auto proj = [&](int q) {
          return CHAR ? V[q][0] * L[0] + V[q][1] * L[1] + V[q][2] * L[2] + V[q][3] * L[3] + V[q][4] * L[4] : V[q][n];
        };
        T half;
        if (sch == F3D_SCHEME_WENO7) {
          T val[7];
#pragma unroll
          for (int q = 0; q < 7; ++q) val[q] = proj(q);
          half = weno7(val, s0, (T)kp.eps7);
        } else {
          T val[5];
#pragma unroll
          for (int q = 0; q < 5; ++q) val[q] = proj(q + 1);
          half = weno5(val, s0, (T)kp.eps5);
        }
        const T other = pair_swap(half);
        const T sum = odd ? other + half : half + other;
        flux[n] = sch == F3D_SCHEME_WENO7 ? sum * (T(1) / T(12)) : sum * (T(1) / T(6));
      }
      if (CHAR) {
        fr.project_back(flux, Fs);
      } else {
#pragma unroll
        for (int n = 0; n < 5; ++n) Fs[n] = flux[n];
      }
    }
    if (!odd) {
      const int ci = (int)(o % kp.Nxt), cj = (int)((o / kp.Nxt) % kp.Nyt), ck = (int)(o / ((long long)kp.Nxt * kp.Nyt));
      const int I = ci - kp.NG, Jj = cj - kp.NG, K = ck - kp.NG;
      const int f = D == 0 ? I : (D == 1 ? Jj : K);
      const int Np = D == 0 ? kp.Nxp : (D == 1 ? kp.Nyp : kp.Nzp);
      T dF[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) dF[q] = Fs[q] - Fu[q];
      heavy_keep(kp, hm, hcorr, D, I, Jj, K, f, Np, dF);
    }
  });
}

// ------------------------------------------------------------------ viscous flux (k march)
// viscousFlux (viscous.jl:1-134) on interior +-2 columns, one lane per (i, j) column marching k:
// u, v, w, T of the column are kept in a 5-plane register window for the zeta derivatives, so each
// plane is read from HBM once; the xi/eta neighbours come from L1/L2.  Output in the reference Fv
// layout (viscous.jl:119-132).  LES: the eddy viscosity model (F3D_LES_*).
template <typename T, int LES>
__global__ __launch_bounds__(256) void k_viscous_march(KP kp, FArgs<T> a, T* __restrict__ Fvx, T* __restrict__ Fvy,
                                                       T* __restrict__ Fvz, int KC, int remap) {
  const Blk B = xcd_block(remap);
  const int ii = B.x * 64 + threadIdx.x, jj = B.y * 4 + threadIdx.y;  // 0.. Np+3
  const int NX = kp.Nxp + 4, NY = kp.Nyp + 4, NZ = kp.Nzp + 4;
  const bool act = ii < NX && jj < NY;
  const int i = 2 + min(ii, NX - 1), j = 2 + min(jj, NY - 1);  // absolute, in [2, Nt-3]
  const int kk0 = B.z * KC;
  const int nk = min(KC, NZ - kk0);
  const long long nc = kp.nc, sy = kp.Nxt, sz = (long long)kp.Nxt * kp.Nyt;
  const long long vn = (long long)NX * NY * NZ;
  const T c1 = T(1) / T(12), c2 = T(-2) / T(3);
  // window w[v][s]: variable v (u,v,w,T) at plane kabs-2+s
  T w[4][5];
  const long long ocol = cidx(kp, i, j, 0);
  const int kabs0 = 2 + kk0;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const long long o = ocol + (long long)(kabs0 - 2 + s) * sz;
    w[0][s + 1] = a.Q[o + nc];
    w[1][s + 1] = a.Q[o + 2 * nc];
    w[2][s + 1] = a.Q[o + 3 * nc];
    w[3][s + 1] = a.Q[o + 5 * nc];
  }
#pragma unroll 1
  for (int t = 0; t < nk; ++t) {
    const int kabs = kabs0 + t;
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int s = 0; s < 4; ++s) w[v][s] = w[v][s + 1];
    const long long o = ocol + (long long)kabs * sz;
    {
      const long long on = o + 2 * sz;
      w[0][4] = a.Q[on + nc];
      w[1][4] = a.Q[on + 2 * nc];
      w[2][4] = a.Q[on + 3 * nc];
      w[3][4] = a.Q[on + 5 * nc];
    }
    T g[4][3];
    const int var[4] = {1, 2, 3, 5};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const T* f = a.Q + var[v] * nc + o;
      g[v][0] = c1 * (f[-2] - f[2]) + c2 * (f[-1] - f[1]);
      g[v][1] = c1 * (f[-2 * sy] - f[2 * sy]) + c2 * (f[-sy] - f[sy]);
      g[v][2] = c1 * (w[v][0] - w[v][4]) + c2 * (w[v][1] - w[v][3]);
    }
    T m[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = a.m.m[q][o];
    T out[3][4];
    viscous_cell<T, LES>(kp, w[0][2], w[1][2], w[2][2], w[3][2], g, m, (T)a.J[o], out,
                         LES != F3D_LES_NONE ? a.Q[o] : T(0));
    if (act) {
      const long long vo = (i - 2) + (long long)NX * ((j - 2) + (long long)NY * (kabs - 2));
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        Fvx[vo + n * vn] = out[0][n];
        Fvy[vo + n * vn] = out[1][n];
        Fvz[vo + n * vn] = out[2][n];
      }
    }
  }
}

// ------------------------------------------------------------------ update
// Viscous divergence (div.jl:14-39, 45-48) + conservative update + RK3 blend + c2Prim, one lane
// per interior (i, j) column marching k; Fv_z of the column is kept in a 5-plane window.
// MODE 0: RHS out (U layout, double) = J * (R + dFv).  MODE 1..3: RK3 stage KRK = MODE.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_update(KP kp, FArgs<T> a, const T* __restrict__ Fvx,
                                                const T* __restrict__ Fvy, const T* __restrict__ Fvz,
                                                double* U, double* __restrict__ Un, T* __restrict__ Q,
                                                double* __restrict__ Rout, double dt, int KC, int remap) {
  const Blk B = xcd_block(remap);
  double* const Uw = a.Uo ? a.Uo : U;
  const int i = B.x * 64 + threadIdx.x, j = B.y * 4 + threadIdx.y;
  const bool act = i < kp.Nxp && j < kp.Nyp;
  const int ic = min(i, kp.Nxp - 1), jc = min(j, kp.Nyp - 1);
  const int k0 = B.z * KC;
  const int nk = min(KC, kp.Nzp - k0);
  const int NG = kp.NG;
  const long long nc = kp.nc, sz = (long long)kp.Nxt * kp.Nyt;
  const long long rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
  const T c1 = T(1) / T(12), c2 = T(-2) / T(3);
  const int Nv1 = kp.Nxp + NG, Nv2 = kp.Nyp + NG;
  const long long vn = (long long)Nv1 * Nv2 * (kp.Nzp + NG), vsy = Nv1, vsz = (long long)Nv1 * Nv2;
  // Fv_z window: wz[n][s] = Fv_z at Fv plane (k+2) - 2 + s of this column
  T wz[4][5];
  const long long vcol = (ic + 2) + (long long)Nv1 * (jc + 2);
  if (kp.viscous) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int n = 0; n < 4; ++n) wz[n][s + 1] = Fvz[vcol + (long long)(k0 + s) * vsz + n * vn];
  }
#pragma unroll 1
  for (int t = 0; t < nk; ++t) {
    const int k = k0 + t;
    const long long o = cidx(kp, ic + NG, jc + NG, k + NG), ro = ridx(kp, ic, jc, k);
    T Rt[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) Rt[n] = a.R[ro + n * rn];
    if (kp.viscous) {
#pragma unroll
      for (int n = 0; n < 4; ++n) {
#pragma unroll
        for (int s = 0; s < 4; ++s) wz[n][s] = wz[n][s + 1];
        wz[n][4] = Fvz[vcol + (long long)(k + 4) * vsz + n * vn];
      }
      const long long vo = vcol + (long long)(k + 2) * vsz;
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        const T* fx = Fvx + vo + n * vn;
        const T* fy = Fvy + vo + n * vn;
        T sdiv = c1 * (fx[-2] - fx[2]) + c2 * (fx[-1] - fx[1]);
        sdiv += c1 * (fy[-2 * vsy] - fy[2 * vsy]) + c2 * (fy[-vsy] - fy[vsy]);
        sdiv += c1 * (wz[n][0] - wz[n][4]) + c2 * (wz[n][1] - wz[n][3]);
        Rt[n + 1] += sdiv;
      }
    }
    if (!act) continue;
    const double Jc = a.J[o];
    if (MODE == 0) {
#pragma unroll
      for (int n = 0; n < 5; ++n) Rout[o + n * nc] = (double)Rt[n] * Jc;
      continue;
    }
    const double Jact = Jc * (a.ldt ? a.ldt[o] : dt);
    double un[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      const double u0 = U[o + n * nc];
      const double u1 = u0 + (double)Rt[n] * Jact;
      double r;
      if (MODE == 1) {
        if (Un) Un[o + n * nc] = u0;  // (null: the caller keeps U^n in U, Uo takes the stage)
        r = u1;
      } else if (MODE == 2) {
        r = u1 * 0.25 + Un[o + n * nc] * 0.75;
      } else {
        r = u1 * (2.0 / 3.0) + Un[o + n * nc] * (1.0 / 3.0);
      }
      un[n] = r;
      Uw[o + n * nc] = r;
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
}

// ------------------------------------------------------------------ viscous divergence + update
// k_update with the viscous divergence computed in place (k_visc_update_qt below): the viscous fluxes
// never go to HBM.  Tile rows of W lanes (64 or 32) over W - 4 interior columns plus 2 halo lanes each
// side; with W = 32 a wave holds two rows.

// Conservative update + RK3 blend + c2Prim of one interior cell (utils.jl:2-70, div.jl:40-48);
// MODE 0 writes the RHS instead.
template <typename T, int MODE>
__device__ __forceinline__ void update_cell(const KP& kp, const T* Rt, long long oc, const double* U, double* Uw,
                                            double* __restrict__ Un, T* __restrict__ Q, double* __restrict__ Rout,
                                            double Jc, double dt, const double* __restrict__ ldt) {
  const long long nc = kp.nc;
  if (MODE == 0) {
#pragma unroll
    for (int n = 0; n < 5; ++n) Rout[oc + n * nc] = (double)Rt[n] * Jc;
    return;
  }
  const double Jact = Jc * (ldt ? ldt[oc] : dt);
  double un[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const double u0 = U[oc + n * nc];
    const double u1 = u0 + (double)Rt[n] * Jact;
    double r;
    if (MODE == 1) {
      if (Un) Un[oc + n * nc] = u0;
      r = u1;
    } else if (MODE == 2) {
      r = u1 * 0.25 + Un[oc + n * nc] * 0.75;
    } else {
      r = u1 * (2.0 / 3.0) + Un[oc + n * nc] * (1.0 / 3.0);
    }
    un[n] = r;
    Uw[oc + n * nc] = r;
  }
  const double eps = 2.220446049250313e-16;
  const double rho = jmax(un[0], eps);
  const double ri = 1.0 / rho;
  const double u = un[1] * ri, v = un[2] * ri, ww = un[3] * ri;
  const double ei = jmax(un[4] - 0.5 * rho * (u * u + v * v + ww * ww), eps);
  const double pr = (kp.gamma - 1) * ei;
  Q[oc] = (T)rho;
  Q[oc + nc] = (T)u;
  Q[oc + 2 * nc] = (T)v;
  Q[oc + 3 * nc] = (T)ww;
  Q[oc + 4 * nc] = (T)pr;
  Q[oc + 5 * nc] = (T)(pr / (rho * kp.Rg));
}

// Operands of one cell's conservative update (RHS increment, U, Un, J, dt), loaded one march
// iteration before they are used so that their memory latency overlaps the viscous evaluation.
template <typename T>
struct UpdOps {
  T R[5];
  double U[5], Un[5], J, dtc;
};

template <typename T, int MODE>
__device__ __forceinline__ void load_ops(const KP& kp, const FArgs<T>& a, const double* __restrict__ U,
                                         const double* __restrict__ Un, long long oc, long long ro, double dt,
                                         UpdOps<T>& op) {
  const long long nc = kp.nc, rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
#pragma unroll
  for (int n = 0; n < 5; ++n) op.R[n] = a.R[ro + n * rn];
  op.J = a.J[oc];
  if (MODE != 0) {
    op.dtc = a.ldt ? a.ldt[oc] : dt;
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      op.U[n] = U[oc + n * nc];
      if (MODE >= 2) op.Un[n] = Un[oc + n * nc];
    }
  }
}

// update_cell from preloaded operands (same arithmetic).
template <typename T, int MODE>
__device__ __forceinline__ void update_cell_ops(const KP& kp, const T* Rt, long long oc, const UpdOps<T>& op,
                                                double* __restrict__ U, double* __restrict__ Un, T* __restrict__ Q,
                                                double* __restrict__ Rout) {
  const long long nc = kp.nc;
  if (MODE == 0) {
#pragma unroll
    for (int n = 0; n < 5; ++n) Rout[oc + n * nc] = (double)Rt[n] * op.J;
    return;
  }
  const double Jact = op.J * op.dtc;
  double un[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const double u0 = op.U[n];
    const double u1 = u0 + (double)Rt[n] * Jact;
    double r;
    if (MODE == 1) {
      if (Un) Un[oc + n * nc] = u0;
      r = u1;
    } else if (MODE == 2) {
      r = u1 * 0.25 + op.Un[n] * 0.75;
    } else {
      r = u1 * (2.0 / 3.0) + op.Un[n] * (1.0 / 3.0);
    }
    un[n] = r;
    U[oc + n * nc] = r;
  }
  const double eps = 2.220446049250313e-16;
  const double rho = jmax(un[0], eps);
  const double ri = 1.0 / rho;
  const double u = un[1] * ri, v = un[2] * ri, ww = un[3] * ri;
  const double ei = jmax(un[4] - 0.5 * rho * (u * u + v * v + ww * ww), eps);
  const double pr = (kp.gamma - 1) * ei;
  Q[oc] = (T)rho;
  Q[oc + nc] = (T)u;
  Q[oc + 2 * nc] = (T)v;
  Q[oc + 3 * nc] = (T)ww;
  Q[oc + 4 * nc] = (T)pr;
  Q[oc + 5 * nc] = (T)(pr / (rho * kp.Rg));
}

// Block = W x (VI + 4) lanes over a (W - 4) x VI tile of interior cells; rows 0, 1, VI + 2, VI + 3 are
// y-halo rows (Fv_y only; waves of their own: with W = 32, VI is even so that a wave's two rows are both
// halo or both interior).  The block marches k: at iteration p (2 ahead of the
// updated plane k) every lane evaluates the velocity/temperature gradient and viscous_cell for its cell;
// Fv_x of the interior rows and Fv_y of the interior columns are exchanged through LDS for the x/y
// divergence of plane p, kept two iterations in a register ring; Fv_z enters a 5-plane register window for
// the z divergence of plane k; u,v,w,T of the column sit in a 5-plane window for the zeta derivatives.
// The xi and eta neighbours of u,v,w,T come from a plane tile in LDS: every lane stores its column's
// plane-p values (from its window) into the tile, the lanes at the x edges add the two cells beyond
// (x0-4, x0-3 and x0+W-2, x0+W-1) and the halo rows the two rows beyond (y0-4, y0-3 and y0+VI+2,
// y0+VI+3); after a barrier the 4th-order xi/eta differences read the tile.  Summation order of the
// divergence as k_update.
// The tile is written before barrier A and read before barrier B (the viscous-flux exchange), so one
// buffer suffices; likewise the flux buffers, written after A and read after B.
template <typename T, int MODE, int VI, bool PF, int W>
__global__ __launch_bounds__(W * (VI + 4)) void k_visc_update_qt(KP kp, FArgs<T> a, double* __restrict__ U,
                                                        double* __restrict__ Un, T* __restrict__ Q,
                                                        double* __restrict__ Rout, double dt, int KC, int remap,
                                                        int nfull) {
  static_assert(W == 64 || (W == 32 && VI % 2 == 0), "tile rows of 64 lanes, or two rows of 32 per wave");
  constexpr int WI = W - 4;              // interior columns of the tile
  __shared__ T sQ[4][VI + 8][W + 4];  // u,v,w,T of plane p: rows y0-4 .. y0+VI+3, columns x0-4 .. x0+W-1
  __shared__ T sFx[4][VI][W];
  __shared__ T sFy[4][VI + 4][WI];
  // Work of the block (1-D grid, started in workgroup order): the first nfull workgroups (normally whole
  // rounds of one per CU) march tile columns 0 .. nfull - 1 bottom to top; the others march one KC-plane
  // segment of the remaining columns, segment by segment, so that the blocks that run at one time march
  // neighbouring columns over the same planes.  Each part is renumbered apart (a bijection on the part;
  // it gives each XCD neighbouring columns when the part starts at a multiple of 8 workgroups).
  const unsigned id = blockIdx.x;
  const int L = !remap ? (int)id
                       : (id < (unsigned)nfull ? (int)xcd_rank(id, nfull)
                                               : nfull + (int)xcd_rank(id - nfull, gridDim.x - nfull));
  const int gx = (kp.Nxp + WI - 1) / WI, ncol = gx * ((kp.Nyp + VI - 1) / VI);
  int c, kk, nkk;
  if (L < nfull) {
    c = L;
    kk = 0;
    nkk = kp.Nzp;
  } else {
    const int j = L - nfull, rest = ncol - nfull;
    c = nfull + j % rest;
    kk = (j / rest) * KC;
    nkk = min(KC, kp.Nzp - kk);
  }
  const int lane = threadIdx.x, row = threadIdx.y;
  const int NG = kp.NG;
  const int x = NG + (c % gx) * WI - 2 + lane, y = NG + (c / gx) * VI - 2 + row;
  const int k0 = NG + kk;
  const int nk = nkk;
  const int xmax = NG + kp.Nxp + 1, ymax = NG + kp.Nyp + 1;
  const bool xin = lane >= 2 && lane < 2 + WI && x < NG + kp.Nxp;
  const bool yin = row >= 2 && row < 2 + VI && y < NG + kp.Nyp;
  const bool needFx = yin && x <= xmax;
  const bool needFy = xin && y <= ymax;
  const bool inner = xin && yin;
  // tile columns are loaded up to the array bound (the +1/+2 neighbours of the box's last cell)
  const int xl = min(x, kp.Nxt - 1), yl = min(y, kp.Nyt - 1);
  const long long nc = kp.nc, sy = kp.Nxt, sz = (long long)kp.Nxt * kp.Nyt;
  const long long rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
  const long long col = cidx(kp, xl, yl, 0);
  // extra cell of the edge lanes (x -+ 2) and of the halo rows (y -+ 2), clamped into the array
  const bool xe = lane < 2 || lane >= W - 2, ye = row < 2 || row >= VI + 2;
  const int xe_x = min(max(lane < 2 ? x - 2 : x + 2, 0), kp.Nxt - 1);
  const int ye_y = min(max(row < 2 ? y - 2 : y + 2, 0), kp.Nyt - 1);
  const long long colx = cidx(kp, xe_x, yl, 0), coly = cidx(kp, xl, ye_y, 0);
  const int cxe = lane < 2 ? lane : lane + 4;   // tile column of the edge cell
  const int rye = row < 2 ? row : row + 4;      // tile row of the extra row
  const int tc = lane + 2, tr = row + 2;        // own tile column / row
  const T c1 = T(1) / T(12), c2 = T(-2) / T(3);
  const int var[4] = {1, 2, 3, 5};
  T w[4][5], fz[4][5], xy[3][4];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int q = 0; q < 5; ++q) w[v][q] = fz[v][q] = T(0);
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int n = 0; n < 4; ++n) xy[q][n] = T(0);
#pragma unroll
  for (int q = 1; q < 5; ++q) {
    const long long o = col + (long long)(k0 - 5 + q) * sz;
#pragma unroll
    for (int v = 0; v < 4; ++v) w[v][q] = a.Q[o + var[v] * nc];
  }
  // PF: the metrics and J of the next plane and the operands of the next update are loaded one
  // iteration ahead, so that their latency overlaps this plane's evaluation
  // (FP32-bulk, whose registers allow it: the edge columns and extra rows too)
  constexpr bool PFE = PF && sizeof(T) == 4;
  T mn[9], exn[4], eyn[4];
  double Jn = 0;
  UpdOps<T> op;
  if constexpr (PF) {
    const long long o0 = col + (long long)(k0 - 2) * sz;
#pragma unroll
    for (int q = 0; q < 9; ++q) mn[q] = a.m.m[q][o0];
    Jn = a.J[o0];
    if constexpr (PFE) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        exn[v] = xe ? a.Q[colx + (long long)(k0 - 2) * sz + var[v] * nc] : T(0);
        eyn[v] = ye ? a.Q[coly + (long long)(k0 - 2) * sz + var[v] * nc] : T(0);
      }
    }
    if (inner && nk > 0)
      load_ops<T, MODE>(kp, a, U, Un, col + (long long)k0 * sz, ridx(kp, x - NG, y - NG, k0 - NG), dt, op);
  }
  // FP64 (no second edge set in registers): the edge cells of plane p, and the window's new plane p + 2, are
  // requested after the previous plane's flux exchange (barrier B), so their latency overlaps that plane's
  // update instead of stalling the block before barrier A (the edges) and the zeta differences after it
  // (measured on 512x192x64: 662 -> 638 us per FP64 launch; the new plane a step ahead costs FP32-bulk 6 us)
  constexpr bool WPF = sizeof(T) == 8;
  T ex[4], ey[4], wn[4];
  if constexpr (!PFE) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      ex[v] = xe ? a.Q[colx + (long long)(k0 - 2) * sz + var[v] * nc] : T(0);
      ey[v] = ye ? a.Q[coly + (long long)(k0 - 2) * sz + var[v] * nc] : T(0);
    }
  }
#pragma unroll 1
  for (int s = -4; s < nk; ++s) {
    const int p = k0 + s + 2;
    if constexpr (PFE) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        ex[v] = exn[v];
        ey[v] = eyn[v];
      }
      if (s + 1 < nk) {
        if (xe) {
#pragma unroll
          for (int v = 0; v < 4; ++v) exn[v] = a.Q[colx + (long long)(p + 1) * sz + var[v] * nc];
        }
        if (ye) {
#pragma unroll
          for (int v = 0; v < 4; ++v) eyn[v] = a.Q[coly + (long long)(p + 1) * sz + var[v] * nc];
        }
      }
    }
    if constexpr (WPF) {
      if (s == -4) {
        const long long o = col + (long long)(p + 2) * sz;
#pragma unroll
        for (int v = 0; v < 4; ++v) wn[v] = a.Q[o + var[v] * nc];
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
#pragma unroll
        for (int q = 0; q < 4; ++q) w[v][q] = w[v][q + 1];
        w[v][4] = wn[v];
      }
    } else {
      const long long o = col + (long long)(p + 2) * sz;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
#pragma unroll
        for (int q = 0; q < 4; ++q) w[v][q] = w[v][q + 1];
        w[v][4] = a.Q[o + var[v] * nc];
      }
    }
    const long long opl = col + (long long)p * sz;
    T m[9];
    T Jc;
    if constexpr (PF) {
#pragma unroll
      for (int q = 0; q < 9; ++q) m[q] = mn[q];
      Jc = (T)Jn;
      if (s + 1 < nk) {
#pragma unroll
        for (int q = 0; q < 9; ++q) mn[q] = a.m.m[q][opl + sz];
        Jn = a.J[opl + sz];
      }
    } else {
#pragma unroll
      for (int q = 0; q < 9; ++q) m[q] = a.m.m[q][opl];
      Jc = (T)a.J[opl];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      sQ[v][tr][tc] = w[v][2];
      if (xe) sQ[v][tr][cxe] = ex[v];
      if (ye) sQ[v][rye][tc] = ey[v];
    }
    __syncthreads();
    T out[3][4];
    if (needFx || needFy) {
      T g[4][3];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        g[v][0] = c1 * (sQ[v][tr][tc - 2] - sQ[v][tr][tc + 2]) + c2 * (sQ[v][tr][tc - 1] - sQ[v][tr][tc + 1]);
        g[v][1] = c1 * (sQ[v][tr - 2][tc] - sQ[v][tr + 2][tc]) + c2 * (sQ[v][tr - 1][tc] - sQ[v][tr + 1][tc]);
        g[v][2] = c1 * (w[v][0] - w[v][4]) + c2 * (w[v][1] - w[v][3]);
      }
      viscous_cell<T>(kp, w[0][2], w[1][2], w[2][2], w[3][2], g, m, Jc, out);
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int n = 0; n < 4; ++n) out[d][n] = T(0);
    }
#pragma unroll
    for (int n = 0; n < 4; ++n) {
#pragma unroll
      for (int q = 0; q < 4; ++q) fz[n][q] = fz[n][q + 1];
      fz[n][4] = out[2][n];
      if (needFx) sFx[n][row - 2][lane] = out[0][n];
      if (needFy) sFy[n][row][lane - 2] = out[1][n];
    }
    __syncthreads();
    if constexpr (WPF) {
      if (s + 1 < nk) {
        const long long o = col + (long long)(p + 3) * sz;
#pragma unroll
        for (int v = 0; v < 4; ++v) wn[v] = a.Q[o + var[v] * nc];
      }
    }
    if constexpr (!PFE) {
      if (s + 1 < nk) {
        if (xe) {
#pragma unroll
          for (int v = 0; v < 4; ++v) ex[v] = a.Q[colx + (long long)(p + 1) * sz + var[v] * nc];
        }
        if (ye) {
#pragma unroll
          for (int v = 0; v < 4; ++v) ey[v] = a.Q[coly + (long long)(p + 1) * sz + var[v] * nc];
        }
      }
    }
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      xy[0][n] = xy[1][n];
      xy[1][n] = xy[2][n];
    }
    if (inner) {
      const int r = row - 2, l = lane - 2;
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        T d = c1 * (sFx[n][r][lane - 2] - sFx[n][r][lane + 2]) + c2 * (sFx[n][r][lane - 1] - sFx[n][r][lane + 1]);
        d += c1 * (sFy[n][row - 2][l] - sFy[n][row + 2][l]) + c2 * (sFy[n][row - 1][l] - sFy[n][row + 1][l]);
        xy[2][n] = d;
      }
    }
    if (s < 0 || !inner) continue;
    const int k = k0 + s;
    const long long oc = col + (long long)k * sz, ro = ridx(kp, x - NG, y - NG, k - NG);
    T Rt[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) Rt[n] = PF ? op.R[n] : a.R[ro + n * rn];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      T sdiv = xy[0][n];
      sdiv += c1 * (fz[n][0] - fz[n][4]) + c2 * (fz[n][1] - fz[n][3]);
      Rt[n + 1] += sdiv;
    }
    if constexpr (PF) {
      update_cell_ops<T, MODE>(kp, Rt, oc, op, a.Uo ? a.Uo : U, Un, Q, Rout);
      if (s + 1 < nk) load_ops<T, MODE>(kp, a, U, Un, oc + sz, ro + (long long)kp.Nxp * kp.Nyp, dt, op);
    } else {
      update_cell<T, MODE>(kp, Rt, oc, U, a.Uo ? a.Uo : U, Un, Q, Rout, a.J[oc], dt, a.ldt);
    }
  }
}

// ------------------------------------------------------------------ orchestration
// Segment length of the y/z marches; F3D_MARCH_SEG_Y / F3D_MARCH_SEG_Z override the defaults.
static int march_seg(int D, int es) {
  static int v[2][2] = {{-1, -1}, {-1, -1}};
  int& r = v[D - 1][es == 8];
  if (r < 0) {
    const char* e = std::getenv(D == 1 ? "F3D_MARCH_SEG_Y" : "F3D_MARCH_SEG_Z");
    // defaults measured on the 512x192x64 ramp (one MI355X): FP64 y 48, z 32; FP32 y 24, z 32
    const int def = es == 8 ? (D == 1 ? 48 : 32) : (D == 1 ? 24 : 32);
    r = e ? std::atoi(e) : def;
    if (r < 8) r = def;
  }
  return r;
}

static int seg_len(int Np, int max_len) {
  const int nseg = (Np + max_len - 1) / max_len;
  return (Np + nseg - 1) / nseg;
}

static dim3 visc_grid(const KP& kp, int KC, int rows, int W) {
  return dim3((kp.Nxp + W - 5) / (W - 4), (kp.Nyp + rows - 1) / rows, (kp.Nzp + KC - 1) / KC);
}

// Listed faces formed by a pair of lanes each (k_heavy_pair) or one lane each (k_heavy).  Default: pairs in
// FP64 (one wave per SIMD becomes two: forming + gather 503 -> 480 us per stage with 4.6 % of the faces off
// UP7 on the 512x192x64 ramp), one lane in FP32-bulk (two waves per SIMD already; pairs 405 against 395
// us).  F3D_HEAVY_PAIR=0 / 1 forces either.
static int heavy_pairs(int es) {
  static int v = -2;
  if (v == -2) {
    const char* e = std::getenv("F3D_HEAVY_PAIR");
    v = e ? (std::strcmp(e, "0") == 0 ? 0 : 1) : -1;
  }
  return v >= 0 ? v : (es == 8 ? 1 : 0);
}

static int env_int(const char* name, int def) {
  const char* e = std::getenv(name);
  return e ? std::atoi(e) : def;
}

// Compute units of the current device (one viscous-update block per CU).
static int cu_count() {
  static int n[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (n[dev] == 0) {
    int v = 0;
    n[dev] = (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) ? v : 256;
  }
  return n[dev];
}

// 1 (default): the next plane's metrics and update operands prefetched an iteration ahead; 0 without
// (A/B measurement).  F3D_VISC_PF overrides.
static int visc_pf(int es) {
  static int v[2] = {-2, -2};
  int& r = v[es == 8];
  if (r == -2) r = env_int("F3D_VISC_PF", 1) != 0;
  return r;
}

// Tile of the viscous-update march: VI interior rows of W lanes per block (plus 4 halo rows; a wave is
// one row of 64 or two rows of 32) and KC planes per segment (each segment re-evaluates 4 planes of
// viscous fluxes).  One block runs per CU (FP64 with the prefetch: 242 VGPRs, two waves per SIMD, 8-wave
// blocks; FP32-bulk: 152 VGPRs, three, up to 12-wave blocks), so the march's duration is about rounds x
// waves x (KC + 4) iterations, rounds = ceil(blocks / CUs); the chooser minimises that over the forms
// built.  On the 512x192x64 ramp: FP64 VI = 12 rows of 32 (28 x 12 interior cells per 8 waves instead of
// 60 x 4: half the y-halo loads and evaluations), KC = 16: 679 us against 762 for 4 rows of 64 (KC = 8,
// 11, 13, 22, 32: 741, 752, 681, 698, 721 us, the order the estimate gives); FP32-bulk VI = 7 rows of 64,
// KC = 64: 252 blocks for 256 CUs, 405 us (8 rows: 428; rows of 32 lanes: 441-490).  F3D_VISC_FORM (VI),
// F3D_VISC_W (32 or 64) and F3D_VISC_KC override; forms that are not built fall back to the default.
struct ViscTile {
  int VI, KC, W;
  int nfull;  // tile columns marched whole (the first rounds) before the KC-plane segments of the rest
};

// (VI, W) forms built: with the prefetch FP64 (4, 64), (8, 32), (12, 32) and FP32-bulk (4|6|7|8, 64),
// (12|16|20, 32); without it (4|8, 64) and FP32-bulk (12, 64).
static bool visc_form_built(int es, bool pf, int vi, int w) {
  if (pf) {
    if (es == 8) return (vi == 4 && w == 64) || ((vi == 8 || vi == 12) && w == 32);
    return (w == 64 && (vi == 4 || vi == 6 || vi == 7 || vi == 8)) || (w == 32 && (vi == 12 || vi == 16 || vi == 20));
  }
  return w == 64 && (vi == 4 || vi == 8 || (vi == 12 && es == 4));
}

static ViscTile visc_tile(const KP& kp, int es) {
  const int fenv = env_int("F3D_VISC_FORM", 0), kenv = env_int("F3D_VISC_KC", 0);
  const int wenv = env_int("F3D_VISC_W", 0) == 32 ? 32 : 64;
  const bool pf = visc_pf(es) != 0;
  struct Form {
    int vi, w;
  };
  static const Form f64pf[2] = {{12, 32}, {4, 64}}, f64[1] = {{4, 64}};
  static const Form f32pf[3] = {{8, 64}, {7, 64}, {6, 64}}, f32[1] = {{8, 64}};
  static const int kc64[6] = {64, 32, 22, 16, 13, 8}, kc32[3] = {64, 32, 16};
  const Form forced{fenv, wenv};
  const Form* forms = es == 8 ? (pf ? f64pf : f64) : (pf ? f32pf : f32);
  int nf = es == 8 ? (pf ? 2 : 1) : (pf ? 3 : 1);
  if (fenv && visc_form_built(es, pf, fenv, wenv)) forms = &forced, nf = 1;
  const int* kcs = es == 8 ? kc64 : kc32;
  const int nkc = kenv >= 4 ? 1 : (es == 8 ? 6 : 3);
  const long long ncu = cu_count();
  // F3D_VISC_FULL = n: n whole columns (0: segments only; unset: whole rounds of one column per CU, as
  // many as the estimate takes, the last of them with the columns left when fewer than the CUs)
  const int fenv2 = env_int("F3D_VISC_FULL", -1);
  ViscTile best{forms[0].vi, 16, forms[0].w, 0};
  double bestc = -1;
  for (int a = 0; a < nf; ++a) {
    const int vi = forms[a].vi, w = forms[a].w;
    const long long cols = (long long)((kp.Nxp + w - 5) / (w - 4)) * ((kp.Nyp + vi - 1) / vi);
    const long long maxr = fenv2 >= 0 ? 0 : (cols + ncu - 1) / ncu;
    for (long long r = 0; r <= maxr; ++r) {
      const long long whole = std::min<long long>(fenv2 >= 0 ? fenv2 : r * ncu, cols), rest = cols - whole;
      for (int b = 0; b < nkc; ++b) {
        const int kc = seg_len(kp.Nzp, kenv >= 4 ? kenv : kcs[b]);
        const long long blocks = rest * ((kp.Nzp + kc - 1) / kc);
        const double c = ((double)((whole + ncu - 1) / ncu) * (kp.Nzp + 4) +
                          (double)((blocks + ncu - 1) / ncu) * (kc + 4)) * ((vi + 4) * w / 64.0);
        if (bestc < 0 || c < bestc * 0.999) {
          bestc = c;
          best = ViscTile{vi, kc, w, (int)whole};
        }
      }
    }
  }
  return best;
}

template <typename T, int MODE, int VI, bool PF, int W>
static void launch_qt(const KP& kp, const FArgs<T>& a, double* U, double* Un, T* Q, double* Rout, double dt, int KC,
                      int nfull, int rm, hipStream_t s) {
  const dim3 g = visc_grid(kp, KC, VI, W);
  const int cols = (int)(g.x * g.y);
  if (nfull > cols) nfull = cols;
  const unsigned nb = (unsigned)(nfull + (cols - nfull) * (int)g.z);
  hipLaunchKernelGGL((k_visc_update_qt<T, MODE, VI, PF, W>), dim3(nb, 1, 1), dim3(W, VI + 4, 1), 0, s, kp, a, U, Un,
                     Q, Rout, dt, KC, rm, nfull);
}

template <typename T, int MODE>
static void launch_visc_update(const KP& kp, const FArgs<T>& a, double* U, double* Un, T* Q, double* Rout, double dt,
                               const ViscTile& vt, int rm, hipStream_t s) {
  constexpr int es = (int)sizeof(T);
  const bool pf = visc_pf(es) != 0;
  int vi = vt.VI, w = vt.W;
  if (!visc_form_built(es, pf, vi, w)) {
    // the forms built without the prefetch: 4, 8 or (FP32-bulk) 12 rows of 64
    w = 64;
    vi = pf ? (es == 8 ? 4 : 8) : (vi <= 4 ? 4 : (vi <= 8 || es == 8) ? 8 : 12);
  }
  const int KC = vt.KC;
  // whole columns counted for the form chosen; a form not built falls back to segments only
  const int nfull = (vi == vt.VI && w == vt.W) ? vt.nfull : 0;
#define F3D_QT(V, P, WW) launch_qt<T, MODE, V, P, WW>(kp, a, U, Un, Q, Rout, dt, KC, nfull, rm, s)
  if (pf) {
    if (w == 64 && vi == 4) F3D_QT(4, true, 64);
    if constexpr (es == 8) {
      if (w == 32 && vi == 8) F3D_QT(8, true, 32);
      if (w == 32 && vi == 12) F3D_QT(12, true, 32);
    } else {
      if (w == 64 && vi == 6) F3D_QT(6, true, 64);
      if (w == 64 && vi == 7) F3D_QT(7, true, 64);
      if (w == 64 && vi == 8) F3D_QT(8, true, 64);
      if (w == 32 && vi == 12) F3D_QT(12, true, 32);
      if (w == 32 && vi == 16) F3D_QT(16, true, 32);
      if (w == 32 && vi == 20) F3D_QT(20, true, 32);
    }
    return;
  }
  if (vi == 4) F3D_QT(4, false, 64);
  if (vi == 8) F3D_QT(8, false, 64);
  if constexpr (es == 4) {
    if (vi == 12) F3D_QT(12, false, 64);
  }
#undef F3D_QT
}

// Planes per block of the marching sensor (F3D_SENSOR_KS; 0 = one plane per block, k_sensor_f).
static int sensor_ks() {
  static int v = -2;
  if (v == -2) {
    const char* e = std::getenv("F3D_SENSOR_KS");
    v = e ? std::atoi(e) : 16;
    if (v < 0) v = 16;
  }
  return v;
}

template <typename T>
void fused_sensor(const KP& kp, const WsView<T>& ws, const T* Q, T* phi, hipStream_t s) {
  TimedScope ts("fused_sensor", s);
  const int KS = sensor_ks();
  // a new generation for the band flag (generation 0 is never current: the flag starts cleared)
  if (ws.pgen && ++*ws.pgen == 0) *ws.pgen = 1;
  const unsigned int gen = ws.pgen ? *ws.pgen : 0;
  unsigned int* pf = ws.pgen ? ws.pflag : nullptr;
  if (KS > 0)
    hipLaunchKernelGGL(k_sensor_march<T>, dim3((kp.Nxt - 2 + 63) / 64, (kp.Nyt - 2 + 3) / 4, (kp.Nzt - 2 + KS - 1) / KS),
                       dim3(64, 4, 1), 0, s, kp, phi, ws.cls, Q, KS, pf, gen);
  else
    hipLaunchKernelGGL(k_sensor_f<T>, dim3((kp.Nxt - 2 + 63) / 64, (kp.Nyt - 2 + 3) / 4, kp.Nzt - 2), dim3(64, 4, 1),
                       0, s, kp, phi, ws.cls, Q, pf, gen);
}

// Launch shapes of the finite-difference sweeps (shared by the launches and their heavy-list layout).
static dim3 grid_sweep_x(const KP& kp, int L) { return dim3((kp.Nxp + L - 1) / L, (kp.Nyp + 1) / 2, kp.Nzp); }
static dim3 grid_march(const KP& kp, int D, int L) {
  const int Nt2 = D == 1 ? kp.Nzp : kp.Nyp, Np = D == 1 ? kp.Nyp : kp.Nzp;
  return dim3((kp.Nxp + 63) / 64, (Nt2 + 3) / 4, (Np + L - 1) / L);
}
static dim3 grid_xy(const KP& kp, int L) {
  return dim3((kp.Nxp + XYI - 1) / XYI, (kp.Nzp + XYL - 1) / XYL, (kp.Nyp + L - 1) / L);
}
static unsigned long long nblocks(const dim3& g) { return (unsigned long long)g.x * g.y * g.z; }

// 1 (default): x and y faces in one march (k_sweep_xy); F3D_SWEEP_XY=0 runs k_sweep_x and the y march
// (A/B measurement; the two forms agree to rounding: the same sums in the same order, contracted into
// FMAs differently by the compiler).
static int sweep_xy() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("F3D_SWEEP_XY");
    v = (e && std::strcmp(e, "0") == 0) ? 0 : 1;
  }
  return v;
}

// Resident waves per CU of a kernel launched in blocks of `threads` lanes (the occupancy query;
// `fallback` when it fails, e.g. without a device).
static int waves_per_cu(const void* kern, int threads, int fallback) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, threads, 0) != hipSuccess || nb <= 0) {
    (void)hipGetLastError();
    return fallback;
  }
  return nb * (threads / 64);
}

// Row-segment length of the x+y march.  Its launch has few enough waves that their count against the
// chip's resident-wave slots sets the duration: w waves over s slots take about (L + 8) * ceil(w / s)
// marching steps when they overfill the chip, and (L + 8) / sqrt(w / s) when they fill only part of it
// (fewer waves keep fewer bytes in flight); the chooser minimises that over L = 16 .. 64.  Measured on
// the 512x192x64 ramp, FP64 L = 16, 24, 32, 48, 64, 96: 281, 285, 271, 297, 264, 344 us; FP32-bulk 164,
// 166, 187, 161, 199, 282 us; the estimate ranks them the same way.  F3D_MARCH_SEG_Y overrides.
static int xy_seg(const KP& kp, int es) {
  const char* e = std::getenv("F3D_MARCH_SEG_Y");
  if (e && std::atoi(e) >= 8) return seg_len(kp.Nyp, std::atoi(e));
  static int wpc[2] = {0, 0};
  int& w = wpc[es == 8];
  if (w == 0)
    w = es == 8 ? waves_per_cu((const void*)&k_sweep_xy<double, F3D_SPLIT_SW>, XYW * XYL, 8)
                : waves_per_cu((const void*)&k_sweep_xy<float, F3D_SPLIT_SW>, XYW * XYL, 12);
  const double slots = (double)cu_count() * w;
  const double per_seg = (double)XYL * ((kp.Nxp + XYI - 1) / XYI) * ((kp.Nzp + XYL - 1) / XYL);
  static const int cand[5] = {16, 24, 32, 48, 64};
  int best = seg_len(kp.Nyp, 64);
  double bc = -1;
  for (int c : cand) {
    const int L = seg_len(kp.Nyp, c);
    const double f = per_seg * ((kp.Nyp + L - 1) / L) / slots;
    const double cost = (L + 8) * (f > 1 ? std::ceil(f) : 1.0 / std::sqrt(f));
    if (bc < 0 || cost < bc * 0.999) {
      bc = cost;
      best = L;
    }
  }
  return best;
}

HList fdm_heavy_layout(const KP& kp, int es) {
  const int Lx = seg_len(kp.Nxp, XW - 8);
  const int Ly = seg_len(kp.Nyp, march_seg(1, es)), Lz = seg_len(kp.Nzp, march_seg(2, es));
  if (sweep_xy()) {
    // one wave per z line of a block; per row at most one x face per lane, per closed y face one per lane
    const int Lxy = xy_seg(kp, es);
    const unsigned long long w = XYL * nblocks(grid_xy(kp, Lxy));
    const unsigned long long nw[3] = {w, w, 4 * nblocks(grid_march(kp, 2, Lz))};
    const unsigned long long cap[3] = {64ULL * Lxy, 64ULL * (Lxy + 1), 64ULL * (Lz + 1)};
    return make_hlist(nw, cap);
  }
  // 4 waves per block; an x wave lists at most one face per lane, a march wave one per lane and face
  const unsigned long long nw[3] = {4 * nblocks(grid_sweep_x(kp, Lx)), 4 * nblocks(grid_march(kp, 1, Ly)),
                                    4 * nblocks(grid_march(kp, 2, Lz))};
  const unsigned long long cap[3] = {64, 64ULL * (Ly + 1), 64ULL * (Lz + 1)};
  return make_hlist(nw, cap);
}

template <typename T>
int check_hlist(const WsView<T>& ws, const HList& h, const char* who) {
  if (h.slots() > ws.hcap || h.nruns() > ws.hncap) {
    set_error("%s: heavy-face list of %llu slots / %u runs exceeds the workspace's %llu / %llu (workspace created "
              "with other segment lengths)", who, h.slots(), h.nruns(), ws.hcap, ws.hncap);
    return F3D_E_ARG;
  }
  return F3D_OK;
}
template int check_hlist<double>(const WsView<double>&, const HList&, const char*);
template int check_hlist<float>(const WsView<float>&, const HList&, const char*);

// Inviscid increments of the finite-difference branch into ws.R: x, y, z sweeps and the heavy-face
// corrections (after fused_sensor), for the splitting SPLIT and the reconstruction kp.character.
template <typename T, int SPLIT>
static int fused_flux_split(const KP& kp, const WsView<T>& ws, const T* Q, const MetricPtrs<T>& m, const double* J,
                            T* phi, hipStream_t s) {
  FArgs<T> a{Q, m, J, ws.R, ws.heavy, ws.hcount};
  a.cls = ws.cls;
  a.pflag = ws.pgen ? ws.pflag : nullptr;
  a.pgen = ws.pgen ? *ws.pgen : 0u;
  const HList hl = fdm_heavy_layout(kp, (int)sizeof(T));
  if (int rc = check_hlist(ws, hl, "fused sweeps")) return rc;
  if (sweep_xy()) {
    TimedScope ts("fused_sweep_xy", s);
    const int L = xy_seg(kp, (int)sizeof(T));
    hipLaunchKernelGGL((k_sweep_xy<T, SPLIT>), grid_xy(kp, L), dim3(XYW, XYL, 1), 0, s, kp, a, L, hl.d[0], hl.d[1]);
  } else {
    {
      TimedScope ts("fused_sweep_x", s);
      const int L = seg_len(kp.Nxp, XW - 8);
      hipLaunchKernelGGL((k_sweep_x<T, SPLIT>), grid_sweep_x(kp, L), dim3(XW, 2, 1), 0, s, kp, a, L, hl.d[0]);
    }
    {
      TimedScope ts("fused_sweep_y", s);
      const int L = seg_len(kp.Nyp, march_seg(1, (int)sizeof(T)));
      hipLaunchKernelGGL((k_sweep_march<T, 1, SPLIT>), grid_march(kp, 1, L), dim3(64, 4, 1), 0, s, kp, a, L, hl.d[1]);
    }
  }
  {
    TimedScope ts("fused_sweep_z", s);
    const int L = seg_len(kp.Nzp, march_seg(2, (int)sizeof(T)));
    hipLaunchKernelGGL((k_sweep_march<T, 2, SPLIT>), grid_march(kp, 2, L), dim3(64, 4, 1), 0, s, kp, a, L, hl.d[2]);
  }
  {
    TimedScope ts("fused_heavy", s);
    const unsigned int nb = hl.items() < 8192u ? (hl.items() ? hl.items() : 1u) : 8192u;
    const HMark hm = hmarks(ws.hidx);
    if (heavy_pairs((int)sizeof(T))) {
      if (kp.character)
        hipLaunchKernelGGL((k_heavy_pair<T, SPLIT, true>), dim3(nb), dim3(64), 0, s, kp, a, hl, ws.hflux, hm);
      else
        hipLaunchKernelGGL((k_heavy_pair<T, SPLIT, false>), dim3(nb), dim3(64), 0, s, kp, a, hl, ws.hflux, hm);
    } else if (kp.character) {
      hipLaunchKernelGGL((k_heavy<T, SPLIT, true>), dim3(nb), dim3(64), 0, s, kp, a, hl, ws.hflux, hm);
    } else {
      hipLaunchKernelGGL((k_heavy<T, SPLIT, false>), dim3(nb), dim3(64), 0, s, kp, a, hl, ws.hflux, hm);
    }
    launch_heavy_apply<T>(kp, ws.R, hm, ws.hflux, s, a.pflag, a.pgen);
  }
  return check_launch("fused sweeps");
}

template <typename T>
static int fused_flux(const KP& kp, const WsView<T>& ws, const T* Q, const MetricPtrs<T>& m, const double* J,
                      T* phi, hipStream_t s) {
  switch (kp.split) {
    case F3D_SPLIT_LF: return fused_flux_split<T, F3D_SPLIT_LF>(kp, ws, Q, m, J, phi, s);
    case F3D_SPLIT_VL: return fused_flux_split<T, F3D_SPLIT_VL>(kp, ws, Q, m, J, phi, s);
    case F3D_SPLIT_AUSM: return fused_flux_split<T, F3D_SPLIT_AUSM>(kp, ws, Q, m, J, phi, s);
    default: return fused_flux_split<T, F3D_SPLIT_SW>(kp, ws, Q, m, J, phi, s);
  }
}

template <typename T>
int fused_update(const KP& kp, const WsView<T>& ws, int mode, double* U, double* Un, const T* Qin, T* Qout,
                 const MetricPtrs<T>& m, const double* J, const T* phi, double* Rout, double dt, hipStream_t s,
                 double* Uo) {
  if (mode != 0 && Qin == Qout && stage_needs_separate_q(kp)) {
    set_error("fused stage: the output state must not alias the input state");
    return F3D_E_ARG;
  }
  FArgs<T> a{Qin, m, J, ws.R, ws.heavy, ws.hcount};
  if (mode != 0 && kp.lts) a.ldt = ws.ldt;
  if (mode != 0) a.Uo = Uo;
  const int rm = xcd_remap();
  // LES configurations take the two-kernel form (viscous fluxes through HBM): the in-place viscous
  // update is sized to the laminar cell's registers
  const bool inplace = visc_fused() && kp.les == F3D_LES_NONE;
  if (kp.viscous && !inplace) {
    TimedScope ts("fused_viscous", s);
    const int KC = seg_len(kp.Nzp + 4, 32);
    const dim3 g((kp.Nxp + 4 + 63) / 64, (kp.Nyp + 4 + 3) / 4, (kp.Nzp + 4 + KC - 1) / KC), b(64, 4, 1);
    if (kp.les == F3D_LES_SMAG)
      hipLaunchKernelGGL((k_viscous_march<T, F3D_LES_SMAG>), g, b, 0, s, kp, a, ws.Fvx, ws.Fvy, ws.Fvz, KC, rm);
    else if (kp.les == F3D_LES_WALE)
      hipLaunchKernelGGL((k_viscous_march<T, F3D_LES_WALE>), g, b, 0, s, kp, a, ws.Fvx, ws.Fvy, ws.Fvz, KC, rm);
    else
      hipLaunchKernelGGL((k_viscous_march<T, F3D_LES_NONE>), g, b, 0, s, kp, a, ws.Fvx, ws.Fvy, ws.Fvz, KC, rm);
  }
  TimedScope ts("fused_update", s);
  if (kp.viscous && inplace) {
    const ViscTile KV = visc_tile(kp, (int)sizeof(T));
    if (mode == 0)
      launch_visc_update<T, 0>(kp, a, nullptr, nullptr, nullptr, Rout, 0.0, KV, rm, s);
    else if (mode == 1)
      launch_visc_update<T, 1>(kp, a, U, Un, Qout, nullptr, dt, KV, rm, s);
    else if (mode == 2)
      launch_visc_update<T, 2>(kp, a, U, Un, Qout, nullptr, dt, KV, rm, s);
    else
      launch_visc_update<T, 3>(kp, a, U, Un, Qout, nullptr, dt, KV, rm, s);
    return check_launch("fused update");
  }
  const int KC = seg_len(kp.Nzp, 16);
  const dim3 g((kp.Nxp + 63) / 64, (kp.Nyp + 3) / 4, (kp.Nzp + KC - 1) / KC), b(64, 4, 1);
  if (mode == 0)
    hipLaunchKernelGGL((k_update<T, 0>), g, b, 0, s, kp, a, ws.Fvx, ws.Fvy, ws.Fvz, nullptr, nullptr, nullptr, Rout,
                       0.0, KC, rm);
  else if (mode == 1)
    hipLaunchKernelGGL((k_update<T, 1>), g, b, 0, s, kp, a, ws.Fvx, ws.Fvy, ws.Fvz, U, Un, Qout, nullptr, dt, KC, rm);
  else if (mode == 2)
    hipLaunchKernelGGL((k_update<T, 2>), g, b, 0, s, kp, a, ws.Fvx, ws.Fvy, ws.Fvz, U, Un, Qout, nullptr, dt, KC, rm);
  else
    hipLaunchKernelGGL((k_update<T, 3>), g, b, 0, s, kp, a, ws.Fvx, ws.Fvy, ws.Fvz, U, Un, Qout, nullptr, dt, KC, rm);
  return check_launch("fused update");
}

template <typename T>
int fused_rhs(const KP& kp, const WsView<T>& ws, const T* Q, const MetricPtrs<T>& m, const double* J, T* phi,
              double* R, hipStream_t s) {
  fused_sensor<T>(kp, ws, Q, phi, s);
  int rc = fused_flux<T>(kp, ws, Q, m, J, phi, s);
  if (rc) return rc;
  return fused_update<T>(kp, ws, 0, nullptr, nullptr, Q, nullptr, m, J, phi, R, 0.0, s);
}

bool stage_needs_separate_q(const KP& kp) {
  return use_fused() && kp.viscous && visc_fused() && kp.les == F3D_LES_NONE;
}

template <typename T>
int fused_stage(const KP& kp, const WsView<T>& ws, int KRK, double* U, double* Un, const T* Qin, T* Qout,
                const MetricPtrs<T>& m, const double* J, double dt, hipStream_t s, double* Uo) {
  if (Qin == Qout && stage_needs_separate_q(kp)) {
    set_error("fused stage: the output state must not alias the input state");
    return F3D_E_ARG;
  }
  fused_sensor<T>(kp, ws, Qin, nullptr, s);  // the sweeps read the sensor classes; phi is not kept
  int rc = fused_flux<T>(kp, ws, Qin, m, J, ws.phi, s);
  if (rc) return rc;
  return fused_update<T>(kp, ws, KRK, U, Un, Qin, Qout, m, J, ws.phi, nullptr, dt, s, Uo);
}

template int fused_rhs<double>(const KP&, const WsView<double>&, const double*, const MetricPtrs<double>&,
                               const double*, double*, double*, hipStream_t);
template int fused_rhs<float>(const KP&, const WsView<float>&, const float*, const MetricPtrs<float>&, const double*,
                              float*, double*, hipStream_t);
template int fused_stage<double>(const KP&, const WsView<double>&, int, double*, double*, const double*, double*,
                                 const MetricPtrs<double>&, const double*, double, hipStream_t, double*);
template int fused_stage<float>(const KP&, const WsView<float>&, int, double*, double*, const float*, float*,
                                const MetricPtrs<float>&, const double*, double, hipStream_t, double*);
template void fused_sensor<double>(const KP&, const WsView<double>&, const double*, double*, hipStream_t);
template void fused_sensor<float>(const KP&, const WsView<float>&, const float*, float*, hipStream_t);
template int fused_update<double>(const KP&, const WsView<double>&, int, double*, double*, const double*, double*,
                                  const MetricPtrs<double>&, const double*, const double*, double*, double,
                                  hipStream_t, double*);
template int fused_update<float>(const KP&, const WsView<float>&, int, double*, double*, const float*, float*,
                                 const MetricPtrs<float>&, const double*, const float*, double*, double, hipStream_t, double*);
}  // namespace f3d
