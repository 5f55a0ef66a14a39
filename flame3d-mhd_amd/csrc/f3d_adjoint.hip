// Adjoint of the FP64 RHS / RK3 stage (hand-written reverse mode, scheme mask frozen) and the
// checkpointed multi-step adjoint.
//
// RHS VJP, lambda (interior, U layout) -> lamQ = (dR/dQ)^T lambda over every cell of Q:
//   k_adj_lambda     Lam = scale * J * lambda on interior cells (ghost entries stay zero)
//   k_adj_heavy_list faces whose 6-cell sensor maximum selects WENO7/WENO5/NND, compacted
//   k_adj_inv_z, k_adj_inv_xy  every face taken as UP7: the transposed difference-of-UP7 stencil
//                    gathers dFp, dFm per cell and direction from Lam (register windows along the
//                    marches, an LDS row for x), then the reverse of the Steger-Warming split gives
//                    (rho,u,v,w,T) adjoints; lamQ written for all cells (k_adj_inviscid: the same in
//                    one kernel gathering from the cache, F3D_ADJ_INV=gather)
//   k_adj_heavy_form per listed face: (scheme - UP7) reverse through the characteristic WENO /
//                    NND reconstruction, the frame and the eight splits, kept per face (its stencil
//                    cells flagged); the dense inviscid kernels add each flagged cell's kept
//                    contributions in a fixed order (no floating-point atomics); k_adj_heavy_rest:
//                    faces beyond the buffer, atomically
//   k_adj_visc1      per cell of the interior+-2 box: transposed viscous divergence -> adjoint of
//                    the cell's viscous fluxes -> adjoint of its 12 derivatives and (u,v,w,T)
//   k_adj_visc2      transposed 4th-order derivative stencils gather those into lamQ
// Stage VJP: RHS VJP with scale a*dt, reverse of the ghost procedure (fill z,y,x then exchange
// z,y,x; ghost adjoints fold into their sources, Dirichlet / inflow / isothermal-wall values
// into parameter gradients) and k_adj_finalize (reverse of c2Prim + the RK3 blend).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "f3d_adjoint.hpp"
#include "f3d_fused.hpp"
#include "f3d_internal.hpp"

using namespace f3d;

struct f3d_adj_workspace {
  f3d_params p;
  size_t bytes;
  void* base;
  f3d_workspace* fws;  // FP64 forward workspace: sensor, heavy-face list, stage recomputation
  double* Lam;         // scale * J * lambda (5 vars, cell layout, ghosts zero)
  double* lamQ;        // adjoint of Q (6 vars, cell layout)
  double* G;           // viscous partial adjoints (16 vars, cell layout, interior+-2 box)
  double* Qs[3];       // primitive states with ghosts of the stage inputs U^n, U^1, U^2 of one step
  double* Un;          // Un of the step being recomputed
  double* lam;         // adjoint of the state (5 vars)
  double* lamUn;       // adjoint of Un (5 vars)
  double* Ucur;        // forward sweep state (5 vars)
  double* Qcur;        // forward sweep primitive state (6 vars)
  double* red;         // reduction partials
  int nred;
  // listed faces' reverse contributions, gathered by the dense inviscid kernels in a fixed order per cell
  // (no floating-point atomics): hbuf[40 c + 5 s + q] = contribution of the face with compact index c to
  // variable q (rho, u, v, w, T) of its stencil cell s; hrec[d][o] = (stamp << 24) | (c + 1) for the face
  // of direction d whose right cell is o; hctr counts the compact indices taken (faces beyond hcap keep
  // the atomic path).
  double* hbuf;
  unsigned int* hrec;
  unsigned char* hflag;
  unsigned int* hctr;
  unsigned int hcap;
  unsigned int hstamp;
};

namespace {

constexpr int kObjBlocks = 1024;

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

__device__ __forceinline__ bool interior(int a, int Np, int NG) { return a >= NG && a < NG + Np; }

// ------------------------------------------------------------------ small kernels
__global__ __launch_bounds__(256) void k_adj_zero(double* __restrict__ a, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    a[t] = 0.0;
}


__global__ __launch_bounds__(256) void k_adj_lambda(KP kp, const double* __restrict__ lam, const double* __restrict__ J,
                                                    double scale, double* __restrict__ Lam) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= kp.Nxp || j >= kp.Nyp) return;
  const long long o = cidx(kp, i + kp.NG, j + kp.NG, k + kp.NG);
  const double s = scale * J[o];
#pragma unroll
  for (int n = 0; n < 5; ++n) Lam[o + n * kp.nc] = s * lam[o + n * kp.nc];
}

// Faces whose largest sensor class over cells c-3..c+2 (c = right cell) is not UP7, the three directions in
// one launch: a lane takes the x, y and z faces whose right cell is its cell and appends them to its wave's
// run of each direction.
__global__ __launch_bounds__(256) void k_adj_heavy_list(KP kp, const unsigned char* __restrict__ cls,
                                                        unsigned long long* __restrict__ list,
                                                        unsigned int* __restrict__ count, HList hl,
                                                        unsigned int* __restrict__ hctr, const unsigned int* pflag,
                                                        unsigned int pgen) {
  // lanes run over the (Nxp+1) x (Nyp+1) face-index positions of plane k in order (a wave's 64 positions
  // may span two rows), so that no wave is left mostly idle at the end of a row
  const long long t = (long long)(blockIdx.x * 4 + threadIdx.y) * 64 + threadIdx.x;
  const int i = (int)(t % (kp.Nxp + 1)), j = (int)(t / (kp.Nxp + 1)), k = blockIdx.z;
  const unsigned int w = launch_wave(4, threadIdx.y);
  // no sensor value of this state left the UP7 band: no face is listed, and the kernels that walk the list
  // (k_adj_heavy_form, k_adj_heavy_rest) check the same flag and return at once
  if (!heavy_possible(pflag, pgen)) return;
  if (blockIdx.x == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0) *hctr = 0u;
#pragma unroll
  for (int D = 0; D < 3; ++D) {
    const bool act = i < kp.Nxp + (D == 0) && j < kp.Nyp + (D == 1) && k < kp.Nzp + (D == 2);
    bool heavy = false;
    unsigned long long entry = 0;
    if (act) {
      const long long o = cidx(kp, i + kp.NG, j + kp.NG, k + kp.NG), st = dstride(kp, D);
      int sch = cls[o - 3 * st];
#pragma unroll
      for (int s = -2; s <= 2; ++s) sch = max(sch, (int)cls[o + s * st]);
      heavy = sch != F3D_SCHEME_UP7;
      entry = ((unsigned long long)o << 4) | ((unsigned long long)sch << 2) | (unsigned long long)D;
    }
    unsigned int cnt = 0;
    run_append(list, hl.d[D].base + (unsigned long long)w * hl.d[D].cap, cnt, heavy, entry);
    if (threadIdx.x == 0) count[hl.d[D].first + w] = cnt;
  }
}

// HRec: the listed faces' contributions of one RHS VJP (workspace arrays, this evaluation's stamp).
struct HRec {
  const unsigned int* pflag;  // sensor band flag of the state and its generation (heavy_possible)
  unsigned int pgen;
  double* buf;
  unsigned int* rec;     // 3 x nc, direction-major
  unsigned char* flag;   // per cell: the stamp when some kept face's stencil covers the cell
  unsigned int* ctr;
  unsigned int cap, stamp;  // stamp in bits 24..31
  __device__ bool covered(long long o) const { return flag[o] == (unsigned char)(stamp >> 24); }
};

// Compact index of the record of face (D, right cell o), or -1 when this evaluation kept none for it.
__device__ __forceinline__ int hrec_slot(const HRec& hr, unsigned int e) {
  return (e & 0xFF000000u) == hr.stamp ? (int)(e & 0x00FFFFFFu) - 1 : -1;
}

// Adds to g[0..4] the kept contributions of the listed faces of one direction around a cell, in the order of
// their right cells' offsets m = -3..4 along the direction (the cell is stencil cell 4 - m of each):
// RW(m) = the record of the face whose right cell is at offset m.
template <class RW>
__device__ __forceinline__ void adj_heavy_add(const HRec& hr, const RW& rw, double* g) {
#pragma unroll 1
  for (int m = -3; m <= 4; ++m) {
    const int cs = hrec_slot(hr, rw(m));
    if (cs >= 0) {
      const double* c = hr.buf + 40ULL * (unsigned int)cs + 5 * (4 - m);
#pragma unroll
      for (int q = 0; q < 5; ++q) g[q] += c[q];
    }
  }
}

// ------------------------------------------------------------------ inviscid, dense part
// lamQ(rho,u,v,w,T) of every cell from the UP7 form of every face: for direction d the RHS is
//   R[c] = J_c * (sum_r da[r+4] Fp[c+r] + sum_r db[r+3] Fm[c+r]),
// so dFp[e] = sum_{r=-3..4} da[4-r] Lam[e+r] and dFm[e] = sum_{r=-4..3} db[3-r] Lam[e+r].
__global__ __launch_bounds__(256) void k_adj_inviscid(KP kp, const double* __restrict__ Q, MetricPtrs<double> mp,
                                                      const double* __restrict__ Lam, double* __restrict__ lamQ,
                                                      HRec hr) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= kp.Nxt || j >= kp.Nyt) return;
  const long long o = cidx(kp, i, j, k), nc = kp.nc;
  const int NG = kp.NG;
  const bool inx = interior(i, kp.Nxp, NG), iny = interior(j, kp.Nyp, NG), inz = interior(k, kp.Nzp, NG);
  double acc[5] = {0, 0, 0, 0, 0};
  const bool slab[3] = {iny && inz, inx && inz, inx && iny};
  if (slab[0] || slab[1] || slab[2]) {
    const double rho = Q[o], u = Q[o + nc], v = Q[o + 2 * nc], w = Q[o + 3 * nc], Tt = Q[o + 5 * nc];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (!slab[d]) continue;
      const int a = d == 0 ? i : (d == 1 ? j : k);
      const int Nt = d == 0 ? kp.Nxt : (d == 1 ? kp.Nyt : kp.Nzt);
      const long long st = dstride(kp, d);
      double gp[5] = {0, 0, 0, 0, 0}, gm[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int r = -4; r <= 4; ++r) {
        if (a + r < 0 || a + r >= Nt) continue;
        const double ca = r >= -3 ? kp.da[4 - r] : 0.0;
        const double cb = r <= 3 ? kp.db[3 - r] : 0.0;
        const long long oo = o + r * st;
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          const double L = Lam[oo + n * nc];
          gp[n] += ca * L;
          gm[n] += cb * L;
        }
      }
      const double A1 = mp.m[3 * d][o], A2 = mp.m[3 * d + 1][o], A3 = mp.m[3 * d + 2][o];
      const double S = sqrt(A1 * A1 + A2 * A2 + A3 * A3);
      double g[5];
      sw_split_vjp(kp, rho, u, v, w, Tt, A1, A2, A3, S, gp, gm, g);
#pragma unroll
      for (int n = 0; n < 5; ++n) acc[n] += g[n];
      if (hr.covered(o)) {
        const unsigned int* rd = hr.rec + (long long)d * nc + o;
        adj_heavy_add(hr, [&](int m) { return a + m >= 0 && a + m < Nt ? rd[m * st] : 0u; }, acc);
      }
    }
  }
  lamQ[o] = acc[0];
  lamQ[o + nc] = acc[1];
  lamQ[o + 2 * nc] = acc[2];
  lamQ[o + 3 * nc] = acc[3];
  lamQ[o + 4 * nc] = 0.0;
  lamQ[o + 5 * nc] = acc[4];
}

// The same VJP as two marches (a direction's share of a cell needs Lam at the 8 cells around it along that
// direction, which k_adj_inviscid gathers from the cache per cell and direction):
//   k_adj_inv_z  lane = (i, j) column, marching k with Lam of the column in a 9-plane register window;
//                writes lamQ of every cell: the z share where i and j are interior, else zero;
//   k_adj_inv_xy lane = i in a 64-cell window (4 halo cells each side, 56 cells written), threadIdx.y = an
//                interior k line, marching j with Lam of the column in a 9-row register window; the row's Lam
//                is exchanged through LDS for the x share; adds the x share (interior j) and the y share
//                (interior i) to lamQ of the cells with interior k.
// Each Lam value is read from HBM once per march.  Same sums, in the order z, x, y per cell.
constexpr int kAIL = 2;   // k lines per block of the x+y march
constexpr int kAIW = 64;  // lanes of a row window
constexpr int kAII = kAIW - 8;

// dFp, dFm of a cell from Lam at the nine cells around it along one direction: LW(r, n) = Lam of variable n
// at offset r - 4 (the sums of k_adj_inviscid, in its order).
template <class F>
__device__ __forceinline__ void adj_gather(const KP& kp, const F& LW, double* gp, double* gm) {
#pragma unroll
  for (int n = 0; n < 5; ++n) gp[n] = gm[n] = 0;
#pragma unroll
  for (int r = -4; r <= 4; ++r) {
    const double ca = r >= -3 ? kp.da[4 - r] : 0.0;
    const double cb = r <= 3 ? kp.db[3 - r] : 0.0;
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      const double l = LW(r + 4, n);
      gp[n] += ca * l;
      gm[n] += cb * l;
    }
  }
}

__global__ __launch_bounds__(256) void k_adj_inv_z(KP kp, const double* __restrict__ Q, MetricPtrs<double> mp,
                                                   const double* __restrict__ Lam, double* __restrict__ lamQ, int L,
                                                   HRec hr) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= kp.Nxt || j >= kp.Nyt) return;
  const int NG = kp.NG, k0 = blockIdx.z * L, k1 = min(k0 + L, kp.Nzt);
  const long long nc = kp.nc, sz = (long long)kp.Nxt * kp.Nyt, col = cidx(kp, i, j, 0);
  const bool zcol = interior(i, kp.Nxp, NG) && interior(j, kp.Nyp, NG);
  double Lw[9][5];
  auto load = [&](int k, double* d) {
    const bool in = zcol && k >= 0 && k < kp.Nzt;
#pragma unroll
    for (int n = 0; n < 5; ++n) d[n] = in ? Lam[col + (long long)k * sz + n * nc] : 0.0;
  };
  // records of the listed faces (read where a kept face's stencil covers the cell)
  const unsigned int* rz = hr.rec + 2 * nc + col;
  const unsigned int* ry = hr.rec + nc + col;
  const unsigned int* rx = hr.rec + col;
  const unsigned char f8 = (unsigned char)(hr.stamp >> 24);
  unsigned char fnext = k0 < k1 ? hr.flag[col + (long long)k0 * sz] : 0;  // coverage flag, a step ahead
#pragma unroll
  for (int r = 0; r < 8; ++r) load(k0 - 4 + r, Lw[r]);
  // the cell's state and zeta metrics, requested a step ahead (their latency overlaps the step before)
  double nq[5] = {0, 0, 0, 0, 0}, na[3] = {0, 0, 0};
  auto load_cell = [&](long long oc) {
    nq[0] = Q[oc]; nq[1] = Q[oc + nc]; nq[2] = Q[oc + 2 * nc]; nq[3] = Q[oc + 3 * nc]; nq[4] = Q[oc + 5 * nc];
    na[0] = mp.m[6][oc]; na[1] = mp.m[7][oc]; na[2] = mp.m[8][oc];
  };
  if (zcol && k0 < k1) load_cell(col + (long long)k0 * sz);
#pragma unroll 1
  for (int k = k0; k < k1; ++k) {
    load(k + 4, Lw[8]);
    const long long o = col + (long long)k * sz;
    const bool covered = fnext == f8;
    if (k + 1 < k1) fnext = hr.flag[o + sz];
    const double q0 = nq[0], q1 = nq[1], q2 = nq[2], q3 = nq[3], q5 = nq[4], A1 = na[0], A2 = na[1], A3 = na[2];
    if (zcol && k + 1 < k1) load_cell(o + sz);
    double g[5] = {0, 0, 0, 0, 0};
    if (zcol) {
      double gp[5], gm[5];
      adj_gather(kp, [&](int r, int n) { return Lw[r][n]; }, gp, gm);
      sw_split_vjp(kp, q0, q1, q2, q3, q5, A1, A2, A3, sqrt(A1 * A1 + A2 * A2 + A3 * A3), gp, gm, g);
    }
    // the listed faces' kept contributions to this cell, in the order z, x, y (k_adj_inv_xy adds the dense
    // x and y shares afterwards)
    if (covered) {
      const long long ok = (long long)k * sz;
      adj_heavy_add(hr, [&](int m) { return k + m >= 0 && k + m < kp.Nzt ? rz[(long long)(k + m) * sz] : 0u; }, g);
      adj_heavy_add(hr, [&](int m) { return i + m >= 0 && i + m < kp.Nxt ? rx[ok + m] : 0u; }, g);
      adj_heavy_add(hr, [&](int m) { return j + m >= 0 && j + m < kp.Nyt ? ry[ok + (long long)m * kp.Nxt] : 0u; }, g);
    }
    lamQ[o] = g[0];
    lamQ[o + nc] = g[1];
    lamQ[o + 2 * nc] = g[2];
    lamQ[o + 3 * nc] = g[3];
    lamQ[o + 4 * nc] = 0.0;
    lamQ[o + 5 * nc] = g[4];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int n = 0; n < 5; ++n) Lw[r][n] = Lw[r + 1][n];
  }
}

__global__ __launch_bounds__(kAIW * kAIL) void k_adj_inv_xy(KP kp, const double* __restrict__ Q, MetricPtrs<double> mp,
                                                            const double* __restrict__ Lam, double* __restrict__ lamQ,
                                                            int L) {
  __shared__ double sL[5][kAIL][kAIW];
  const int t = threadIdx.x, ln = threadIdx.y, NG = kp.NG;
  const int i = blockIdx.x * kAII - 4 + t;            // the lane's cell (may lie outside [0, Nxt))
  const bool iok = i >= 0 && i < kp.Nxt;
  const bool out = t >= 4 && t < 4 + kAII && iok;     // lamQ of this lane's cells is updated
  const bool iin = interior(i, kp.Nxp, NG);
  const int kk = blockIdx.y * kAIL + ln;
  const bool kok = kk < kp.Nzp;
  const int k = NG + (kok ? kk : kp.Nzp - 1);
  const int j0 = blockIdx.z * L, j1 = min(j0 + L, kp.Nyt);
  const long long nc = kp.nc, sy = kp.Nxt;
  const long long row0 = cidx(kp, iok ? i : 0, 0, k);
  double Lw[9][5];
  // Lam of row j of this lane's column (zero outside the array: Lam is zero on ghost cells anyway)
  auto load = [&](int j, double* d) {
    const bool in = iok && kok && j >= 0 && j < kp.Nyt;
#pragma unroll
    for (int n = 0; n < 5; ++n) d[n] = in ? Lam[row0 + (long long)j * sy + n * nc] : 0.0;
  };
#pragma unroll
  for (int r = 0; r < 8; ++r) load(j0 - 4 + r, Lw[r]);
#pragma unroll 1
  for (int j = j0; j < j1; ++j) {
    load(j + 4, Lw[8]);
    const bool jin = interior(j, kp.Nyp, NG);
#pragma unroll
    for (int n = 0; n < 5; ++n) sL[n][ln][t] = Lw[4][n];
    // the cell's operands are requested before the barrier, so that their latency overlaps the wait
    const bool upd = out && kok && (jin || iin);
    const long long o = row0 + (long long)j * sy;
    double rho = 0, u = 0, v = 0, w = 0, Tt = 0, acc[5] = {0, 0, 0, 0, 0}, X[3] = {0, 0, 0}, Y[3] = {0, 0, 0};
    if (upd) {
      rho = Q[o]; u = Q[o + nc]; v = Q[o + 2 * nc]; w = Q[o + 3 * nc]; Tt = Q[o + 5 * nc];
      acc[0] = lamQ[o]; acc[1] = lamQ[o + nc]; acc[2] = lamQ[o + 2 * nc]; acc[3] = lamQ[o + 3 * nc];
      acc[4] = lamQ[o + 5 * nc];
      if (jin) { X[0] = mp.m[0][o]; X[1] = mp.m[1][o]; X[2] = mp.m[2][o]; }
      if (iin) { Y[0] = mp.m[3][o]; Y[1] = mp.m[4][o]; Y[2] = mp.m[5][o]; }
    }
    __syncthreads();
    if (upd) {
      if (jin) {  // x share: Lam of the row from the neighbouring lanes
        double gp[5], gm[5], g[5];
        adj_gather(kp, [&](int r, int n) { return sL[n][ln][t - 4 + r]; }, gp, gm);
        const double A1 = X[0], A2 = X[1], A3 = X[2];
        sw_split_vjp(kp, rho, u, v, w, Tt, A1, A2, A3, sqrt(A1 * A1 + A2 * A2 + A3 * A3), gp, gm, g);
#pragma unroll
        for (int n = 0; n < 5; ++n) acc[n] += g[n];
      }
      if (iin) {  // y share from the column window
        double gp[5], gm[5], g[5];
        adj_gather(kp, [&](int r, int n) { return Lw[r][n]; }, gp, gm);
        const double A1 = Y[0], A2 = Y[1], A3 = Y[2];
        sw_split_vjp(kp, rho, u, v, w, Tt, A1, A2, A3, sqrt(A1 * A1 + A2 * A2 + A3 * A3), gp, gm, g);
#pragma unroll
        for (int n = 0; n < 5; ++n) acc[n] += g[n];
      }
      lamQ[o] = acc[0];
      lamQ[o + nc] = acc[1];
      lamQ[o + 2 * nc] = acc[2];
      lamQ[o + 3 * nc] = acc[3];
      lamQ[o + 5 * nc] = acc[4];
    }
    __syncthreads();  // the row in LDS is read before the next row overwrites it
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int n = 0; n < 5; ++n) Lw[r][n] = Lw[r + 1][n];
  }
}

// ------------------------------------------------------------------ inviscid, listed faces
// Adjoint of a listed face's flux, gF = Lam[o] - Lam[o - st] (the face is the left face of o, the right one
// of o - st); false when it vanishes.
__device__ __forceinline__ bool adj_face_seed(const KP& kp, const double* __restrict__ Lam, long long o, long long st,
                                              double* gF) {
  bool any = false;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    gF[q] = Lam[o + q * kp.nc] - Lam[o - st + q * kp.nc];
    any |= gF[q] != 0.0;
  }
  return any;
}

// Reverse of one listed face's (scheme - UP7) flux from its adjoint gF, cell by cell: emit(s, g) receives the
// contribution g[0..4] to (rho, u, v, w, T) of stencil cell s (offset s - 4 along D from the right cell o), the
// frame's dependence on (u, v, w, T) of cells 3 and 4 included in theirs.
template <class Emit>
__device__ __forceinline__ void adj_face_reverse(const KP& kp, const double* __restrict__ Q, const MetricPtrs<double>& mp,
                                                 int D, int sch, long long o, const double* gF, Emit&& emit) {
  const long long nc = kp.nc, st = dstride(kp, D);
  const double* A1p = mp.m[3 * D];
  const double* A2p = mp.m[3 * D + 1];
  const double* A3p = mp.m[3 * D + 2];
  double fp[8][5], fm[8][5], gfp[8][5], gfm[8][5];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const long long oc = o + (s - 4) * st;
    const double A1 = A1p[oc], A2 = A2p[oc], A3 = A3p[oc];
    sw_split<double>(kp, Q[oc], Q[oc + nc], Q[oc + 2 * nc], Q[oc + 3 * nc], Q[oc + 5 * nc], A1, A2, A3,
                     sqrt(A1 * A1 + A2 * A2 + A3 * A3), fp[s], fm[s]);
#pragma unroll
    for (int q = 0; q < 5; ++q) gfp[s][q] = gfm[s][q] = 0.0;
  }
  const long long oL = o - st;
  const double qL[4] = {Q[oL + nc], Q[oL + 2 * nc], Q[oL + 3 * nc], Q[oL + 5 * nc]};
  const double qR[4] = {Q[o + nc], Q[o + 2 * nc], Q[o + 3 * nc], Q[o + 5 * nc]};
  const double AL[3] = {A1p[oL], A2p[oL], A3p[oL]};
  const double AR[3] = {A1p[o], A2p[o], A3p[o]};
  const double SL = sqrt(AL[0] * AL[0] + AL[1] * AL[1] + AL[2] * AL[2]);
  const double SR = sqrt(AR[0] * AR[0] + AR[1] * AR[1] + AR[2] * AR[2]);
  double gqL[4] = {0, 0, 0, 0}, gqR[4] = {0, 0, 0, 0};
  face_flux_vjp(kp, sch, fp, fm, qL, qR, AL, AR, SL, SR, gF, gfp, gfm, gqL, gqR);
  // the dense kernels already took this face as UP7
#pragma unroll
  for (int q = 0; q < 7; ++q)
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      gfp[q][m] -= kp.lin[q] * gF[m];
      gfm[7 - q][m] -= kp.lin[q] * gF[m];
    }
#pragma unroll 1
  for (int s = 0; s < 8; ++s) {
    const long long oc = o + (s - 4) * st;
    const double A1 = A1p[oc], A2 = A2p[oc], A3 = A3p[oc];
    double g[5];
    sw_split_vjp(kp, Q[oc], Q[oc + nc], Q[oc + 2 * nc], Q[oc + 3 * nc], Q[oc + 5 * nc], A1, A2, A3,
                 sqrt(A1 * A1 + A2 * A2 + A3 * A3), gfp[s], gfm[s], g);
    if (s == 3) {
#pragma unroll
      for (int q = 0; q < 4; ++q) g[q + 1] += gqL[q];
    } else if (s == 4) {
#pragma unroll
      for (int q = 0; q < 4; ++q) g[q + 1] += gqR[q];
    }
    emit(s, g);
  }
}

// Per listed face: the reverse contributions to its eight stencil cells kept in the face's compact slot (a
// wave takes its slots with one counter update) and the slot recorded under the face's right cell; the dense
// inviscid kernels then add them cell by cell in a fixed order.  Faces beyond the buffer's capacity are left
// to k_adj_heavy_rest.
__global__ __launch_bounds__(64) void k_adj_heavy_form(KP kp, const double* __restrict__ Q, MetricPtrs<double> mp,
                                                       const double* __restrict__ Lam,
                                                       const unsigned long long* __restrict__ list,
                                                       const unsigned int* __restrict__ count, HList hl, HRec hr) {
  if (!heavy_possible(hr.pflag, hr.pgen)) return;  // nothing listed
  for_each_heavy(hl, count, list, [&](const unsigned long long v) {
    const int D = (int)(v & 3), sch = (int)((v >> 2) & 3);
    const long long o = (long long)(v >> 4), st = dstride(kp, D);
    double gF[5];
    const bool any = adj_face_seed(kp, Lam, o, st, gF);
    const unsigned long long m = __ballot(any);
    if (!m) return;
    const int lane = __lane_id(), leader = __ffsll((unsigned long long)m) - 1;
    unsigned int b = 0;
    if (lane == leader) b = atomicAdd(hr.ctr, (unsigned int)__popcll(m));
    b = __shfl(b, leader);
    if (!any) return;
    const unsigned int cs = b + (unsigned int)__popcll(m & ((1ULL << lane) - 1ULL));
    if (cs >= hr.cap) return;
    double* dst = hr.buf + 40ULL * cs;
    const unsigned char f8 = (unsigned char)(hr.stamp >> 24);
    adj_face_reverse(kp, Q, mp, D, sch, o, gF, [&](int s, const double* g) {
#pragma unroll
      for (int q = 0; q < 5; ++q) dst[5 * s + q] = g[q];
      hr.flag[o + (s - 4) * st] = f8;
    });
    hr.rec[(long long)D * kp.nc + o] = hr.stamp | (cs + 1);
  });
}

// Faces the buffer could not keep (only when more than `cap` listed faces have an adjoint): their
// contributions added to lamQ with atomics after the dense kernels (summation order not fixed).
__global__ __launch_bounds__(64) void k_adj_heavy_rest(KP kp, const double* __restrict__ Q, MetricPtrs<double> mp,
                                                       const double* __restrict__ Lam,
                                                       const unsigned long long* __restrict__ list,
                                                       const unsigned int* __restrict__ count, double* __restrict__ lamQ,
                                                       HList hl, HRec hr) {
  if (!heavy_possible(hr.pflag, hr.pgen) || *hr.ctr <= hr.cap) return;
  const long long nc = kp.nc;
  for_each_heavy(hl, count, list, [&](const unsigned long long v) {
    const int D = (int)(v & 3), sch = (int)((v >> 2) & 3);
    const long long o = (long long)(v >> 4), st = dstride(kp, D);
    if (hrec_slot(hr, hr.rec[(long long)D * nc + o]) >= 0) return;
    double gF[5];
    if (!adj_face_seed(kp, Lam, o, st, gF)) return;
    adj_face_reverse(kp, Q, mp, D, sch, o, gF, [&](int s, const double* g) {
      const long long oc = o + (s - 4) * st;
      atomicAdd(&lamQ[oc], g[0]);
      atomicAdd(&lamQ[oc + nc], g[1]);
      atomicAdd(&lamQ[oc + 2 * nc], g[2]);
      atomicAdd(&lamQ[oc + 3 * nc], g[3]);
      atomicAdd(&lamQ[oc + 5 * nc], g[4]);
    });
  });
}

// ------------------------------------------------------------------ viscous
__device__ __forceinline__ bool in_box(const KP& kp, int i, int j, int k) {
  return i >= 2 && i <= kp.Nxt - 3 && j >= 2 && j <= kp.Nyt - 3 && k >= 2 && k <= kp.Nzt - 3;
}

// Per cell of the interior+-2 box: the adjoint of its viscous fluxes (from the transposed
// 4th-order divergence of Lam) taken back through viscous_cell.  G[3v+d] = adjoint of the
// d-derivative of (u,v,w,T)[v]; G[12+v] = adjoint of (u,v,w,T)[v] at the cell.
__global__ __launch_bounds__(256) void k_adj_visc1(KP kp, const double* __restrict__ Q, MetricPtrs<double> mp,
                                                   const double* __restrict__ J, const double* __restrict__ Lam,
                                                   double* __restrict__ G) {
  const int i = 2 + blockIdx.x * 64 + threadIdx.x, j = 2 + blockIdx.y * 4 + threadIdx.y, k = 2 + blockIdx.z;
  if (i > kp.Nxt - 3 || j > kp.Nyt - 3) return;
  const long long o = cidx(kp, i, j, k), nc = kp.nc;
  const int NG = kp.NG;
  const bool inx = interior(i, kp.Nxp, NG), iny = interior(j, kp.Nyp, NG), inz = interior(k, kp.Nzp, NG);
  const bool slab[3] = {iny && inz, inx && inz, inx && iny};
  if (!(slab[0] || slab[1] || slab[2])) {
#pragma unroll
    for (int q = 0; q < 16; ++q) G[o + q * nc] = 0.0;
    return;
  }
  const double c1 = 1.0 / 12.0, c2 = -2.0 / 3.0;
  double go[3][4];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const long long st = dstride(kp, d);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      if (!slab[d]) {
        go[d][n] = 0.0;
        continue;
      }
      const double* L = Lam + (n + 1) * nc + o;
      go[d][n] = c1 * (L[2 * st] - L[-2 * st]) + c2 * (L[st] - L[-st]);
    }
  }
  const int var[4] = {1, 2, 3, 5};
  double g[4][3];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double* f = Q + var[v] * nc + o;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const long long st = dstride(kp, d);
      g[v][d] = c1 * (f[-2 * st] - f[2 * st]) + c2 * (f[-st] - f[st]);
    }
  }
  double m[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) m[q] = mp.m[q][o];
  double gg[4][3], gval[4];
  viscous_cell_vjp(kp, Q[o + nc], Q[o + 2 * nc], Q[o + 3 * nc], Q[o + 5 * nc], g, m, J[o], go, gg, gval);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
#pragma unroll
    for (int d = 0; d < 3; ++d) G[o + (3 * v + d) * nc] = gg[v][d];
    G[o + (12 + v) * nc] = gval[v];
  }
}

// lamQ(u,v,w,T) += own G[12+v] + transposed derivative stencils of G[3v+d] (box cells only).
__global__ __launch_bounds__(256) void k_adj_visc2(KP kp, const double* __restrict__ G, double* __restrict__ lamQ) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= kp.Nxt || j >= kp.Nyt) return;
  const long long o = cidx(kp, i, j, k), nc = kp.nc;
  const double c1 = 1.0 / 12.0, c2 = -2.0 / 3.0;
  const int var[4] = {1, 2, 3, 5};
  const bool self = in_box(kp, i, j, k);
  double acc[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) acc[v] = self ? G[o + (12 + v) * nc] : 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const long long st = dstride(kp, d);
#pragma unroll
    for (int r = -2; r <= 2; ++r) {
      if (r == 0) continue;
      const int ii = i + (d == 0 ? r : 0), jj = j + (d == 1 ? r : 0), kk = k + (d == 2 ? r : 0);
      if (!in_box(kp, ii, jj, kk)) continue;
      // d(g[c])/d(f[a]) for c = a + r:  r=+2 -> c1, r=-2 -> -c1, r=+1 -> c2, r=-1 -> -c2
      const double cf = r == 2 ? c1 : (r == -2 ? -c1 : (r == 1 ? c2 : -c2));
      const long long oo = o + r * st;
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] += cf * G[oo + (3 * v + d) * nc];
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) lamQ[o + var[v] * nc] += acc[v];
}

// k_adj_visc1 + k_adj_visc2 in one k-march: the 16 partial adjoints G of a cell never reach HBM.  A block
// holds a tile of 16 rows of 32 lanes (two rows per wave), the outer 2 lanes and rows a halo, and marches
// the planes of its segment: per plane p every lane forms G at its cell (u, v, w, T and Lam of the column
// in 5-plane register rings for the zeta differences, the xi/eta neighbours from L1/L2), the xi and eta
// parts go through LDS to the tile's inner cells, which add the transposed xi/eta stencils and their own
// G[12+v] to lamQ at plane p; the zeta parts stay in a 5-plane ring and reach lamQ at plane p-2.  G is
// zero outside the interior+-2 box, as k_adj_visc1 leaves it, so every cell of lamQ gets what
// k_adj_visc2 gathers.
constexpr int kAVW = 32, kAVR = 16;  // row width, rows (inner: kAVW-4 columns, kAVR-4 rows)
__global__ __launch_bounds__(512) void k_adj_visc_march(KP kp, const double* __restrict__ Q, MetricPtrs<double> mp,
                                                        const double* __restrict__ J, const double* __restrict__ Lam,
                                                        double* __restrict__ lamQ, int KC) {
  __shared__ double sG[8][kAVR][kAVW];            // G[3v] (xi) then G[3v+1] (eta) of plane p
  __shared__ double sLQ[8][kAVR + 4][kAVW + 4];   // Lam (vars 1..4) then u,v,w,T of plane p, tile +-2
  const int lane = threadIdx.x, col = lane & (kAVW - 1), row = 2 * threadIdx.y + (lane >> 5);
  const int x = blockIdx.x * (kAVW - 4) - 2 + col, y = blockIdx.y * (kAVR - 4) - 2 + row;
  const int kz0 = blockIdx.z * KC, kz1 = min(kz0 + KC, kp.Nzt);  // lamQ planes of this block: [kz0, kz1)
  const bool act = x >= 0 && x < kp.Nxt && y >= 0 && y < kp.Nyt;
  const bool outl = act && col >= 2 && col < kAVW - 2 && row >= 2 && row < kAVR - 2;
  const bool xybox = x >= 2 && x <= kp.Nxt - 3 && y >= 2 && y <= kp.Nyt - 3;
  const int NG = kp.NG;
  const bool inx = interior(x, kp.Nxp, NG), iny = interior(y, kp.Nyp, NG);
  const long long nc = kp.nc, sy = kp.Nxt, sz = (long long)kp.Nxt * kp.Nyt;
  auto cl = [](int a, int n) { return min(max(a, 0), n - 1); };
  const long long colo = (long long)cl(x, kp.Nxt) + sy * cl(y, kp.Nyt);
  // the extra cell an edge lane (x -+ 2) or a halo row (y -+ 2) stages for the tile's +-2 neighbours
  const bool xe = col < 2 || col >= kAVW - 2, ye = row < 2 || row >= kAVR - 2;
  const long long colx = (long long)cl(col < 2 ? x - 2 : x + 2, kp.Nxt) + sy * cl(y, kp.Nyt);
  const long long coly = (long long)cl(x, kp.Nxt) + sy * cl(row < 2 ? y - 2 : y + 2, kp.Nyt);
  const int tcx = col < 2 ? col : col + 4, trx = row + 2;  // tile position of the x-extra cell
  const int tcy = col + 2, try_ = row < 2 ? row : row + 4;  // and of the y-extra cell
  const int tc = col + 2, tr = row + 2;
  const double c1 = 1.0 / 12.0, c2 = -2.0 / 3.0;
  const int var[4] = {1, 2, 3, 5};
  // rings: slot s holds plane p-2+s of Lam (vars 1..4) and of u,v,w,T (slot 5: plane p+3, requested a
  // step ahead); gz slot s the zeta part of G at plane p-4+s
  double L[4][6], F[4][6], gz[4][5];
  // the xi / eta part (with the own G[12+v]) of planes p-2 and p-1, added to lamQ together with the zeta
  // part of the same plane two iterations after it is formed: one read-modify-write of lamQ per cell
  double pend[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  auto load = [&](int pl, int sl) {
    const long long o = colo + (long long)cl(pl, kp.Nzt) * sz;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      L[v][sl] = Lam[o + (v + 1) * nc];
      F[v][sl] = Q[o + var[v] * nc];
    }
  };
  const int p0 = kz0 - 2;
#pragma unroll
  for (int sl = 0; sl < 5; ++sl) load(p0 - 2 + sl, sl);
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int sl = 0; sl < 5; ++sl) gz[v][sl] = 0.0;
#pragma unroll 1
  for (int p = p0; p <= kz1 + 1; ++p) {
    load(p + 3, 5);
    const long long o = colo + (long long)cl(p, kp.Nzt) * sz;
    // stage plane p of Lam and u,v,w,T in the tile (own cell, and the edge / halo extra cells)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      sLQ[v][tr][tc] = L[v][2];
      sLQ[4 + v][tr][tc] = F[v][2];
    }
    if (xe) {
      const long long ox = colx + (long long)cl(p, kp.Nzt) * sz;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        sLQ[v][trx][tcx] = Lam[ox + (v + 1) * nc];
        sLQ[4 + v][trx][tcx] = Q[ox + var[v] * nc];
      }
    }
    if (ye) {
      const long long oy = coly + (long long)cl(p, kp.Nzt) * sz;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        sLQ[v][try_][tcy] = Lam[oy + (v + 1) * nc];
        sLQ[4 + v][try_][tcy] = Q[oy + var[v] * nc];
      }
    }
    __syncthreads();
    double Gx[4] = {0, 0, 0, 0}, Gy[4] = {0, 0, 0, 0}, Gs[4] = {0, 0, 0, 0};
    gz[0][4] = gz[1][4] = gz[2][4] = gz[3][4] = 0.0;
    const bool inz = interior(p, kp.Nzp, NG);
    const bool slab[3] = {iny && inz, inx && inz, inx && iny};
    if (act && xybox && p >= 2 && p <= kp.Nzt - 3 && (slab[0] || slab[1] || slab[2])) {
      double go[3][4], g[4][3];
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        go[0][n] = slab[0] ? c1 * (sLQ[n][tr][tc + 2] - sLQ[n][tr][tc - 2]) +
                                 c2 * (sLQ[n][tr][tc + 1] - sLQ[n][tr][tc - 1])
                           : 0.0;
        go[1][n] = slab[1] ? c1 * (sLQ[n][tr + 2][tc] - sLQ[n][tr - 2][tc]) +
                                 c2 * (sLQ[n][tr + 1][tc] - sLQ[n][tr - 1][tc])
                           : 0.0;
        go[2][n] = slab[2] ? c1 * (L[n][4] - L[n][0]) + c2 * (L[n][3] - L[n][1]) : 0.0;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        g[v][0] = c1 * (sLQ[4 + v][tr][tc - 2] - sLQ[4 + v][tr][tc + 2]) +
                  c2 * (sLQ[4 + v][tr][tc - 1] - sLQ[4 + v][tr][tc + 1]);
        g[v][1] = c1 * (sLQ[4 + v][tr - 2][tc] - sLQ[4 + v][tr + 2][tc]) +
                  c2 * (sLQ[4 + v][tr - 1][tc] - sLQ[4 + v][tr + 1][tc]);
        g[v][2] = c1 * (F[v][0] - F[v][4]) + c2 * (F[v][1] - F[v][3]);
      }
      double m[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) m[q] = mp.m[q][o];
      double gg[4][3];
      viscous_cell_vjp(kp, F[0][2], F[1][2], F[2][2], F[3][2], g, m, J[o], go, gg, Gs);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        Gx[v] = gg[v][0];
        Gy[v] = gg[v][1];
        gz[v][4] = gg[v][2];
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      sG[v][row][col] = Gx[v];
      sG[4 + v][row][col] = Gy[v];
    }
    __syncthreads();
    // plane p: own G[12+v] + transposed xi / eta stencils (d g[c] / d f[a] for c = a + r: r = +2 -> c1,
    // -2 -> -c1, +1 -> c2, -1 -> -c2), kept two iterations
    double accp[4] = {0, 0, 0, 0};
    if (outl && p >= kz0 && p < kz1) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double acc = Gs[v];
        acc += c1 * (sG[v][row][col + 2] - sG[v][row][col - 2]) + c2 * (sG[v][row][col + 1] - sG[v][row][col - 1]);
        acc += c1 * (sG[4 + v][row + 2][col] - sG[4 + v][row - 2][col]) +
               c2 * (sG[4 + v][row + 1][col] - sG[4 + v][row - 1][col]);
        accp[v] = acc;
      }
    }
    // plane p-2: its xi / eta part plus the transposed zeta stencil of the ring
    const int q = p - 2;
    if (outl && q >= kz0 && q < kz1) {
      const long long oq = colo + (long long)q * sz;
#pragma unroll
      for (int v = 0; v < 4; ++v)
        lamQ[oq + var[v] * nc] += pend[v][0] + (c1 * (gz[v][4] - gz[v][0]) + c2 * (gz[v][3] - gz[v][1]));
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      pend[v][0] = pend[v][1];
      pend[v][1] = accp[v];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int sl = 0; sl < 5; ++sl) {
        L[v][sl] = L[v][sl + 1];
        F[v][sl] = F[v][sl + 1];
      }
#pragma unroll
      for (int sl = 0; sl < 4; ++sl) gz[v][sl] = gz[v][sl + 1];
    }
    __syncthreads();  // sG and sLQ are rewritten for the next plane
  }
}

// ------------------------------------------------------------------ ghost procedure, reverse
// Cell of the lamQ ghost slab of face (D, side) that holds partial n (0..6) of the block whose first line
// is (a0, b0): variables 0..5 of the line's first ghost cell, then variable 0 of its second.
template <int D>
__device__ __forceinline__ long long bc_slot(const KP& kp, int side, int a0, int b0, int n) {
  const int Np = D == 0 ? kp.Nxp : (D == 1 ? kp.Nyp : kp.Nzp);
  const int g = (side ? Np + kp.NG : 0) + (n < 6 ? 0 : 1);
  const long long o = D == 0 ? cidx(kp, g, a0, b0) : (D == 1 ? cidx(kp, a0, g, b0) : cidx(kp, a0, b0, g));
  return o + (n < 6 ? (long long)n * kp.nc : 0);
}

// lam_bc[n] += the partials of the nbx x nby blocks of the reverse fills of dimension D (the sides with
// red[side] set), summed in a fixed order (a strided sum per thread, then a fixed tree): the boundary-data
// gradient is the same bit for bit on every run.  The partial slots are cleared again (every ghost adjoint
// is zero after a reverse fill).
template <int D>
__global__ __launch_bounds__(256) void k_adj_bc_final(KP kp, int red0, int red1, int lw, int nbx, int nby,
                                                      double* __restrict__ lamQ, double* __restrict__ lam_bc) {
  // the seven components side by side; each one's partials are added in the same order as one at a time
  // (per thread: side 1's blocks, then side 0's, q ascending; then the same halving tree)
  __shared__ double sm[7][256];
  const int nb = nbx * nby;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int side = 1; side >= 0; --side) {
    if (!(side ? red1 : red0)) continue;
    for (int q = threadIdx.x; q < nb; q += 256) {
#pragma unroll
      for (int n = 0; n < 7; ++n) {
        const long long o = bc_slot<D>(kp, side, (q % nbx) * lw, (q / nbx) * 4, n);
        v[n] += lamQ[o];
        lamQ[o] = 0.0;
      }
    }
  }
#pragma unroll
  for (int n = 0; n < 7; ++n) sm[n][threadIdx.x] = v[n];
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int n = 0; n < 7; ++n) sm[n][threadIdx.x] += sm[n][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 7) lam_bc[threadIdx.x] += sm[threadIdx.x][0];
}

// Reverse of one side of one dimension's ghost fill (boundary.jl:1-156 with the kinds of f3d.h):
// each lane owns one line of NG ghost cells, folds their adjoints into the source cells and
// clears them.  lam_bc[0..5] += adjoint of q_inf, lam_bc[6] += adjoint of T_wall.
// Both sides in one launch when their sources cannot overlap (side = side0 + blockIdx.z, kind bc0 / bc1).
template <int D>
__global__ __launch_bounds__(256) void k_adj_ghost(KP kp, int side0, int bc0, int bc1, double* __restrict__ lamQ,
                                                   const double* __restrict__ Q, double* __restrict__ lam_inflow,
                                                   double* __restrict__ lam_bc) {
  const int side = side0 + (int)blockIdx.z, bc = side ? bc1 : bc0;
  const int a = blockIdx.x * 64 + threadIdx.x, b = blockIdx.y * 4 + threadIdx.y;
  // the other two dimensions, full extents
  const int na = D == 0 ? kp.Nyt : kp.Nxt, nb = D == 2 ? kp.Nyt : kp.Nzt;
  const bool act = a < na && b < nb;
  const int NG = kp.NG, Np = D == 0 ? kp.Nxp : (D == 1 ? kp.Nyp : kp.Nzp);
  const long long nc = kp.nc;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  if (act) {
    auto at = [&](int t) {
      return D == 0 ? cidx(kp, t, a, b) : (D == 1 ? cidx(kp, a, t, b) : cidx(kp, a, b, t));
    };
    for (int l = 0; l < NG; ++l) {
      const int g = side ? Np + NG + l : l;
      const long long o = at(g);
      double lg[6];
#pragma unroll
      for (int n = 0; n < 6; ++n) {
        lg[n] = lamQ[o + n * nc];
        lamQ[o + n * nc] = 0.0;
      }
      if (bc == F3D_BC_PERIODIC || bc == F3D_BC_ZEROGRAD) {
        const int t = bc == F3D_BC_PERIODIC ? (side ? g - Np : g + Np) : (side ? Np + NG - 1 : NG);
        const long long s = at(t);
#pragma unroll
        for (int n = 0; n < 6; ++n) lamQ[s + n * nc] += lg[n];
      } else if (bc == F3D_BC_DIRICHLET) {
#pragma unroll
        for (int n = 0; n < 6; ++n) acc[n] += lg[n];
      } else if (bc == F3D_BC_INFLOW) {
        if (lam_inflow) {
          const long long io = a + (long long)kp.Nyt * b, in = (long long)kp.Nyt * kp.Nzt;  // D == 0: (j, k)
#pragma unroll
          for (int n = 0; n < 6; ++n) lam_inflow[io + n * in] += lg[n];
        }
      } else {
        const int mi = side ? 2 * (Np + NG) - 1 - g : 2 * NG - 1 - g;
        const long long s = at(mi);
        double gr = lg[0], gpv = lg[4], gT = lg[5];
        if (bc == F3D_BC_WALL_ISOTHERMAL) {
          // ghost: T = 2 Tw - T_s, rho = p_s / (Rg T)
          const double q0 = Q[o], q5 = Q[o + 5 * nc];
          gpv += lg[0] / (kp.Rg * q5);
          const double gq5 = lg[5] - lg[0] * q0 / q5;
          gT = -gq5;
          acc[6] += 2 * gq5;
          gr = 0.0;
        }
        lamQ[s] += gr;
        lamQ[s + nc] -= lg[1];
        lamQ[s + 2 * nc] -= lg[2];
        lamQ[s + 3 * nc] -= lg[3];
        lamQ[s + 4 * nc] += gpv;
        lamQ[s + 5 * nc] += gT;
      }
    }
  }
  if (bc == F3D_BC_DIRICHLET || bc == F3D_BC_WALL_ISOTHERMAL) {
    __shared__ double sm[7][4];
    const int tid = threadIdx.y * 64 + threadIdx.x;
#pragma unroll
    for (int n = 0; n < 7; ++n) {
      double v = acc[n];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
      if ((tid & 63) == 0) sm[n][tid >> 6] = v;
    }
    __syncthreads();
    // the block's partial sums go into its own first ghost line (cleared above, read by no other block);
    // k_adj_bc_final adds them up in block order and clears them again
    if (tid < 7 && lam_bc) {
      const double v = sm[tid][0] + sm[tid][1] + sm[tid][2] + sm[tid][3];
      lamQ[bc_slot<D>(kp, side, blockIdx.x * 64, blockIdx.y * 4, tid)] = v;
    }
  }
}

// k_adj_ghost<0> with the four ghost cells of a line on four neighbouring lanes (a wave covers 16 lines:
// 32-byte runs instead of one cache line per lane).  Each lane clears its ghost cell; where the four
// ghosts fold into distinct sources (periodic, walls) each lane adds its own, where they share one
// target (zero gradient, inflow, Dirichlet) lane 0 of the line adds them in the order of k_adj_ghost.
__global__ __launch_bounds__(256) void k_adj_ghost_x(KP kp, int side0, int bc0, int bc1, double* __restrict__ lamQ,
                                                     const double* __restrict__ Q, double* __restrict__ lam_inflow,
                                                     double* __restrict__ lam_bc) {
  const int side = side0 + (int)blockIdx.z, bc = side ? bc1 : bc0;
  const int l = threadIdx.x & 3, a = blockIdx.x * 16 + (threadIdx.x >> 2), b = blockIdx.y * 4 + threadIdx.y;
  const bool act = a < kp.Nyt && b < kp.Nzt;
  const int NG = kp.NG, Np = kp.Nxp;
  const long long nc = kp.nc;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const int g = side ? Np + NG + l : l;
  const long long o = act ? cidx(kp, g, a, b) : 0;
  double lg[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    lg[n] = act ? lamQ[o + n * nc] : 0.0;
    if (act) lamQ[o + n * nc] = 0.0;
  }
  // the four ghosts of the line, for lane 0 (shuffles by every lane: the line's lanes are one quad)
  double all[4][6];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int n = 0; n < 6; ++n) all[q][n] = __shfl(lg[n], (threadIdx.x & ~3) | q, 64);
  if (act) {
    if (bc == F3D_BC_PERIODIC) {
      const long long sidx = cidx(kp, side ? g - Np : g + Np, a, b);
#pragma unroll
      for (int n = 0; n < 6; ++n) lamQ[sidx + n * nc] += lg[n];
    } else if (bc == F3D_BC_ZEROGRAD || bc == F3D_BC_INFLOW || bc == F3D_BC_DIRICHLET) {
      if (l == 0) {
        if (bc == F3D_BC_ZEROGRAD) {
          const long long sidx = cidx(kp, side ? Np + NG - 1 : NG, a, b);
          for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int n = 0; n < 6; ++n) lamQ[sidx + n * nc] += all[q][n];
        } else if (bc == F3D_BC_DIRICHLET) {
          for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int n = 0; n < 6; ++n) acc[n] += all[q][n];
        } else if (lam_inflow) {
          const long long io = a + (long long)kp.Nyt * b, in = (long long)kp.Nyt * kp.Nzt;
          for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int n = 0; n < 6; ++n) lam_inflow[io + n * in] += all[q][n];
        }
      }
    } else {
      const int mi = side ? 2 * (Np + NG) - 1 - g : 2 * NG - 1 - g;
      const long long sidx = cidx(kp, mi, a, b);
      double gr = lg[0], gpv = lg[4], gT = lg[5];
      if (bc == F3D_BC_WALL_ISOTHERMAL) {
        const double q0 = Q[o], q5 = Q[o + 5 * nc];
        gpv += lg[0] / (kp.Rg * q5);
        const double gq5 = lg[5] - lg[0] * q0 / q5;
        gT = -gq5;
        acc[6] += 2 * gq5;
        gr = 0.0;
      }
      lamQ[sidx] += gr;
      lamQ[sidx + nc] -= lg[1];
      lamQ[sidx + 2 * nc] -= lg[2];
      lamQ[sidx + 3 * nc] -= lg[3];
      lamQ[sidx + 4 * nc] += gpv;
      lamQ[sidx + 5 * nc] += gT;
    }
  }
  if (bc == F3D_BC_DIRICHLET || bc == F3D_BC_WALL_ISOTHERMAL) {
    __shared__ double sm[7][4];
    const int tid = threadIdx.y * 64 + threadIdx.x;
#pragma unroll
    for (int n = 0; n < 7; ++n) {
      double v = acc[n];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
      if ((tid & 63) == 0) sm[n][tid >> 6] = v;
    }
    __syncthreads();
    // the block's partial sums go into its own first ghost line (cleared above, read by no other block);
    // k_adj_bc_final adds them up in block order and clears them again
    if (tid < 7 && lam_bc) {
      const double v = sm[tid][0] + sm[tid][1] + sm[tid][2] + sm[tid][3];
      lamQ[bc_slot<0>(kp, side, blockIdx.x * 16, blockIdx.y * 4, tid)] = v;
    }
  }
}

// ------------------------------------------------------------------ c2Prim + RK3 blend, reverse
// lam_s = a * lam_out + c2Prim^T(lamQ) at U_s (+ lamUn for KRK 1);  lamUn = b*lam_out (KRK 3),
// += b*lam_out (KRK 2).  lam_s may alias lam_out.
template <int KRK>
__global__ __launch_bounds__(256) void k_adj_finalize(KP kp, const double* __restrict__ lamQ, const double* Us,
                                                      const double* lam_out, double* lam_s, double* lamUn) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= kp.Nxp || j >= kp.Nyp) return;
  const long long o = cidx(kp, i + kp.NG, j + kp.NG, k + kp.NG), nc = kp.nc;
  const double a = KRK == 1 ? 1.0 : (KRK == 2 ? 0.25 : 2.0 / 3.0);
  const double b = KRK == 1 ? 0.0 : (KRK == 2 ? 0.75 : 1.0 / 3.0);
  double U[5], gQ[6], gU[5], lo[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    U[n] = Us[o + n * nc];
    lo[n] = lam_out[o + n * nc];
  }
#pragma unroll
  for (int n = 0; n < 6; ++n) gQ[n] = lamQ[o + n * nc];
  c2prim_vjp(kp, U, gQ, gU);
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    double r = a * lo[n] + gU[n];
    if (KRK == 1) r += lamUn[o + n * nc];
    if (KRK == 2) lamUn[o + n * nc] += b * lo[n];
    if (KRK == 3) lamUn[o + n * nc] = b * lo[n];
    lam_s[o + n * nc] = r;
  }
}

// k_adj_finalize inside the multi-step adjoint, with the two passes that follow it folded in: (SEED, KRK 1)
// the objective seed of the previous step, lam += c2Prim^T(wscale * W) at the same U (c2Prim^T is linear in
// its argument, so it is taken once of lamQ + wscale * W), and Lam = nscale * J * lam_s of the next stage
// VJP (k_adj_lambda's product on the value just formed).
template <int KRK, bool SEED>
__global__ __launch_bounds__(256) void k_adj_finalize_next(KP kp, const double* __restrict__ lamQ, const double* Us,
                                                           const double* lam_out, double* lam_s, double* lamUn,
                                                           const double* __restrict__ W, double wscale,
                                                           const double* __restrict__ J, double nscale,
                                                           double* __restrict__ Lam) {
  // blocks of 256 cells along one row: 2 KB runs of each of the ~30 arrays streamed
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= kp.Nxp || j >= kp.Nyp) return;
  const long long o = cidx(kp, i + kp.NG, j + kp.NG, k + kp.NG), nc = kp.nc;
  const double a = KRK == 1 ? 1.0 : (KRK == 2 ? 0.25 : 2.0 / 3.0);
  const double b = KRK == 1 ? 0.0 : (KRK == 2 ? 0.75 : 1.0 / 3.0);
  double U[5], gQ[6], gU[5], lo[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    U[n] = Us[o + n * nc];
    lo[n] = lam_out[o + n * nc];
  }
#pragma unroll
  for (int n = 0; n < 6; ++n) gQ[n] = SEED ? lamQ[o + n * nc] + wscale * W[o + n * nc] : lamQ[o + n * nc];
  c2prim_vjp(kp, U, gQ, gU);
  const double sj = nscale * J[o];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    double r = a * lo[n] + gU[n];
    if (KRK == 1) r += lamUn[o + n * nc];
    if (KRK == 2) lamUn[o + n * nc] += b * lo[n];
    if (KRK == 3) lamUn[o + n * nc] = b * lo[n];
    lam_s[o + n * nc] = r;
    Lam[o + n * nc] = sj * r;
  }
}

// c2Prim^T of one cell into a U-shaped adjoint: lam += scale * c2Prim^T(W) at U (objective seed).
__global__ __launch_bounds__(256) void k_adj_seed(KP kp, const double* __restrict__ U, const double* __restrict__ W,
                                                  double scale, double* __restrict__ lam) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= kp.Nxp || j >= kp.Nyp) return;
  const long long o = cidx(kp, i + kp.NG, j + kp.NG, k + kp.NG), nc = kp.nc;
  double Uc[5], gQ[6], gU[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) Uc[n] = U[o + n * nc];
#pragma unroll
  for (int n = 0; n < 6; ++n) gQ[n] = scale * W[o + n * nc];
  c2prim_vjp(kp, Uc, gQ, gU);
#pragma unroll
  for (int n = 0; n < 5; ++n) lam[o + n * nc] += gU[n];
}

// c2Prim^T over the interior (f3d_c2prim_vjp_f64): gU = c2Prim^T(lamQ) at U, overwriting gU.
__global__ __launch_bounds__(256) void k_adj_c2prim(KP kp, const double* __restrict__ U,
                                                    const double* __restrict__ lamQ, double* __restrict__ gUo) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= kp.Nxp || j >= kp.Nyp) return;
  const long long o = cidx(kp, i + kp.NG, j + kp.NG, k + kp.NG), nc = kp.nc;
  double Uc[5], gQ[6], gU[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) Uc[n] = U[o + n * nc];
#pragma unroll
  for (int n = 0; n < 6; ++n) gQ[n] = lamQ[o + n * nc];
  c2prim_vjp(kp, Uc, gQ, gU);
#pragma unroll
  for (int n = 0; n < 5; ++n) gUo[o + n * nc] = gU[n];
}

// Objective term sum over interior cells of W . Q (6 vars): per-block partials, then one sum.
__global__ __launch_bounds__(256) void k_obj_partial(KP kp, const double* __restrict__ Q, const double* __restrict__ W,
                                                     double* __restrict__ part) {
  const long long n = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
  double s = 0.0;
  for (long long t = blockIdx.x * 256LL + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
    const int i = (int)(t % kp.Nxp), j = (int)((t / kp.Nxp) % kp.Nyp), k = (int)(t / ((long long)kp.Nxp * kp.Nyp));
    const long long o = cidx(kp, i + kp.NG, j + kp.NG, k + kp.NG);
#pragma unroll
    for (int q = 0; q < 6; ++q) s += W[o + q * kp.nc] * Q[o + q * kp.nc];
  }
  __shared__ double sm[4];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

__global__ __launch_bounds__(256) void k_obj_final(const double* __restrict__ part, int n, double scale,
                                                   double* __restrict__ out) {
  double s = 0.0;
  for (int t = threadIdx.x; t < n; t += 256) s += part[t];
  __shared__ double sm[4];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = scale * (sm[0] + sm[1] + sm[2] + sm[3]);
}

// ------------------------------------------------------------------ host orchestration
dim3 grid_interior(const KP& kp) { return dim3((kp.Nxp + 63) / 64, (kp.Nyp + 3) / 4, kp.Nzp); }
dim3 grid_all(const KP& kp) { return dim3((kp.Nxt + 63) / 64, (kp.Nyt + 3) / 4, kp.Nzt); }
const dim3 kB(64, 4, 1);

int check_fdm(const f3d_params* p, const char* who) {
  if (p->fvm != F3D_FDM) {
    set_error("%s: the adjoint covers the finite-difference branch only (params fvm = %d)", who, p->fvm);
    return F3D_E_ARG;
  }
  if (p->split != F3D_SPLIT_SW || p->character != 1 || p->les != F3D_LES_NONE || p->lts || p->filtering) {
    set_error("%s: the adjoint covers the Steger-Warming, characteristic, constant-viscosity, global-dt, "
              "unfiltered path only (params split = %d, character = %d, les = %d, lts = %d, filtering = %d)",
              who, p->split, p->character, p->les, p->lts, p->filtering);
    return F3D_E_ARG;
  }
  return F3D_OK;
}

int check_aws(const f3d_params* p, const f3d_adj_workspace* aws, const char* who) {
  int rc = f3d_params_check(p);
  if (!rc) rc = check_fdm(p, who);
  if (rc) return rc;
  if (!aws) {
    set_error("%s: adjoint workspace is NULL", who);
    return F3D_E_NULL;
  }
  if (p->Nxp != aws->p.Nxp || p->Nyp != aws->p.Nyp || p->Nzp != aws->p.Nzp || p->NG != aws->p.NG) {
    set_error("%s: adjoint workspace was created for another grid", who);
    return F3D_E_WS;
  }
  return F3D_OK;
}

int check_metrics(const double* const* metrics, const double* J, const char* who) {
  if (!metrics || !J) {
    set_error("%s: NULL metric / Jacobian argument", who);
    return F3D_E_NULL;
  }
  for (int q = 0; q < 9; ++q)
    if (!metrics[q]) {
      set_error("%s: metric %d is NULL", who, q);
      return F3D_E_NULL;
    }
  return F3D_OK;
}

// lamQ = (dR/dQ)^T (scale * lam) over all cells of Q (RHS per unit dt, R = J * sum of differences).
// Grid of the list kernel of direction D: one lane per face of the interior face range.
// 1 unless F3D_ADJ_VISC=split: the viscous reverse as one march (k_adj_visc_march) instead of
// k_adj_visc1 + k_adj_visc2 (A/B measurement).
bool adj_visc_march() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("F3D_ADJ_VISC");
    v = (e && std::strcmp(e, "split") == 0) ? 0 : 1;
  }
  return v == 1;
}

// Planes per block of k_adj_visc_march: one block per CU at a time, so the march takes about
// rounds x (KC + 4) plane steps with rounds = ceil(blocks / CUs); the KC of equal segments minimising it.
// F3D_ADJ_VISC_KC overrides.
int adj_visc_kc(const KP& kp) {
  static int env = -2, ncu = 0;
  if (env == -2) {
    const char* e = std::getenv("F3D_ADJ_VISC_KC");
    env = e ? std::atoi(e) : -1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                hipSuccess || ncu < 1)
      ncu = 256;
  }
  if (env > 0) return env;
  // the k range in nseg equal segments, nseg = 1 .. 32 (on the 256x128x64 ramp: 2 segments of 36 planes, one
  // round of 240 blocks, 301 us per launch against 343 for 6 of 12 in three rounds)
  const long long tiles = (long long)((kp.Nxt + kAVW - 5) / (kAVW - 4)) * ((kp.Nyt + kAVR - 5) / (kAVR - 4));
  int best = kp.Nzt;
  long long cost = -1;
  for (int nseg = 1; nseg <= 32; ++nseg) {
    const int kc = (kp.Nzt + nseg - 1) / nseg;
    const long long nb = tiles * ((kp.Nzt + kc - 1) / kc);
    const long long c = ((nb + ncu - 1) / ncu) * (kc + 4);
    if (cost < 0 || c < cost) {
      cost = c;
      best = kc;
    }
  }
  return best;
}

// One lane per cell of the (Nxp+1) x (Nyp+1) x (Nzp+1) face-index box (its x, y and z faces).
// 1 (default): the inviscid VJP as the z and x+y marches; F3D_ADJ_INV=gather: k_adj_inviscid (A/B).
int adj_inv_march() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("F3D_ADJ_INV");
    v = (e && std::strcmp(e, "gather") == 0) ? 0 : 1;
  }
  return v;
}

// Faces of the grid (x, y and z faces of the interior).
long long face_count(const f3d_params* p) {
  const long long nx = p->Nxp, ny = p->Nyp, nz = p->Nzp;
  return (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1);
}

// Listed faces whose reverse contributions the workspace keeps for the ordered gather: every face of the
// grid up to 2^23 faces (2.5 GiB), else an eighth of the cell count (4.2 % of the faces), at most 2^24 - 2
// (24-bit record slots); F3D_ADJ_HEAVY_CAP overrides (testing the path beyond it).  When it covers every
// face, the atomic path for the faces beyond it is never launched.
unsigned int adj_heavy_cap(const f3d_params* p, size_t nc) {
  const long long faces = face_count(p);
  const char* e = std::getenv("F3D_ADJ_HEAVY_CAP");
  long long c = std::min(faces, std::max((long long)(nc / 8), 1LL << 23));
  if (e) c = std::atoll(e);
  if (c < 1) c = 1;
  if (c > (1LL << 24) - 2) c = (1LL << 24) - 2;
  return (unsigned int)c;
}

int seg_len_adj(int N, int max_len) {
  const int nseg = (N + max_len - 1) / max_len;
  return (N + nseg - 1) / nseg;
}

dim3 adj_list_grid(const KP& kp) {
  const long long npos = (long long)(kp.Nxp + 1) * (kp.Nyp + 1);
  return dim3((unsigned int)((npos + 255) / 256), 1, kp.Nzp + 1);
}

}  // namespace

HList f3d::adj_heavy_layout(const KP& kp) {
  unsigned long long nw[3], cap[3];
  const dim3 g = adj_list_grid(kp);
  for (int d = 0; d < 3; ++d) {
    nw[d] = 4ULL * g.x * g.y * g.z;
    cap[d] = 64;
  }
  return make_hlist(nw, cap);
}

namespace {

// lam_ready: aws->Lam already holds scale * J * lam (formed by the previous finalize pass).
int rhs_vjp_impl(const f3d_params* p, f3d_adj_workspace* aws, const double* lam, double scale, const double* Q,
                 const double* const* metrics, const double* J, double* lamQ, hipStream_t s, bool lam_ready = false) {
  const KP kp = make_kp(p);
  MetricPtrs<double> mp;
  for (int q = 0; q < 9; ++q) mp.m[q] = metrics[q];
  const WsView<double> fv = ws_view<double>(aws->fws);
  const HList hl = f3d::adj_heavy_layout(kp);
  int rc = check_hlist(fv, hl, "rhs vjp");
  if (rc) return rc;
  fused_sensor<double>(kp, fv, Q, nullptr, s);  // the same sensor classes; marks whether any left the UP7 band
  // a new record stamp per evaluation (8 bits; the records are cleared when it wraps)
  if (++aws->hstamp > 255u) {
    // the record and flag arrays lie back to back
    hipLaunchKernelGGL(k_adj_zero, dim3(1024), dim3(256), 0, s, (double*)aws->hrec,
                       (long long)((align_up(3 * (size_t)kp.nc * sizeof(unsigned int)) + (size_t)kp.nc + sizeof(double) - 1) /
                                   sizeof(double)));
    aws->hstamp = 1;
  }
  const HRec hr{fv.pgen ? fv.pflag : nullptr, fv.pgen ? *fv.pgen : 0u, aws->hbuf, aws->hrec, aws->hflag, aws->hctr,
                aws->hcap, aws->hstamp << 24};
  const unsigned int nbh = hl.items() < 8192u ? (hl.items() ? hl.items() : 1u) : 8192u;
  {
    TimedScope ts("adj_heavy_list", s);
    hipLaunchKernelGGL(k_adj_heavy_list, adj_list_grid(kp), kB, 0, s, kp, fv.cls, fv.heavy, fv.hcount, hl, aws->hctr,
                       fv.pgen ? fv.pflag : nullptr, fv.pgen ? *fv.pgen : 0u);
  }
  if (!lam_ready) {
    TimedScope ts("adj_lambda", s);
    hipLaunchKernelGGL(k_adj_lambda, grid_interior(kp), kB, 0, s, kp, lam, J, scale, aws->Lam);
  }
  {
    TimedScope ts("adj_heavy", s);
    hipLaunchKernelGGL(k_adj_heavy_form, dim3(nbh), dim3(64), 0, s, kp, Q, mp, aws->Lam, fv.heavy, fv.hcount, hl, hr);
  }
  if (adj_inv_march()) {
    TimedScope ts("adj_inviscid", s);
    // segments of at most 24 cells (measured on the 256x128x64 ramp: z / y segments of 24 / 24 take 228 us,
    // 32 / 32 246 us, 72 / 68 463 us: more, shorter columns fill the chip); F3D_ADJ_INV_LZ / _LY override
    const char* ez = std::getenv("F3D_ADJ_INV_LZ");
    const char* ey = std::getenv("F3D_ADJ_INV_LY");
    const int Lz = seg_len_adj(kp.Nzt, ez && std::atoi(ez) >= 8 ? std::atoi(ez) : 24);
    const int Ly = seg_len_adj(kp.Nyt, ey && std::atoi(ey) >= 8 ? std::atoi(ey) : 24);
    hipLaunchKernelGGL(k_adj_inv_z, dim3((kp.Nxt + 63) / 64, (kp.Nyt + 3) / 4, (kp.Nzt + Lz - 1) / Lz), kB, 0, s, kp,
                       Q, mp, aws->Lam, lamQ, Lz, hr);
    hipLaunchKernelGGL(k_adj_inv_xy, dim3((kp.Nxt + kAII - 1) / kAII, (kp.Nzp + kAIL - 1) / kAIL, (kp.Nyt + Ly - 1) / Ly),
                       dim3(kAIW, kAIL, 1), 0, s, kp, Q, mp, aws->Lam, lamQ, Ly);
  } else {
    TimedScope ts("adj_inviscid", s);
    hipLaunchKernelGGL(k_adj_inviscid, grid_all(kp), kB, 0, s, kp, Q, mp, aws->Lam, lamQ, hr);
  }
  if ((long long)aws->hcap < face_count(p)) {
    TimedScope ts("adj_heavy_rest", s);
    // a small grid: it exits at once unless the buffer overflowed (the walk strides over the items)
    hipLaunchKernelGGL(k_adj_heavy_rest, dim3(nbh < 256u ? nbh : 256u), dim3(64), 0, s, kp, Q, mp, aws->Lam, fv.heavy,
                       fv.hcount, lamQ, hl, hr);
  }
  if (kp.viscous && adj_visc_march()) {
    TimedScope ts("adj_visc", s);
    const int KC = adj_visc_kc(kp);
    hipLaunchKernelGGL(k_adj_visc_march, dim3((kp.Nxt + kAVW - 5) / (kAVW - 4), (kp.Nyt + kAVR - 5) / (kAVR - 4),
                                              (kp.Nzt + KC - 1) / KC), dim3(64, kAVR / 2, 1), 0, s, kp, Q, mp, J,
                       aws->Lam, lamQ, KC);
  } else if (kp.viscous) {
    {
      TimedScope ts("adj_visc1", s);
      hipLaunchKernelGGL(k_adj_visc1, dim3((kp.Nxt - 4 + 63) / 64, (kp.Nyt - 4 + 3) / 4, kp.Nzt - 4), kB, 0, s, kp,
                         Q, mp, J, aws->Lam, aws->G);
    }
    {
      TimedScope ts("adj_visc2", s);
      hipLaunchKernelGGL(k_adj_visc2, grid_all(kp), kB, 0, s, kp, aws->G, lamQ);
    }
  }
  return check_launch("rhs vjp");
}

// Reverse of the stage ghost procedure (exchange_ghost then fill_ghost, solver.jl:439-443).
int ghost_vjp_impl(const f3d_params* p, double* lamQ, const double* Q, double* lam_inflow, double* lam_bc,
                   hipStream_t s) {
  const KP kp = make_kp(p);
  TimedScope ts("adj_ghost", s);
  // the two sides of a dimension in one launch when their source cells cannot coincide (Np >= 2 NG)
  auto launch = [&](int d) {
    const int bc0 = p->bc[2 * d], bc1 = p->bc[2 * d + 1];
    const int Np = d == 0 ? kp.Nxp : (d == 1 ? kp.Nyp : kp.Nzp);
    const bool both = Np >= 2 * kp.NG;
    auto reduces = [&](int bc) { return lam_bc && (bc == F3D_BC_DIRICHLET || bc == F3D_BC_WALL_ISOTHERMAL); };
    const dim3 g = d == 0 ? dim3((kp.Nyt + 15) / 16, (kp.Nzt + 3) / 4)
                          : (d == 1 ? dim3((kp.Nxt + 63) / 64, (kp.Nzt + 3) / 4) : dim3((kp.Nxt + 63) / 64, (kp.Nyt + 3) / 4));
    auto ghost = [&](int side0, int nside) {
      const dim3 gz(g.x, g.y, nside);
      if (d == 0)
        hipLaunchKernelGGL(k_adj_ghost_x, gz, kB, 0, s, kp, side0, bc0, bc1, lamQ, Q, lam_inflow, lam_bc);
      else if (d == 1)
        hipLaunchKernelGGL(k_adj_ghost<1>, gz, kB, 0, s, kp, side0, bc0, bc1, lamQ, Q, lam_inflow, lam_bc);
      else
        hipLaunchKernelGGL(k_adj_ghost<2>, gz, kB, 0, s, kp, side0, bc0, bc1, lamQ, Q, lam_inflow, lam_bc);
    };
    auto final_ = [&](int red0, int red1) {
      if (!red0 && !red1) return;
      const int lw = d == 0 ? 16 : 64;
      if (d == 0)
        hipLaunchKernelGGL(k_adj_bc_final<0>, dim3(1), dim3(256), 0, s, kp, red0, red1, lw, (int)g.x, (int)g.y, lamQ,
                           lam_bc);
      else if (d == 1)
        hipLaunchKernelGGL(k_adj_bc_final<1>, dim3(1), dim3(256), 0, s, kp, red0, red1, lw, (int)g.x, (int)g.y, lamQ,
                           lam_bc);
      else
        hipLaunchKernelGGL(k_adj_bc_final<2>, dim3(1), dim3(256), 0, s, kp, red0, red1, lw, (int)g.x, (int)g.y, lamQ,
                           lam_bc);
    };
    if (both) {
      ghost(0, 2);
      final_(reduces(bc0), reduces(bc1));
    } else {  // the high side, then the low side (the reverse of the fill's order)
      ghost(1, 1);
      final_(0, reduces(bc1));
      ghost(0, 1);
      final_(reduces(bc0), 0);
    }
  };
  // reverse of fill_z, fill_y, fill_x.  The reverse of exchange_ghost (mpi.jl, one rank) is not launched:
  // each reverse fill clears its whole slab, so every ghost adjoint is zero by then and the periodic
  // folds of the exchange would add zeros.
  for (int d = 2; d >= 0; --d) launch(d);
  return check_launch("ghost vjp");
}

// Next: the multi-step driver's fold of the following passes into the finalize (k_adj_finalize_next);
// nscale == 0 keeps the plain finalize.
struct FinNext {
  bool lam_ready = false;       // aws->Lam already formed for this stage
  double nscale = 0.0;          // scale of the next stage VJP's Lam (0: not formed here)
  const double* W = nullptr;    // objective weights of the seed folded into KRK 1 (nullptr: none)
  double wscale = 0.0;
};

int stage_vjp_impl(const f3d_params* p, f3d_adj_workspace* aws, int KRK, const double* lam_out, double* lam_s,
                   double* lamUn, const double* Us, const double* Qs, const double* const* metrics, const double* J,
                   double dt, double* lam_inflow, double* lam_bc, hipStream_t s, const FinNext& nx = FinNext()) {
  const KP kp = make_kp(p);
  const double a = KRK == 1 ? 1.0 : (KRK == 2 ? 0.25 : 2.0 / 3.0);
  int rc = rhs_vjp_impl(p, aws, lam_out, a * dt, Qs, metrics, J, aws->lamQ, s, nx.lam_ready);
  if (rc) return rc;
  rc = ghost_vjp_impl(p, aws->lamQ, Qs, lam_inflow, lam_bc, s);
  if (rc) return rc;
  TimedScope ts("adj_finalize", s);
  const dim3 g = grid_interior(kp);
  const dim3 gr((kp.Nxp + 255) / 256, kp.Nyp, kp.Nzp), br(256, 1, 1);
  if (nx.nscale != 0.0) {
    if (KRK == 1 && nx.W)
      hipLaunchKernelGGL((k_adj_finalize_next<1, true>), gr, br, 0, s, kp, aws->lamQ, Us, lam_out, lam_s, lamUn, nx.W,
                         nx.wscale, J, nx.nscale, aws->Lam);
    else if (KRK == 1)
      hipLaunchKernelGGL((k_adj_finalize_next<1, false>), gr, br, 0, s, kp, aws->lamQ, Us, lam_out, lam_s, lamUn,
                         nx.W, nx.wscale, J, nx.nscale, aws->Lam);
    else if (KRK == 2)
      hipLaunchKernelGGL((k_adj_finalize_next<2, false>), gr, br, 0, s, kp, aws->lamQ, Us, lam_out, lam_s, lamUn,
                         nx.W, nx.wscale, J, nx.nscale, aws->Lam);
    else
      hipLaunchKernelGGL((k_adj_finalize_next<3, false>), gr, br, 0, s, kp, aws->lamQ, Us, lam_out, lam_s, lamUn,
                         nx.W, nx.wscale, J, nx.nscale, aws->Lam);
    return check_launch("stage vjp finalize");
  }
  if (KRK == 1)
    hipLaunchKernelGGL(k_adj_finalize<1>, g, kB, 0, s, kp, aws->lamQ, Us, lam_out, lam_s, lamUn);
  else if (KRK == 2)
    hipLaunchKernelGGL(k_adj_finalize<2>, g, kB, 0, s, kp, aws->lamQ, Us, lam_out, lam_s, lamUn);
  else
    hipLaunchKernelGGL(k_adj_finalize<3>, g, kB, 0, s, kp, aws->lamQ, Us, lam_out, lam_s, lamUn);
  return check_launch("stage vjp finalize");
}

// Conserved states in the checkpoint buffer: the checkpoints, the every-1 recomputed states of a segment,
// the final state, and the two stage states (U^1, U^2) of each step of a segment (of every step when
// every == 1, kept by the forward sweep itself).
size_t stage_steps(int nsteps, int every) { return (size_t)(every == 1 ? nsteps : every); }
size_t ckpt_slots(int nsteps, int every) {
  return (size_t)((nsteps + every - 1) / every) + (size_t)(every - 1) + 1 + 2 * stage_steps(nsteps, every);
}
// With a checkpoint every step (and the fused stage), the forward sweep also keeps the primitive states
// (ghosts filled) of the three stage inputs of every step, as the stages form them, so the reverse sweep
// needs no c2Prim + ghost fill of its own: 3 x 6 values per cell and step after the conserved states.
bool keep_q(int every) { return every == 1 && use_fused(); }

}  // namespace

// ================================================================ extern "C"
#define F3D_API extern "C" __attribute__((visibility("default")))

F3D_API int f3d_adjoint_workspace_create(const f3d_params* p, f3d_adj_workspace** out) {
  int rc = f3d_params_check(p);
  if (!rc) rc = check_fdm(p, "adjoint_workspace_create");
  if (rc) return rc;
  if (!out) {
    set_error("adjoint_workspace_create: out is NULL");
    return F3D_E_NULL;
  }
  *out = nullptr;
  f3d_workspace* fws = nullptr;
  rc = f3d_workspace_create(p, F3D_PREC_F64, &fws);
  if (rc) return rc;
  const size_t nc = f3d_cell_count(p), v = nc * sizeof(double);
  // Lam 5, lamQ 6, G 16 (the split viscous reverse only), Qs 3x6, Un 5, lam 5, lamUn 5, Ucur 5, Qcur 6, red
  const size_t nvars = 5 + 6 + (adj_visc_march() ? 0 : 16) + 18 + 5 + 5 + 5 + 5 + 6;
  const unsigned int hcap = adj_heavy_cap(p, nc);
  const size_t hbytes = align_up((size_t)hcap * 40 * sizeof(double)) + align_up(3 * nc * sizeof(unsigned int)) +
                        align_up(nc) + 256;
  const size_t total = nvars * align_up(v) + align_up(kObjBlocks * sizeof(double)) + hbytes;
  void* base = nullptr;
  hipError_t e = hipMalloc(&base, total);
  if (e != hipSuccess) {
    f3d_workspace_destroy(fws);
    set_error("adjoint_workspace_create: hipMalloc of %zu bytes failed: %s", total, hipGetErrorString(e));
    return (int)e;
  }
  f3d_adj_workspace* a = new f3d_adj_workspace();
  a->p = *p;
  a->fws = fws;
  a->base = base;
  a->bytes = total + f3d_workspace_bytes(fws);
  char* c = (char*)base;
  auto take = [&](int nv) {
    double* r = (double*)c;
    c += nv * align_up(v);
    return r;
  };
  a->Lam = take(5);
  a->lamQ = take(6);
  a->G = adj_visc_march() ? nullptr : take(16);
  for (int q = 0; q < 3; ++q) a->Qs[q] = take(6);
  a->Un = take(5);
  a->lam = take(5);
  a->lamUn = take(5);
  a->Ucur = take(5);
  a->Qcur = take(6);
  a->red = (double*)c;
  a->nred = kObjBlocks;
  c += align_up(kObjBlocks * sizeof(double));
  a->hbuf = (double*)c;
  c += align_up((size_t)hcap * 40 * sizeof(double));
  a->hrec = (unsigned int*)c;
  c += align_up(3 * nc * sizeof(unsigned int));
  a->hflag = (unsigned char*)c;
  c += align_up(nc);
  a->hctr = (unsigned int*)c;
  a->hcap = hcap;
  a->hstamp = 0;
  // zeroed once: ghost entries of Lam are never written and must read as zero
  hipLaunchKernelGGL(k_adj_zero, dim3(1024), dim3(256), 0, 0, (double*)base, (long long)(total / 8));
  e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(base);
    f3d_workspace_destroy(fws);
    delete a;
    set_error("adjoint_workspace_create: zero fill failed: %s", hipGetErrorString(e));
    return (int)e;
  }
  *out = a;
  return F3D_OK;
}

F3D_API int f3d_adjoint_workspace_destroy(f3d_adj_workspace* a) {
  if (!a) return F3D_OK;
  hipError_t e = hipFree(a->base);
  int rc = f3d_workspace_destroy(a->fws);
  delete a;
  if (e != hipSuccess) {
    set_error("adjoint_workspace_destroy: %s", hipGetErrorString(e));
    return (int)e;
  }
  return rc;
}

F3D_API size_t f3d_adjoint_workspace_bytes(const f3d_adj_workspace* a) { return a ? a->bytes : 0; }

F3D_API int f3d_rhs_vjp_f64(const f3d_params* p, f3d_adj_workspace* aws, const double* lamR, const double* Q,
                            const double* const* metrics, const double* J, double* lamQ, f3d_stream_t s) {
  int rc = check_aws(p, aws, "rhs_vjp");
  if (!rc) rc = check_metrics(metrics, J, "rhs_vjp");
  if (rc) return rc;
  if (!lamR || !Q || !lamQ) {
    set_error("rhs_vjp: NULL array argument");
    return F3D_E_NULL;
  }
  return rhs_vjp_impl(p, aws, lamR, 1.0, Q, metrics, J, lamQ, s);
}

F3D_API int f3d_ghost_vjp_f64(const f3d_params* p, double* lamQ, const double* Q, double* lam_inflow, double* lam_bc,
                              f3d_stream_t s) {
  int rc = f3d_params_check(p);
  if (rc) return rc;
  if (!lamQ || !Q) {
    set_error("ghost_vjp: NULL array argument");
    return F3D_E_NULL;
  }
  return ghost_vjp_impl(p, lamQ, Q, lam_inflow, lam_bc, s);
}

F3D_API int f3d_c2prim_vjp_f64(const f3d_params* p, double* lamU, const double* lamQ, const double* U,
                               f3d_stream_t s) {
  int rc = f3d_params_check(p);
  if (rc) return rc;
  if (!lamU || !lamQ || !U) {
    set_error("c2prim_vjp: NULL array argument");
    return F3D_E_NULL;
  }
  const KP kp = make_kp(p);
  hipLaunchKernelGGL(k_adj_c2prim, grid_interior(kp), kB, 0, s, kp, U, lamQ, lamU);
  return check_launch("c2prim vjp");
}

F3D_API int f3d_rk3_stage_vjp_f64(const f3d_params* p, f3d_adj_workspace* aws, int32_t KRK, const double* lam_out,
                                  double* lam_s, double* lamUn, const double* Us, const double* Qs,
                                  const double* const* metrics, const double* J, double dt, double* lam_inflow,
                                  double* lam_bc, f3d_stream_t s) {
  int rc = check_aws(p, aws, "rk3_stage_vjp");
  if (!rc) rc = check_metrics(metrics, J, "rk3_stage_vjp");
  if (rc) return rc;
  if (KRK < 1 || KRK > 3) {
    set_error("rk3_stage_vjp: KRK must be 1, 2 or 3 (got %d)", KRK);
    return F3D_E_ARG;
  }
  if (!lam_out || !lam_s || !lamUn || !Us || !Qs) {
    set_error("rk3_stage_vjp: NULL array argument");
    return F3D_E_NULL;
  }
  return stage_vjp_impl(p, aws, KRK, lam_out, lam_s, lamUn, Us, Qs, metrics, J, dt, lam_inflow, lam_bc, s);
}

F3D_API size_t f3d_adjoint_ckpt_bytes(const f3d_params* p, int32_t nsteps, int32_t every) {
  if (!p || nsteps < 1 || every < 1) return 0;
  if (every > nsteps) every = nsteps;  // one segment of nsteps steps
  const size_t nc = f3d_cell_count(p);
  return (ckpt_slots(nsteps, every) * 5 + (keep_q(every) ? (size_t)nsteps * 18 : 0)) * nc * sizeof(double);
}

F3D_API int f3d_adjoint_f64(const f3d_params* p, f3d_adj_workspace* aws, int32_t nsteps, int32_t every,
                            const double* U0, const double* const* metrics, const double* J, double dt,
                            const double* inflow, const double* W, double* ckpt, size_t ckpt_bytes, double* obj,
                            double* lamU0, double* lam_inflow, double* lam_bc, f3d_stream_t s) {
  int rc = check_aws(p, aws, "adjoint");
  if (!rc) rc = check_metrics(metrics, J, "adjoint");
  if (rc) return rc;
  if (nsteps < 1 || every < 1) {
    set_error("adjoint: nsteps (%d) and every (%d) must be >= 1", nsteps, every);
    return F3D_E_ARG;
  }
  if (every > nsteps) every = nsteps;  // one segment of nsteps steps (the slot layout below assumes it)
  if (!U0 || !W || !ckpt || !obj || !lamU0) {
    set_error("adjoint: NULL array argument");
    return F3D_E_NULL;
  }
  if (p->bc[0] == F3D_BC_INFLOW && !inflow) {
    set_error("adjoint: F3D_BC_INFLOW needs an inflow profile");
    return F3D_E_NULL;
  }
  const size_t need = f3d_adjoint_ckpt_bytes(p, nsteps, every);
  if (ckpt_bytes < need) {
    set_error("adjoint: checkpoint buffer of %zu bytes, %zu needed", ckpt_bytes, need);
    return F3D_E_ARG;
  }
  const KP kp = make_kp(p);
  const size_t nc = kp.nc, vb = 5 * nc * sizeof(double);
  const int nck = (nsteps + every - 1) / every;
  auto slot = [&](size_t q) { return ckpt + q * 5 * nc; };
  double* seg0 = slot(nck);                 // every-1 recomputed states of a segment
  double* Ufin = slot(nck + every - 1);     // U^nsteps
  double* stg0 = slot(nck + every);         // stage states U^1, U^2 per step (of the segment or of all steps)
  auto seg = [&](int q) { return q == 0 ? nullptr : seg0 + (size_t)(q - 1) * 5 * nc; };
  const bool kq = keep_q(every);
  double* const q0 = ckpt + ckpt_slots(nsteps, every) * 5 * nc;
  auto qk = [&](int n, int q) { return q0 + ((size_t)3 * n + q) * 6 * nc; };  // Q of stage q of step n
  auto stg = [&](int n, int n0, int q) {
    const size_t idx = every == 1 ? (size_t)n : (size_t)(n - n0);
    return stg0 + (2 * idx + (size_t)(q - 1)) * 5 * nc;
  };
  auto obj_term = [&](const double* Qn, double* outp) {
    hipLaunchKernelGGL(k_obj_partial, dim3(aws->nred), dim3(256), 0, s, kp, Qn, W, aws->red);
    hipLaunchKernelGGL(k_obj_final, dim3(1), dim3(256), 0, s, aws->red, aws->nred, dt, outp);
  };
  // c2Prim + ghost fill (the fills of periodic dimensions rewrite what exchange_ghost would copy)
  auto prim = [&](const double* Usrc, double* Qd) {
    int r = f3d_c2prim_f64(p, Usrc, Qd, s);
    if (!r) r = f3d_fill_ghost_f64(p, Qd, nullptr, inflow, s);
    return r;
  };
  auto step = [&](double* U, double* Q) {
    return steps_io<double>(p, aws->fws, 1, U, aws->Un, Q, metrics, J, dt, inflow, s);
  };
  auto step_keep = [&](double* U, double* Q, double* U1, double* U2, bool third) {
    return step_keep_stages(p, aws->fws, U, aws->Un, Q, metrics, J, dt, inflow, U1, U2, s, third);
  };
  const int nlast = (nck - 1) * every;  // first step of the last segment
  // ---- forward sweep: checkpoints every `every` steps, objective per step
  F3D_HIP(hipMemcpyAsync(aws->Ucur, U0, vb, hipMemcpyDeviceToDevice, s));
  rc = prim(aws->Ucur, kq ? qk(0, 0) : aws->Qcur);
  // (keeping the stage states of every step when every == 1, else the states and stage states of the
  // last segment, whose reversal comes first)
  for (int n = 0; n < nsteps && !rc; ++n) {
    if (n % every == 0) F3D_HIP(hipMemcpyAsync(slot(n / every), aws->Ucur, vb, hipMemcpyDeviceToDevice, s));
    if (kq) {
      double* Qn1 = n + 1 < nsteps ? qk(n + 1, 0) : aws->Qcur;
      rc = step_keep_q(p, aws->fws, aws->Ucur, aws->Un, qk(n, 0), metrics, J, dt, inflow, stg(n, 0, 1), stg(n, 0, 2),
                       qk(n, 1), qk(n, 2), Qn1, s);
      if (!rc) obj_term(Qn1, obj + n);
      continue;
    }
    if (every == 1) {
      rc = step_keep(aws->Ucur, aws->Qcur, stg(n, 0, 1), stg(n, 0, 2), true);
    } else if (n >= nlast) {
      rc = step_keep(aws->Ucur, aws->Qcur, stg(n, nlast, 1), stg(n, nlast, 2), true);
      if (!rc && n + 1 < nsteps)
        F3D_HIP(hipMemcpyAsync(seg(n + 1 - nlast), aws->Ucur, vb, hipMemcpyDeviceToDevice, s));
    } else {
      rc = step(aws->Ucur, aws->Qcur);
    }
    if (!rc) obj_term(aws->Qcur, obj + n);
  }
  if (rc) return rc;
  F3D_HIP(hipMemcpyAsync(Ufin, aws->Ucur, vb, hipMemcpyDeviceToDevice, s));
  // ---- reverse sweep
  hipLaunchKernelGGL(k_adj_zero, dim3(1024), dim3(256), 0, s, aws->lam, (long long)(5 * nc));
  bool seeded = false;  // the seed of the step being reversed was folded into the previous finalize
  for (int c = nck - 1; c >= 0 && !rc; --c) {
    const int n0 = c * every, n1 = (n0 + every < nsteps) ? n0 + every : nsteps;
    // recompute the steps of the segment from its checkpoint, keeping U^{n0+1} .. U^{n1-1} and the
    // stage states of every step; its last step needs stages 1 and 2 only (U^{n1} is checkpointed).
    // The forward sweep kept all of this for every == 1 and for the last segment.
    if (every > 1 && c < nck - 1 && use_fused()) {
      // the stages read the kept states where they are and write the next ones there: no copies
      rc = prim(slot(c), aws->Qcur);
      for (int n = n0; n < n1 && !rc; ++n) {
        const bool third = n + 1 < n1;
        rc = step_keep_stages(p, aws->fws, n == n0 ? slot(c) : seg(n - n0), aws->Un, aws->Qcur, metrics, J, dt,
                              inflow, stg(n, n0, 1), stg(n, n0, 2), s, third, third ? seg(n + 1 - n0) : nullptr);
      }
    } else if (every > 1 && c < nck - 1) {
      F3D_HIP(hipMemcpyAsync(aws->Ucur, slot(c), vb, hipMemcpyDeviceToDevice, s));
      rc = prim(aws->Ucur, aws->Qcur);
      for (int n = n0; n < n1 && !rc; ++n) {
        rc = step_keep(aws->Ucur, aws->Qcur, stg(n, n0, 1), stg(n, n0, 2), n + 1 < n1);
        if (!rc && n + 1 < n1) F3D_HIP(hipMemcpyAsync(seg(n + 1 - n0), aws->Ucur, vb, hipMemcpyDeviceToDevice, s));
      }
    }
    for (int n = n1 - 1; n >= n0 && !rc; --n) {
      const double* Ustart = (n == n0) ? slot(c) : seg(n - n0);
      if (!seeded) {  // the last step: seed at U^nsteps (the others are folded into the finalize of step n + 1)
        TimedScope ts("adj_seed", s);
        hipLaunchKernelGGL(k_adj_seed, grid_interior(kp), kB, 0, s, kp, Ufin, W, dt, aws->lam);
      }
      // stage states of step n: (U^n, Q^n), (U^1, Q^1), (U^2, Q^2), U read where it is kept (prim writes
      // only Q); the stages' Q is c2Prim + ghost fill of their U, as the stage update itself forms it
      const double* Ust[3] = {Ustart, stg(n, n0, 1), stg(n, n0, 2)};
      const double* Qst[3];
      for (int q = 0; q < 3 && !rc; ++q) {
        if (kq) {
          Qst[q] = qk(n, q);  // formed by the forward sweep's stages
        } else {
          rc = prim(Ust[q], aws->Qs[q]);
          Qst[q] = aws->Qs[q];
        }
      }
      // each finalize forms the next stage VJP's Lam; that of stage 1 also adds the seed of step n - 1
      // (at U^n, its stage-1 state) unless n == 0
      for (int KRK = 3; KRK >= 1 && !rc; --KRK) {
        FinNext nx;
        nx.lam_ready = KRK < 3 || seeded;
        if (KRK > 1) {
          nx.nscale = (KRK == 3 ? 0.25 : 1.0) * dt;
        } else if (n > 0) {
          nx.nscale = (2.0 / 3.0) * dt;
          nx.W = W;
          nx.wscale = dt;
        }
        rc = stage_vjp_impl(p, aws, KRK, aws->lam, aws->lam, aws->lamUn, Ust[KRK - 1], Qst[KRK - 1], metrics, J, dt,
                            lam_inflow, lam_bc, s, nx);
      }
      seeded = true;
    }
  }
  if (rc) return rc;
  F3D_HIP(hipMemcpyAsync(lamU0, aws->lam, vb, hipMemcpyDeviceToDevice, s));
  return check_launch("adjoint");
}
