// Fused RHS / RK3-stage kernels (the MI355X hot path).
#pragma once
#include "f3d_internal.hpp"

namespace f3d {

// Workspace arrays the fused path uses (bulk type T).
template <typename T>
struct WsView {
  T* R;                      // interior RHS accumulator [5][Nzp][Nyp][Nxp]
  T* Fvx;                    // viscous flux scratch, reference Fv layout [4][Nzp+4][Nyp+4][Nxp+4]
  T* Fvy;
  T* Fvz;
  T* phi;                    // shock sensor [Nzt][Nyt][Nxt]
  unsigned long long* heavy; // non-UP7 faces: (right cell << 4) | (scheme << 2) | dir, in per-wave runs (HList)
  unsigned int* hcount;      // entries used in each wave's run
  unsigned long long hcap;   // capacity of `heavy` (slots)
  unsigned long long hncap;  // capacity of `hcount` (runs)
  double* ldt;               // per-cell dt of local time stepping, cell-shaped (written at stage 1)
  T* hflux;                  // per face position: the 5 values a listed face adds to its cells (heavy_face_count x 5)
  unsigned long long* hidx;  // per-cell mark bytes of the listed faces (hmark_words, see HMark)
  unsigned int* pflag;       // device word: the sensor generation of the last stage in which some cell's
                             // sensor value left the UP7 band (no face can be off UP7 otherwise)
  unsigned int* pgen;        // host counter of the sensor generations
  unsigned char* cls;        // per-cell sensor class: scheme_of of the cell's phi (a face's scheme is the largest
                             // class of its six cells: the class of their largest phi, NaN the largest)
};

// Whether any face of this stage can be off UP7: the sensor marks its generation when some cell's value is
// not below phi1 (NaN included), and every face scheme is chosen from the maximum of six cells' values.  A
// NULL flag (callers without the fused sensor) means "possibly".
__device__ __forceinline__ bool heavy_possible(const unsigned int* pflag, unsigned int gen) {
  return pflag == nullptr || *pflag == gen;
}

// Block coordinates with an XCD-aware order.  Workgroups are dispatched round-robin over the 8 XCDs
// (XCD = linear workgroup id mod 8) and each XCD has its own L2; giving every XCD one contiguous
// range of the grid (x fastest) keeps the neighbouring blocks whose halo rows a block re-reads in
// the same L2.  remap = 0 returns the plain blockIdx.
struct Blk {
  int x, y, z;
};
// Rank of workgroup `id` among n when each XCD (id mod 8) takes one contiguous range of ranks.
__device__ __forceinline__ unsigned xcd_rank(unsigned id, unsigned n) {
  const unsigned xcd = id & 7u, r = id >> 3, q = n >> 3, rem = n & 7u;
  return xcd * q + (xcd < rem ? xcd : rem) + r;
}
__device__ __forceinline__ Blk xcd_block(int remap) {
  if (!remap) return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
  const unsigned gx = gridDim.x, gy = gridDim.y, nb = gx * gy * gridDim.z;
  const unsigned L = xcd_rank(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), nb);
  return Blk{(int)(L % gx), (int)((L / gx) % gy), (int)(L / (gx * gy))};
}

// 1 unless F3D_XCD_REMAP=0 (A/B measurement of the block order).
int xcd_remap();

// Heavy-face list (faces the sensor sends to WENO7/WENO5/NND) in per-wave runs: every wave of a sweep
// owns a fixed run of `cap` slots and counts its entries in a register, so no global atomic is needed
// (a single shared counter serialises the sweeps once a few percent of the faces are heavy: one atomic
// per wave and step).  Run w of a launch starts at slot base + w * cap; hcount[first + w] = entries used.
struct HSeg {
  unsigned long long base;
  unsigned int cap;
  unsigned int first;
  unsigned int nw;  // waves (runs) of the launch
};
// The runs of the three sweeps of one stage, for the kernel that processes the listed faces.
struct HList {
  HSeg d[3];
  unsigned int nruns() const { return d[0].nw + d[1].nw + d[2].nw; }
  // work items of for_each_heavy (a block per item; the grid needs no more blocks than items)
  unsigned int items() const {
    unsigned int n = 0;
    for (int q = 0; q < 3; ++q) n += d[q].cap > 64 ? d[q].nw : (d[q].nw + 63) / 64;
    return n;
  }
  unsigned long long slots() const {
    unsigned long long n = 0;
    for (int q = 0; q < 3; ++q) n = n > d[q].base + (unsigned long long)d[q].nw * d[q].cap ? n : d[q].base + (unsigned long long)d[q].nw * d[q].cap;
    return n;
  }
};
// Runs laid out back to back, direction after direction (nw[q] waves of cap[q] slots).
inline HList make_hlist(const unsigned long long nw[3], const unsigned long long cap[3]) {
  HList h{};
  unsigned long long base = 0, first = 0;
  for (int q = 0; q < 3; ++q) {
    h.d[q] = HSeg{base, (unsigned int)cap[q], (unsigned int)first, (unsigned int)nw[q]};
    base += nw[q] * cap[q];
    first += nw[q];
  }
  return h;
}

// Wave-aggregated append into the wave's run: lanes with pred write their entry after the ones already
// there (lane order); cnt (wave-uniform) advances by the number written.
__device__ __forceinline__ void run_append(unsigned long long* list, unsigned long long run0, unsigned int& cnt,
                                           bool pred, unsigned long long entry) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0) return;
  const int lane = __lane_id();
  if (pred) {
    const unsigned long long below = lane == 0 ? 0ULL : (mask & ((~0ULL) >> (64 - lane)));
    list[run0 + cnt + __popcll(below)] = entry;
  }
  cnt += (unsigned int)__popcll(mask);
}

// Work items of the heavy-face walkers (a block per item): a run of more than 64 slots is an item of its
// own (its entries 64 at a time), runs of at most 64 slots go 64 runs to an item (their counts read in one
// coalesced load, their entries packed onto the lanes).
__host__ __device__ inline unsigned int dir_items(const HSeg& hs) { return hs.cap > 64 ? hs.nw : (hs.nw + 63) / 64; }

// Calls f(entry, slot) for the listed faces of item `loc` of one direction (slot: the entry's index in the
// list, where a later pass finds what an earlier one kept for the same face).  Blocks of one 64-lane wave.
template <typename F>
__device__ __forceinline__ void heavy_item(const HSeg& hs, unsigned int loc, const unsigned int* __restrict__ counts,
                                           const unsigned long long* __restrict__ list, F&& f) {
  const int lane = threadIdx.x;
  if (hs.cap > 64) {
    const unsigned int n = counts[hs.first + loc];
    const unsigned long long run0 = hs.base + (unsigned long long)loc * hs.cap;
    for (unsigned int e = lane; e < n; e += 64) f(list[run0 + e], run0 + e);
  } else {
    // the 64 runs' entries packed onto the lanes: exclusive prefix of their counts, then entry e goes
    // to the last run whose prefix is <= e (empty runs share their successor's prefix)
    __shared__ unsigned int pre[64];
    const unsigned int w0 = loc * 64, mine = w0 + lane;
    const unsigned int myn = mine < hs.nw ? counts[hs.first + mine] : 0u;
    unsigned int incl = myn;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned int t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    const unsigned int tot = __shfl(incl, 63, 64);
    pre[lane] = incl - myn;
    __syncthreads();  // (a one-wave block)
    for (unsigned int e = lane; e < tot; e += 64) {
      int lo = 0, hi = 63;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= e) lo = mid; else hi = mid - 1;
      }
      const unsigned long long slot = hs.base + (unsigned long long)(w0 + lo) * hs.cap + (e - pre[lo]);
      f(list[slot], slot);
    }
    __syncthreads();  // pre is rewritten by the next item
  }
}

// heavy_item with a pair of lanes per entry: lanes 2p and 2p + 1 both receive entry p, p + 32, ... of the
// item (32 entries per pass of the wave), so that the two halves of one face's work can run side by side
// and exchange values with cross-lane moves (the two lanes of a pair are always active together).
template <typename F>
__device__ __forceinline__ void heavy_item_pairs(const HSeg& hs, unsigned int loc, const unsigned int* __restrict__ counts,
                                                 const unsigned long long* __restrict__ list, F&& f) {
  const int lane = threadIdx.x, pr = lane >> 1;
  if (hs.cap > 64) {
    const unsigned int n = counts[hs.first + loc];
    const unsigned long long run0 = hs.base + (unsigned long long)loc * hs.cap;
    for (unsigned int e = pr; e < n; e += 32) f(list[run0 + e], run0 + e);
  } else {
    __shared__ unsigned int pre[64];
    const unsigned int w0 = loc * 64, mine = w0 + lane;
    const unsigned int myn = mine < hs.nw ? counts[hs.first + mine] : 0u;
    unsigned int incl = myn;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned int t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    const unsigned int tot = __shfl(incl, 63, 64);
    pre[lane] = incl - myn;
    __syncthreads();  // (a one-wave block)
    for (unsigned int e = pr; e < tot; e += 32) {
      int lo = 0, hi = 63;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= e) lo = mid; else hi = mid - 1;
      }
      const unsigned long long slot = hs.base + (unsigned long long)(w0 + lo) * hs.cap + (e - pre[lo]);
      f(list[slot], slot);
    }
    __syncthreads();  // pre is rewritten by the next item
  }
}

// f(entry, slot) for every listed face, a pair of lanes per face (heavy_item_pairs).
template <typename F>
__device__ __forceinline__ void for_each_heavy_pairs(const HList& hl, const unsigned int* __restrict__ counts,
                                                     const unsigned long long* __restrict__ list, F&& f) {
  const unsigned int i0 = dir_items(hl.d[0]), i1 = dir_items(hl.d[1]), total = i0 + i1 + dir_items(hl.d[2]);
  for (unsigned int it = blockIdx.x; it < total; it += gridDim.x) {
    const int dq = it < i0 ? 0 : (it < i0 + i1 ? 1 : 2);
    heavy_item_pairs(hl.d[dq], it - (dq == 0 ? 0u : (dq == 1 ? i0 : i0 + i1)), counts, list, f);
  }
}

// f(entry, slot) for every listed face of all three directions, items spread over the grid.
template <typename F>
__device__ __forceinline__ void for_each_heavy_slot(const HList& hl, const unsigned int* __restrict__ counts,
                                                    const unsigned long long* __restrict__ list, F&& f) {
  const unsigned int i0 = dir_items(hl.d[0]), i1 = dir_items(hl.d[1]), total = i0 + i1 + dir_items(hl.d[2]);
  for (unsigned int it = blockIdx.x; it < total; it += gridDim.x) {
    const int dq = it < i0 ? 0 : (it < i0 + i1 ? 1 : 2);
    heavy_item(hl.d[dq], it - (dq == 0 ? 0u : (dq == 1 ? i0 : i0 + i1)), counts, list, f);
  }
}

// f(entry) for every listed face of all three directions.
template <typename F>
__device__ __forceinline__ void for_each_heavy(const HList& hl, const unsigned int* __restrict__ counts,
                                               const unsigned long long* __restrict__ list, F&& f) {
  for_each_heavy_slot(hl, counts, list, [&](const unsigned long long v, unsigned long long) { f(v); });
}

// Heavy-face corrections and the marks of the cells they enter.  The forming pass keeps each listed face's
// F_scheme - F_UP7 at its face position (hcorr[5 f + q], f = face_id) and marks the face in its two cells:
// bit o = 2 d + side of the cell's mark byte (side 0 = its left face along d, 1 = its right face), set with
// an integer atomic OR on the 32-bit word holding four cells' bytes (the same bits whatever the order).
// k_heavy_apply then walks the cells in order (coalesced byte reads; the RHS read and written only where a
// byte is non-zero), adds a cell's corrections in the order o = 0..5 (+ for a left face, - for a right
// face) and clears the byte: one lane per cell and a fixed order, so no floating-point atomics and the same
// bits on every evaluation, and the marks are all zero again for the next one.
struct HMark {
  unsigned int* mark;  // one byte per interior cell, four to a word
};
__host__ __device__ inline long long hmark_count(const KP& kp) { return (long long)kp.Nxp * kp.Nyp * kp.Nzp; }
// 8-byte words holding the mark bytes of every interior cell.
__host__ __device__ inline long long hmark_words(const KP& kp) { return (hmark_count(kp) + 7) / 8; }
// Faces of the three directions: (Nxp + 1) Nyp Nzp along x, then Nxp (Nyp + 1) Nzp along y, Nxp Nyp (Nzp + 1)
// along z.
__host__ __device__ inline long long heavy_face_count(const KP& kp) {
  const long long X = kp.Nxp, Y = kp.Nyp, Z = kp.Nzp;
  return (X + 1) * Y * Z + X * (Y + 1) * Z + X * Y * (Z + 1);
}
// Position of the face of direction D left of interior cell (I, J, K); the coordinate along D runs to Np (the
// right boundary face).
__host__ __device__ inline long long face_id(const KP& kp, int D, int I, int J, int K) {
  const long long X = kp.Nxp, Y = kp.Nyp, Z = kp.Nzp;
  if (D == 0) return I + (X + 1) * (J + Y * (long long)K);
  if (D == 1) return (X + 1) * Y * Z + I + X * (J + (Y + 1) * (long long)K);
  return (X + 1) * Y * Z + X * (Y + 1) * Z + I + X * (J + Y * (long long)K);
}
// Keeps the correction of the listed face of direction D left of interior position (I, J, K) (f = its
// coordinate along D, Np the cells along D) and marks it in the cell right of it (f < Np) and the cell left of
// it (f > 0).
template <typename T>
__device__ __forceinline__ void heavy_keep(const KP& kp, const HMark& hm, T* __restrict__ corr, int D, int I, int J,
                                           int K, int f, int Np, const T* dF) {
  const long long fid = face_id(kp, D, I, J, K);
#pragma unroll
  for (int q = 0; q < 5; ++q) corr[fid * 5 + q] = dF[q];
  const long long rr = (long long)I + (long long)kp.Nxp * ((long long)J + (long long)kp.Nyp * K);
  if (f < Np) atomicOr(&hm.mark[rr >> 2], (1u << (2 * D)) << (8 * (int)(rr & 3)));
  if (f > 0) {
    const long long rl = rr - (D == 0 ? 1 : (D == 1 ? (long long)kp.Nxp : (long long)kp.Nxp * kp.Nyp));
    atomicOr(&hm.mark[rl >> 2], (1u << (2 * D + 1)) << (8 * (int)(rl & 3)));
  }
}
// Adds the corrections of interior cell (I, J, K), whose mark byte is mb, to its RHS Rt, in the order x left,
// x right, y left, y right, z left, z right.
template <typename T>
__device__ __forceinline__ void heavy_add_mark(const KP& kp, unsigned int mb, const T* __restrict__ corr, int I,
                                               int J, int K, T* Rt) {
#pragma unroll
  for (int o = 0; o < 6; ++o) {
    if (!((mb >> o) & 1u)) continue;
    const int D = o >> 1, sd = o & 1;
    const long long f = face_id(kp, D, I + (D == 0 ? sd : 0), J + (D == 1 ? sd : 0), K + (D == 2 ? sd : 0));
#pragma unroll
    for (int q = 0; q < 5; ++q) Rt[q] = sd ? Rt[q] - corr[f * 5 + q] : Rt[q] + corr[f * 5 + q];
  }
}

// The corrections of every marked cell added to the RHS R (interior layout [5][Nzp][Nyp][Nxp]), one lane per
// cell in index order; the marks are cleared as they are used.
template <typename T>
__global__ __launch_bounds__(256) void k_heavy_apply(KP kp, T* __restrict__ R, unsigned char* __restrict__ mark,
                                                     const T* __restrict__ corr, const unsigned int* pflag,
                                                     unsigned int pgen) {
  if (!heavy_possible(pflag, pgen)) return;  // nothing listed (or marked) this stage
  const long long rn = hmark_count(kp), sxy = (long long)kp.Nxp * kp.Nyp;
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < rn; c += (long long)gridDim.x * 256) {
    const unsigned int mb = mark[c];
    if (!mb) continue;
    const int I = (int)(c % kp.Nxp), J = (int)((c / kp.Nxp) % kp.Nyp), K = (int)(c / sxy);
    T Rt[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) Rt[q] = R[c + q * rn];
    heavy_add_mark(kp, mb, corr, I, J, K, Rt);
#pragma unroll
    for (int q = 0; q < 5; ++q) R[c + q * rn] = Rt[q];
    mark[c] = 0;
  }
}
template <typename T>
inline void launch_heavy_apply(const KP& kp, T* R, const HMark& hm, const T* corr, hipStream_t s,
                               const unsigned int* pflag, unsigned int pgen) {
  const long long nb = (hmark_count(kp) + 255) / 256;
  hipLaunchKernelGGL((k_heavy_apply<T>), dim3((unsigned int)(nb < 8192 ? (nb ? nb : 1) : 8192)), dim3(256), 0, s, kp,
                     R, reinterpret_cast<unsigned char*>(hm.mark), corr, pflag, pgen);
}

// Clears n 8-byte words (the no-allocation entry's scratch).
void clear_hidx(unsigned long long* idx, long long n, hipStream_t s);

// The marks of one heavy-face evaluation (all zero between evaluations: the apply pass clears what it reads).
inline HMark hmarks(unsigned long long* words) { return HMark{reinterpret_cast<unsigned int*>(words)}; }

// Linear id of the calling wave in a launch of blocks of `wpb` waves (wave index wib in its block).
__device__ __forceinline__ unsigned int launch_wave(int wpb, int wib) {
  return ((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * (unsigned int)wpb + (unsigned int)wib;
}

}  // namespace f3d

struct f3d_workspace;

namespace f3d {
template <typename T>
WsView<T> ws_view(f3d_workspace* ws);

// true when the fused sweeps serve f3d_rhs_* / f3d_rk3_stage_* (F3D_RHS_ENGINE=chain selects the
// kernel-by-kernel chain instead, for A/B measurement).
bool use_fused();
// true when the fused stage routes the viscous fluxes through the workspace's Fv arrays (LES models, or
// F3D_VISC_FUSED=0) instead of computing them inside the update march.
bool fused_needs_fv(const KP& kp);

// Heavy-face list layouts: the finite-difference sweeps of bulk element size es, the finite-volume
// sweeps, and the adjoint's list kernels.  The workspace holds the largest of the three.
HList fdm_heavy_layout(const KP& kp, int es);
HList fvm_heavy_layout(const KP& kp, int es);
HList adj_heavy_layout(const KP& kp);
// F3D_OK when the workspace view holds layout h, else F3D_E_ARG with a message.
template <typename T>
int check_hlist(const WsView<T>& ws, const HList& h, const char* who);

// RHS into R (U layout, interior, double): R = J*(-dF + dFv).
template <typename T>
int fused_rhs(const KP& kp, const WsView<T>& ws, const T* Q, const MetricPtrs<T>& m, const double* J, T* phi,
              double* R, hipStream_t s);

// One RK3 stage up to and including c2Prim of the interior (ghost fill done by the caller): reads the
// primitive state Qin, writes the new one into Qout.  Qout must differ from Qin when
// stage_needs_separate_q(kp) (the viscous divergence is then computed in the update march, which
// still reads Qin of neighbouring columns while other blocks write Qout).
template <typename T>
int fused_stage(const KP& kp, const WsView<T>& ws, int KRK, double* U, double* Un, const T* Qin, T* Qout,
                const MetricPtrs<T>& m, const double* J, double dt, hipStream_t s, double* Uo = nullptr);
bool stage_needs_separate_q(const KP& kp);

// Building blocks shared by the two inviscid branches.  fused_sensor: shock sensor into phi (and the
// heavy-face counter cleared).  fused_update: viscous divergence + conservative update from the
// inviscid increments in ws.R; mode 0 writes the RHS into Rout, mode KRK = 1..3 performs the stage
// update (+ RK3 blend + c2Prim of the interior into Qout).
template <typename T>
void fused_sensor(const KP& kp, const WsView<T>& ws, const T* Q, T* phi, hipStream_t s);
template <typename T>
// Uo != nullptr (stages): the updated conserved state goes to Uo and U is only read.
int fused_update(const KP& kp, const WsView<T>& ws, int mode, double* U, double* Un, const T* Qin, T* Qout,
                 const MetricPtrs<T>& m, const double* J, const T* phi, double* Rout, double dt, hipStream_t s,
                 double* Uo = nullptr);

// Finite-volume branch (f3d_fvm.hip): RHS and stage as fused_rhs / fused_stage, reconstructing the
// conserved state U (ghost cells included); geo[3]: face geometry arrays (f3d_fvm_geometry_*).
template <typename T>
int fvm_fused_rhs(const KP& kp, const WsView<T>& ws, const T* Q, const double* U, const MetricPtrs<T>& m,
                  const double* J, const T* const* geo, T* phi, double* R, hipStream_t s);
template <typename T>
int fvm_fused_stage(const KP& kp, const WsView<T>& ws, int KRK, double* U, double* Un, const T* Qin, T* Qout,
                    const MetricPtrs<T>& m, const double* J, const T* const* geo, double dt, hipStream_t s);

// Stage / multi-step drivers of f3d_stage.hip (with the ghost procedure of the output state).
// geo == nullptr: finite-difference branch; otherwise finite-volume (p->fvm) with these face geometry
// arrays, and the ghost procedure refills U's ghost cells as well.
template <typename T>
// Uo != nullptr (fused finite-difference path only): the stage writes its conserved state to Uo and
// leaves U as it was.
int stage_io(const f3d_params* p, f3d_workspace* ws, int KRK, double* U, double* Un, const T* Qin, T* Qout,
             const T* const* metrics, const double* J, double dt, const T* inflow, hipStream_t s,
             const T* const* geo = nullptr, double* Uo = nullptr);
template <typename T>
int steps_io(const f3d_params* p, f3d_workspace* ws, int nsteps, double* U, double* Un, T* Q,
             const T* const* metrics, const double* J, double dt, const T* inflow, hipStream_t s,
             const T* const* geo = nullptr);
// One FP64 RK3 step of the finite-difference branch as steps_io that also leaves the conserved state
// after stage 1 in U1 and after stage 2 in U2 (the adjoint's stage states); third = false stops after
// stage 2.  Fused path: the stages write U1, U2 and then Uout (U when Uout is null) directly, U is only
// read.  Other paths: in place on U with copies into U1, U2 (Uout must be null; after third = false U
// holds the second stage's state).  Q: the primitive state of U on entry, of the step's result on return
// (of the second stage when third = false).  No filtering block.
int step_keep_stages(const f3d_params* p, f3d_workspace* ws, double* U, double* Un, double* Q,
                     const double* const* metrics, const double* J, double dt, const double* inflow, double* U1,
                     double* U2, hipStream_t s, bool third = true, double* Uout = nullptr);
// One fused FP64 step that keeps the stage states where the caller wants them: stage 1 reads Qin and
// writes U^1 to U1 and its primitive state (ghosts filled) to Q1, stage 2 reads (U1, Q1) and writes
// (U2, Q2), stage 3 reads (U2, Q2) and writes U^{n+1} to U and its primitive state to Qnext.  The Q's are
// the states the stages themselves form (c2Prim in the update, then the ghost fill).  Fused engine only.
int step_keep_q(const f3d_params* p, f3d_workspace* ws, double* U, double* Un, const double* Qin,
                const double* const* metrics, const double* J, double dt, const double* inflow, double* U1,
                double* U2, double* Q1, double* Q2, double* Qnext, hipStream_t s);
}  // namespace f3d
