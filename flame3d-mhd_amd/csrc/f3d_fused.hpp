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
  unsigned long long* heavy; // compacted non-UP7 faces: (right cell << 4) | (scheme << 2) | dir
  unsigned int* hcount;      // number of entries in `heavy`
  unsigned long long hcap;   // capacity of `heavy` (total face count of the three directions)
  double* ldt;               // per-cell dt of local time stepping, cell-shaped (written at stage 1)
};

// Block coordinates with an XCD-aware order.  Workgroups are dispatched round-robin over the 8 XCDs
// (XCD = linear workgroup id mod 8) and each XCD has its own L2; giving every XCD one contiguous
// range of the grid (x fastest) keeps the neighbouring blocks whose halo rows a block re-reads in
// the same L2.  remap = 0 returns the plain blockIdx.
struct Blk {
  int x, y, z;
};
__device__ __forceinline__ Blk xcd_block(int remap) {
  if (!remap) return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
  const unsigned gx = gridDim.x, gy = gridDim.y, nb = gx * gy * gridDim.z;
  const unsigned id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned xcd = id & 7u, r = id >> 3, q = nb >> 3, rem = nb & 7u;
  const unsigned L = xcd * q + (xcd < rem ? xcd : rem) + r;
  return Blk{(int)(L % gx), (int)((L / gx) % gy), (int)(L / (gx * gy))};
}

// 1 unless F3D_XCD_REMAP=0 (A/B measurement of the block order).
int xcd_remap();

// Wave-aggregated append to the compacted heavy-face list: one atomic per wave.
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

}  // namespace f3d

struct f3d_workspace;

namespace f3d {
template <typename T>
WsView<T> ws_view(f3d_workspace* ws);

// true when the fused sweeps serve f3d_rhs_* / f3d_rk3_stage_* (F3D_RHS_ENGINE=chain selects the
// kernel-by-kernel chain instead, for A/B measurement).
bool use_fused();

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
                const MetricPtrs<T>& m, const double* J, double dt, hipStream_t s);
bool stage_needs_separate_q(const KP& kp);

// Building blocks shared by the two inviscid branches.  fused_sensor: shock sensor into phi (and the
// heavy-face counter cleared).  fused_update: viscous divergence + conservative update from the
// inviscid increments in ws.R; mode 0 writes the RHS into Rout, mode KRK = 1..3 performs the stage
// update (+ RK3 blend + c2Prim of the interior into Qout).
template <typename T>
void fused_sensor(const KP& kp, const WsView<T>& ws, const T* Q, T* phi, hipStream_t s);
template <typename T>
int fused_update(const KP& kp, const WsView<T>& ws, int mode, double* U, double* Un, const T* Qin, T* Qout,
                 const MetricPtrs<T>& m, const double* J, const T* phi, double* Rout, double dt, hipStream_t s);

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
int stage_io(const f3d_params* p, f3d_workspace* ws, int KRK, double* U, double* Un, const T* Qin, T* Qout,
             const T* const* metrics, const double* J, double dt, const T* inflow, hipStream_t s,
             const T* const* geo = nullptr);
template <typename T>
int steps_io(const f3d_params* p, f3d_workspace* ws, int nsteps, double* U, double* Un, T* Q,
             const T* const* metrics, const double* J, double dt, const T* inflow, hipStream_t s,
             const T* const* geo = nullptr);
// One FP64 RK3 step of the finite-difference branch as steps_io, copying the conserved state after
// stage 1 into U1 and after stage 2 into U2 (the adjoint's stage states); third = false stops after
// stage 2 (U, Q then hold the second stage's state).  No filtering block.
int step_keep_stages(const f3d_params* p, f3d_workspace* ws, double* U, double* Un, double* Q,
                     const double* const* metrics, const double* J, double dt, const double* inflow, double* U1,
                     double* U2, hipStream_t s, bool third = true);
}  // namespace f3d
