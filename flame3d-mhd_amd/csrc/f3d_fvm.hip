// Finite-volume branch of flowAdvance (solver.jl:21-31; FVM.jl) for gfx950.
//
// Kernel-by-kernel (Eigen_/Conser_reconstruct_{i,j,k}, flowAdvance):
//   k_fvm_face       one lane per face: 8-cell stencil of U, sensor window, geometry (from the mesh
//                    points, or from the face geometry arrays), states (UP7 / WENO7 / WENO5 / minmod,
//                    characteristic or component-wise), HLLC, times the face area, into the
//                    reference's face array.
//   k_fvm_geometry   face unit normals and areas of one direction from the mesh points, once per mesh
//                    (f3d_fvm_geometry_*): the fused path reads 4 values per face instead of the 24
//                    mesh points around it.
// Fused RHS / RK3 stage (the same split as the finite-difference fused path):
//   fused_sensor     shock sensor (f3d_fused.hip), heavy-face counter cleared
//   k_fvm_sweep_x    x faces of a 120-cell row segment staged in LDS (U, sensor): UP7 states + HLLC
//                    for the faces the sensor leaves linear, R = dF_x; other faces contribute zero and
//                    go to the compacted heavy list
//   k_fvm_march      y and z faces: lanes march along the direction with the last 8 cells' U in a
//                    register window, next cell and next face geometry prefetched; R += dF_y, dF_z
//   k_fvm_heavy      the listed faces: full characteristic / component-wise reconstruction + HLLC,
//                    kept at the face's position, then added to the two cells by k_heavy_apply
//                    (fixed order per cell, no floating-point atomics)
//   fused_update     viscous divergence + update + RK3 blend + c2Prim (f3d_fused.hip)
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "f3d_fused.hpp"
#include "f3d_fvm.hpp"

namespace f3d {

namespace {

constexpr int kXW = 128;  // lanes per x row segment (cells incl. the 4+4 halo)

template <typename T>
struct FvmArgs {
  const double* U;
  const T* geo[3];  // face geometry (nx, ny, nz, Area) per direction, face-array layout
  T* R;
  unsigned long long* heavy;
  unsigned int* hcount;
  const unsigned int* pflag = nullptr;  // sensor band flag of this stage (heavy_possible), and its generation
  unsigned int pgen = 0;
  const unsigned char* cls = nullptr;   // per-cell sensor class (scheme_of of the cell's phi)
};

__device__ __forceinline__ long long rix(const KP& kp, int i, int j, int k) {
  // interior RHS layout [5][Nzp][Nyp][Nxp], 0-based interior indices
  return (long long)i + (long long)kp.Nxp * ((long long)j + (long long)kp.Nyp * k);
}

template <typename T>
__device__ __forceinline__ T row_norm(const T* __restrict__ A1, const T* __restrict__ A2, const T* __restrict__ A3,
                                      long long o) {
  const T a = A1[o], b = A2[o], c = A3[o];
  return sqrt(a * a + b * b + c * c);
}

// ------------------------------------------------------------------ one lane per face
template <typename T, int D, bool EIGEN, bool SFROM_A>
__global__ __launch_bounds__(256) void k_fvm_face(KP kp, int ni, int nj, int nk, const T* __restrict__ Q,
                                                  const double* __restrict__ U, const T* __restrict__ phi,
                                                  const T* __restrict__ S, const T* __restrict__ Ax,
                                                  const T* __restrict__ Ay, const T* __restrict__ Az,
                                                  T* __restrict__ F, const double* __restrict__ x,
                                                  const double* __restrict__ y, const double* __restrict__ z,
                                                  const T* __restrict__ geo) {
  const int ii = blockIdx.x * 64 + threadIdx.x, jj = blockIdx.y * 4 + threadIdx.y, kk = blockIdx.z;
  if (ii >= ni || jj >= nj || kk >= nk) return;
  // face (ii, jj, kk) of direction D lies between cells (left) NG+idx-1 and NG+idx along D
  const int i = kp.NG + ii - (D == 0), j = kp.NG + jj - (D == 1), k = kp.NG + kk - (D == 2);
  const FvmDir dd = fvm_dir(kp, D);
  const long long o = cidx(kp, i, j, k);
  const int sch = fvm_scheme<T>(kp, phi, o, dd.st);
  T Sc = 0, Sc1 = 0;
  if (sch != F3D_SCHEME_UP7) {
    Sc = SFROM_A ? row_norm(Ax, Ay, Az, o) : S[o];
    Sc1 = SFROM_A ? row_norm(Ax, Ay, Az, o + dd.st) : S[o + dd.st];
  }
  T out[5];
  const long long fo = ii + (long long)ni * (jj + (long long)nj * kk), fn = (long long)ni * nj * nk;
  if (geo) {
    const T n[3] = {geo[fo], geo[fo + fn], geo[fo + 2 * fn]};
    fvm_face_flux_n<T, EIGEN>(kp, sch, Q, U, o, dd.st, n, geo[fo + 3 * fn], Sc, Sc1, out);
  } else {
    fvm_face_flux<T, EIGEN>(kp, sch, Q, U, x, y, z, o, dd, Sc, Sc1, out);
  }
#pragma unroll
  for (int n = 0; n < 5; ++n) F[fo + n * fn] = out[n];
}

// ------------------------------------------------------------------ face geometry
template <typename T, int D>
__global__ __launch_bounds__(256) void k_fvm_geometry(KP kp, const double* __restrict__ x, const double* __restrict__ y,
                                                      const double* __restrict__ z, T* __restrict__ geo) {
  const int ni = kp.Nxp + (D == 0), nj = kp.Nyp + (D == 1), nk = kp.Nzp + (D == 2);
  const int ii = blockIdx.x * 64 + threadIdx.x, jj = blockIdx.y * 4 + threadIdx.y, kk = blockIdx.z;
  if (ii >= ni || jj >= nj || kk >= nk) return;
  const FvmDir dd = fvm_dir(kp, D);
  const long long o = cidx(kp, kp.NG + ii - (D == 0), kp.NG + jj - (D == 1), kp.NG + kk - (D == 2));
  double Pc[6], Pn[6];
  fvm_partials(x, y, z, o, dd.sa, dd.sb, Pc);
  fvm_partials(x, y, z, o + dd.st, dd.sa, dd.sb, Pn);
  T n[3], A;
  fvm_geom<T>(Pc, Pn, dd.flip, n, &A);
  const long long fo = ii + (long long)ni * (jj + (long long)nj * kk), fn = (long long)ni * nj * nk;
  geo[fo] = n[0];
  geo[fo + fn] = n[1];
  geo[fo + 2 * fn] = n[2];
  geo[fo + 3 * fn] = A;
}

// ------------------------------------------------------------------ fused x sweep (LDS rows)
template <typename T>
__global__ __launch_bounds__(256) void k_fvm_sweep_x(KP kp, FvmArgs<T> a, int L, HSeg hs) {
  __shared__ T sU[5][2][kXW];
  __shared__ unsigned char scl[2][kXW];
  __shared__ T sF[5][2][kXW];
  const int l = threadIdx.x, r = threadIdx.y;
  const int jj = blockIdx.y * 2 + r, kk = blockIdx.z;
  const bool rowok = jj < kp.Nyp;
  const int j = kp.NG + (rowok ? jj : kp.Nyp - 1), k = kp.NG + kk;
  const int f0 = blockIdx.x * L;
  const int ncell = min(L, kp.Nxp - f0);
  const int nface = ncell + 1;
  const int xc = kp.NG + f0 - 4 + l;  // cell this lane stages
  const long long nc = kp.nc;
  // lane l handles the face between cells l-1 and l of the segment (face index f); its geometry is
  // requested before the row is staged
  const bool fact = l >= 4 && l < 4 + nface;
  const bool hp = heavy_possible(a.pflag, a.pgen);  // some face of the stage may be off UP7
  const int f = f0 + l - 4;
  T gn[4] = {0, 0, 0, 0};
  if (fact) {
    const long long go = fidx(kp, 0, f, rowok ? jj : kp.Nyp - 1, kk), gs = fcount(kp, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) gn[q] = a.geo[0][go + q * gs];
  }
  if (l < nface + 7) {
    const long long o = cidx(kp, xc, j, k);
#pragma unroll
    for (int n = 0; n < 5; ++n) sU[n][r][l] = (T)a.U[o + n * nc];
    scl[r][l] = hp ? a.cls[o] : (unsigned char)0;
  }
  __syncthreads();
  int sch = F3D_SCHEME_UP7;
  if (fact) {
    if (hp) {  // the largest sensor class of the face's six cells
      sch = scl[r][l - 3];
#pragma unroll
      for (int s = -2; s <= 2; ++s) sch = max(sch, (int)scl[r][l + s]);
    }
    T Fl[5] = {0, 0, 0, 0, 0};
    if (sch == F3D_SCHEME_UP7) {
      T W[8][5];
#pragma unroll
      for (int s = 0; s < 8; ++s)
#pragma unroll
        for (int n = 0; n < 5; ++n) W[s][n] = sU[n][r][l - 4 + s];
      T UL[5], UR[5];
      fvm_up7_states<T>(kp, W, UL, UR);
      fvm_hllc<T>(kp, UL, UR, gn[0], gn[1], gn[2], Fl);
#pragma unroll
      for (int n = 0; n < 5; ++n) Fl[n] *= gn[3];
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) sF[n][r][l] = Fl[n];
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
    const long long ro = rix(kp, f, jj, kk), rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
#pragma unroll
    for (int n = 0; n < 5; ++n) a.R[ro + n * rn] = sF[n][r][l] - sF[n][r][l + 1];
  }
}

// ------------------------------------------------------------------ fused y / z march
// Lane = interior i (64 consecutive), threadIdx.y = transverse line; a block marches a segment of
// L cells along D.  W[s] holds U of the last 8 cells streamed; once 8 are in, W[0..7] is the stencil
// c-3..c+4 of the face between W[3] and W[4] (face f = f0 + step - 7).  The next cell's U and the
// next face's geometry are requested a step ahead.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_fvm_march(KP kp, FvmArgs<T> a, int L, HSeg hs) {
  const int ii = blockIdx.x * 64 + threadIdx.x;
  const int tt = blockIdx.y * 4 + threadIdx.y;
  const int Nt2 = D == 1 ? kp.Nzp : kp.Nyp;
  const int Np = D == 1 ? kp.Nyp : kp.Nzp;
  const bool act = ii < kp.Nxp && tt < Nt2;
  const int i = kp.NG + min(ii, kp.Nxp - 1), t = kp.NG + min(tt, Nt2 - 1);
  const int f0 = blockIdx.z * L;
  const int ncell = min(L, Np - f0);
  const int nstep = ncell + 8;
  const long long st = dstride(kp, D), nc = kp.nc;
  const long long obase = (D == 1 ? cidx(kp, i, 0, t) : cidx(kp, i, t, 0)) + (long long)(kp.NG + f0 - 4) * st;
  const long long rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
  const int ic = i - kp.NG, tc = t - kp.NG;
  const long long gs = fcount(kp, D);
  const long long gstride = D == 1 ? (long long)kp.Nxp : (long long)kp.Nxp * kp.Nyp;  // next face along D
  const long long gbase = D == 1 ? fidx(kp, 1, ic, f0, tc) : fidx(kp, 2, ic, tc, f0);
  T W[8][5];
  int ph[8];  // sensor classes of the window's cells
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    ph[s] = 0;
#pragma unroll
    for (int n = 0; n < 5; ++n) W[s][n] = 0;
  }
  T Fprev[5] = {0, 0, 0, 0, 0};
  T Rn[5] = {0, 0, 0, 0, 0};  // RHS of the row the next step closes, loaded a step ahead
  // the wave's run of the heavy list: at most one face per lane and completed face (<= 64 (L + 1))
  const unsigned int hw = launch_wave(4, threadIdx.y);
  const unsigned long long hrun = hs.base + (unsigned long long)hw * hs.cap;
  unsigned int hcnt = 0;
  T gn[4] = {0, 0, 0, 0};  // geometry of the next face to complete
  double nu[5];
  int nph;
  const bool hp = heavy_possible(a.pflag, a.pgen);  // some face of the stage may be off UP7
#pragma unroll
  for (int n = 0; n < 5; ++n) nu[n] = a.U[obase + n * nc];
  nph = hp ? (int)a.cls[obase] : 0;
#pragma unroll 1
  for (int step = 0; step < nstep; ++step) {
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      ph[s] = ph[s + 1];
#pragma unroll
      for (int n = 0; n < 5; ++n) W[s][n] = W[s + 1][n];
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) W[7][n] = (T)nu[n];
    ph[7] = nph;
    if (step + 1 < nstep) {
      const long long o = obase + (long long)(step + 1) * st;
#pragma unroll
      for (int n = 0; n < 5; ++n) nu[n] = a.U[o + n * nc];
      if (hp) nph = a.cls[o];
    }
    T Rc[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) Rc[n] = Rn[n];
    if (step >= 7 && step + 1 < nstep && act) {  // row f0 + step - 7, closed at step + 1
      const long long ro = D == 1 ? rix(kp, ii, f0 + step - 7, tt) : rix(kp, ii, tt, f0 + step - 7);
#pragma unroll
      for (int n = 0; n < 5; ++n) Rn[n] = a.R[ro + n * rn];
    }
    if (step >= 6) {
      const T g0 = gn[0], g1 = gn[1], g2 = gn[2], g3 = gn[3];  // face f0 + step - 7 (step >= 7)
      if (step - 6 <= ncell) {  // request face f0 + step - 6
        const long long go = gbase + (long long)(step - 6) * gstride;
#pragma unroll
        for (int q = 0; q < 4; ++q) gn[q] = a.geo[D][go + q * gs];
      }
      if (step >= 7) {  // face f0 + step - 7 complete
        const int fi = step - 7;
        const int f = f0 + fi;
        int sch = F3D_SCHEME_UP7;
        if (hp) {
          sch = ph[1];
#pragma unroll
          for (int s = 2; s <= 6; ++s) sch = max(sch, ph[s]);
        }
        T Fc[5] = {0, 0, 0, 0, 0};
        if (sch == F3D_SCHEME_UP7) {
          T UL[5], UR[5];
          fvm_up7_states<T>(kp, W, UL, UR);
          fvm_hllc<T>(kp, UL, UR, g0, g1, g2, Fc);
#pragma unroll
          for (int n = 0; n < 5; ++n) Fc[n] *= g3;
        }
        if (hp)
          run_append(a.heavy, hrun, hcnt, act && (fi < L || f == Np) && sch != F3D_SCHEME_UP7,
                   ((unsigned long long)(obase + (long long)(fi + 4) * st) << 4) |
                       ((unsigned long long)sch << 2) | (unsigned long long)D);
        if (fi >= 1 && act) {
          const long long ro = D == 1 ? rix(kp, ii, f - 1, tt) : rix(kp, ii, tt, f - 1);
#pragma unroll
          for (int n = 0; n < 5; ++n) a.R[ro + n * rn] = Rc[n] + (Fprev[n] - Fc[n]);
        }
#pragma unroll
        for (int n = 0; n < 5; ++n) Fprev[n] = Fc[n];
      }
    }
  }
  if (threadIdx.x == 0) a.hcount[hs.first + hw] = hcnt;
}

// ------------------------------------------------------------------ fused x + y march
// Lanes = a window of FXW cells along x (4 halo cells each side, FXW - 8 interior columns), threadIdx.y = a
// z line; the block marches a segment of L rows along y.  Per step the new row's U and sensor go to LDS and
// the x faces of the row are formed there (k_fvm_sweep_x's states and HLLC), kept four steps in an LDS
// ring; the y faces come from the column's 8-cell U window (k_fvm_march).  When a y face closes row r,
//   R[r] = (F_x,left - F_x,right) + (F_y,prev - F_y,new),
// the x sweep's store and the y march's update in one expression: R is written once (not written, read and
// written again) and U, the sensor and the row's geometry are read once.  The x faces go to one run of the
// heavy list per wave, the y faces to another.
constexpr int FXW = 64;         // lanes of a row window (one wave)
constexpr int FXI = FXW - 8;    // interior columns per window
constexpr int FXL = 2;          // z lines (waves) per block

template <typename T>
__global__ __launch_bounds__(FXW * FXL, 2) void k_fvm_sweep_xy(KP kp, FvmArgs<T> a, int L, HSeg hx, HSeg hy) {
  __shared__ T sU[5][FXL][FXW];
  __shared__ unsigned char scl[FXL][FXW];
  __shared__ T sF[4][5][FXL][FXW];  // x-face fluxes of the last four rows (slot = step & 3)
  const int l = threadIdx.x, r = threadIdx.y;
  const int fx0 = blockIdx.x * FXI;
  const int ncx = min(FXI, kp.Nxp - fx0);  // interior columns of the window
  const int nfx = ncx + 1;                 // x faces formed, fx0 .. fx0 + ncx (lanes 4 .. 4 + ncx)
  const int tt = blockIdx.y * FXL + r;
  const bool lineok = tt < kp.Nzp;
  const int tc = min(tt, kp.Nzp - 1);
  const int k = kp.NG + tc;
  const int x = min(kp.NG + fx0 - 4 + l, kp.Nxt - 1);  // the lane's cell (lanes past the stencil clamped)
  const int ii = fx0 + l - 4;                           // interior column index of the lane
  const bool yin = l >= 4 && l < 4 + ncx;               // interior column: its y line is updated
  const bool act = lineok && yin;
  const bool fact = l >= 4 && l < 4 + nfx;              // the lane forms the x face between cells l-1, l
  const int f0 = blockIdx.z * L;
  const int ncell = min(L, kp.Nyp - f0);
  const int nstep = ncell + 8;
  const long long st = kp.Nxt, nc = kp.nc;
  const long long obase = cidx(kp, x, kp.NG + f0 - 4, k);
  const long long rn = (long long)kp.Nxp * kp.Nyp * kp.Nzp;
  const long long gs0 = fcount(kp, 0), gs1 = fcount(kp, 1);
  const long long gy0 = fidx(kp, 1, yin ? ii : 0, f0, tc);  // geometry of this column's y face f0
  T W[8][5];
  int ph[8];  // sensor classes of the window's cells
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    ph[q] = 0;
#pragma unroll
    for (int n = 0; n < 5; ++n) W[q][n] = 0;
  }
  T Fprev[5] = {0, 0, 0, 0, 0};
  T gn[4] = {0, 0, 0, 0};  // geometry of the next y face to complete
  T gx[4] = {0, 0, 0, 0};  // geometry of the x face of the next interior row
  const unsigned int hw = launch_wave(FXL, r);
  const unsigned long long xrun = hx.base + (unsigned long long)hw * hx.cap;
  const unsigned long long yrun = hy.base + (unsigned long long)hw * hy.cap;
  unsigned int xcnt = 0, ycnt = 0;
  double nu[5];
  int nph;
  const bool hp = heavy_possible(a.pflag, a.pgen);  // some face of the stage may be off UP7
#pragma unroll
  for (int n = 0; n < 5; ++n) nu[n] = a.U[obase + n * nc];
  nph = hp ? (int)a.cls[obase] : 0;
#pragma unroll 1
  for (int step = 0; step < nstep; ++step) {
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      ph[q] = ph[q + 1];
#pragma unroll
      for (int n = 0; n < 5; ++n) W[q][n] = W[q + 1][n];
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) W[7][n] = (T)nu[n];
    ph[7] = nph;
    const bool xstep = step >= 4 && step < 4 + ncell;  // this step's row is interior: form its x faces
    const T gx0 = gx[0], gx1 = gx[1], gx2 = gx[2], gx3 = gx[3];
    if (step + 1 < nstep) {
      const long long o = obase + (long long)(step + 1) * st;
#pragma unroll
      for (int n = 0; n < 5; ++n) nu[n] = a.U[o + n * nc];
      if (hp) nph = a.cls[o];
      if (fact && lineok && step + 1 >= 4 && step + 1 < 4 + ncell) {
        const long long go = fidx(kp, 0, ii, f0 + step + 1 - 4, tc);
#pragma unroll
        for (int q = 0; q < 4; ++q) gx[q] = a.geo[0][go + q * gs0];
      }
    }
    if (xstep) {
#pragma unroll
      for (int n = 0; n < 5; ++n) sU[n][r][l] = W[7][n];
      scl[r][l] = (unsigned char)ph[7];
    }
    if (step >= 6) {
      const T g0 = gn[0], g1 = gn[1], g2 = gn[2], g3 = gn[3];  // face f0 + step - 7 (step >= 7)
      if (act && step - 6 <= ncell) {  // request face f0 + step - 6
        const long long go = gy0 + (long long)(step - 6) * kp.Nxp;
#pragma unroll
        for (int q = 0; q < 4; ++q) gn[q] = a.geo[1][go + q * gs1];
      }
      if (step >= 7) {  // y face f0 + step - 7 complete: it closes row f - 1
        const int fi = step - 7;
        const int f = f0 + fi;
        int sch = F3D_SCHEME_UP7;
        if (hp) {
          sch = ph[1];
#pragma unroll
          for (int q = 2; q <= 6; ++q) sch = max(sch, ph[q]);
        }
        T Fc[5] = {0, 0, 0, 0, 0};
        if (sch == F3D_SCHEME_UP7) {
          T UL[5], UR[5];
          fvm_up7_states<T>(kp, W, UL, UR);
          fvm_hllc<T>(kp, UL, UR, g0, g1, g2, Fc);
#pragma unroll
          for (int n = 0; n < 5; ++n) Fc[n] *= g3;
        }
        if (hp)
          run_append(a.heavy, yrun, ycnt, act && (fi < L || f == kp.Nyp) && sch != F3D_SCHEME_UP7,
                   ((unsigned long long)(obase + (long long)(fi + 4) * st) << 4) | ((unsigned long long)sch << 2) | 1ULL);
        if (fi >= 1 && act) {  // row f - 1: its x faces were formed four steps ago (slot step & 3)
          const long long ro = rix(kp, ii, f - 1, tc);
          const int sl = step & 3;
#pragma unroll
          for (int n = 0; n < 5; ++n) a.R[ro + n * rn] = (sF[sl][n][r][l] - sF[sl][n][r][l + 1]) + (Fprev[n] - Fc[n]);
        }
#pragma unroll
        for (int n = 0; n < 5; ++n) Fprev[n] = Fc[n];
      }
    }
    __syncthreads();  // this row's U and sensor are in LDS; the ring slot read above is free
    if (xstep) {
      const int sl = step & 3;
      int sch = F3D_SCHEME_UP7;
      if (fact) {
        if (hp) {
          sch = scl[r][l - 3];
#pragma unroll
          for (int q = -2; q <= 2; ++q) sch = max(sch, (int)scl[r][l + q]);
        }
        T Fl[5] = {0, 0, 0, 0, 0};
        if (sch == F3D_SCHEME_UP7) {
          T Wx[8][5];
#pragma unroll
          for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int n = 0; n < 5; ++n) Wx[q][n] = sU[n][r][l - 4 + q];
          T UL[5], UR[5];
          fvm_up7_states<T>(kp, Wx, UL, UR);
          fvm_hllc<T>(kp, UL, UR, gx0, gx1, gx2, Fl);
#pragma unroll
          for (int n = 0; n < 5; ++n) Fl[n] *= gx3;
        }
#pragma unroll
        for (int n = 0; n < 5; ++n) sF[sl][n][r][l] = Fl[n];
      }
      const int j = kp.NG + f0 + step - 4;
      if (hp)
        run_append(a.heavy, xrun, xcnt, lineok && fact && (ii < fx0 + FXI || ii == kp.Nxp) && sch != F3D_SCHEME_UP7,
                 ((unsigned long long)cidx(kp, kp.NG + ii, j, k) << 4) | ((unsigned long long)sch << 2) | 0ULL);
    }
    __syncthreads();  // the row's U is read before the next row overwrites it
  }
  if (l == 0) {
    a.hcount[hx.first + hw] = xcnt;
    a.hcount[hy.first + hw] = ycnt;
  }
}

// ------------------------------------------------------------------ heavy faces
// Every listed face (one launch, all directions): the full characteristic / component-wise
// reconstruction + HLLC flux (the sweep gave these faces zero), kept at the face's position and marked in its
// two cells (heavy_keep); k_heavy_apply adds each cell's fluxes in a fixed order (no floating-point atomics:
// the same summation order on every evaluation).
template <typename T, bool EIGEN>
__global__ __launch_bounds__(64) void k_fvm_heavy(KP kp, FvmArgs<T> a, const T* __restrict__ Q, MetricPtrs<T> m,
                                                  HList hl, T* __restrict__ hcorr, HMark hm) {
  if (!heavy_possible(a.pflag, a.pgen)) return;  // nothing listed this stage
  for_each_heavy(hl, a.hcount, a.heavy, [&](const unsigned long long v) {
    const int D = (int)(v & 3), sch = (int)((v >> 2) & 3);
    const long long o = (long long)(v >> 4);  // right cell of the face
    const FvmDir dd = fvm_dir(kp, D);
    const long long oc = o - dd.st;
    const T Sc = row_norm(m.m[3 * D], m.m[3 * D + 1], m.m[3 * D + 2], oc);
    const T Sc1 = row_norm(m.m[3 * D], m.m[3 * D + 1], m.m[3 * D + 2], o);
    const int ci = (int)(o % kp.Nxt), cj = (int)((o / kp.Nxt) % kp.Nyt), ck = (int)(o / ((long long)kp.Nxt * kp.Nyt));
    const int I = ci - kp.NG, Jj = cj - kp.NG, K = ck - kp.NG;
    const long long go = fidx(kp, D, I, Jj, K), gs = fcount(kp, D);
    const T* g = a.geo[D];
    const T n3[3] = {g[go], g[go + gs], g[go + 2 * gs]};
    T F[5];
    fvm_face_flux_n<T, EIGEN>(kp, sch, Q, a.U, oc, dd.st, n3, g[go + 3 * gs], Sc, Sc1, F);
    const int f = D == 0 ? I : (D == 1 ? Jj : K);
    const int Np = D == 0 ? kp.Nxp : (D == 1 ? kp.Nyp : kp.Nzp);
    heavy_keep(kp, hm, hcorr, D, I, Jj, K, f, Np, F);
  });
}

int seg_len(int Np, int max_len) {
  const int nseg = (Np + max_len - 1) / max_len;
  return (Np + nseg - 1) / nseg;
}

// Segment length of the y / z marches; F3D_FVM_SEG_Y / F3D_FVM_SEG_Z override.
int fvm_seg(int D) {
  static int v[2] = {-1, -1};
  int& r = v[D - 1];
  if (r < 0) {
    const char* e = std::getenv(D == 1 ? "F3D_FVM_SEG_Y" : "F3D_FVM_SEG_Z");
    const int def = D == 1 ? 32 : 16;
    r = e ? std::atoi(e) : def;
    if (r < 8) r = def;
  }
  return r;
}

static dim3 fvm_grid_x(const KP& kp, int L) { return dim3((kp.Nxp + L - 1) / L, (kp.Nyp + 1) / 2, kp.Nzp); }
static dim3 fvm_grid_xy(const KP& kp, int L) {
  return dim3((kp.Nxp + FXI - 1) / FXI, (kp.Nzp + FXL - 1) / FXL, (kp.Nyp + L - 1) / L);
}

// 1 (default): x and y faces in one march (k_fvm_sweep_xy); F3D_FVM_SWEEP_XY=0 runs k_fvm_sweep_x and the y
// march (A/B measurement; the two forms give the same sums in the same order).
int fvm_sweep_xy() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("F3D_FVM_SWEEP_XY");
    v = (e && std::strcmp(e, "0") == 0) ? 0 : 1;
  }
  return v;
}

int cu_count_fvm() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      v <= 0)
    return 256;
  return v;
}

// Row-segment length of the x+y march (as for the finite-difference x+y march): w waves over s resident-wave
// slots take about (L + 8) x ceil(w / s) steps when they overfill the chip and (L + 8) / sqrt(w / s) when they
// fill part of it; minimised over L = 16 .. 64.  F3D_FVM_SEG_Y overrides.
int fvm_xy_seg(const KP& kp, int es) {
  const char* e = std::getenv("F3D_FVM_SEG_Y");
  if (e && std::atoi(e) >= 8) return seg_len(kp.Nyp, std::atoi(e));
  static int wpc[2] = {0, 0};
  int& w = wpc[es == 8];
  if (w == 0) {
    int nb = 0;
    const void* kern = es == 8 ? (const void*)&k_fvm_sweep_xy<double> : (const void*)&k_fvm_sweep_xy<float>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, FXW * FXL, 0) != hipSuccess || nb <= 0) {
      (void)hipGetLastError();
      nb = es == 8 ? 4 : 6;
    }
    w = nb * FXL;
  }
  const double slots = (double)cu_count_fvm() * w;
  const double per_seg = (double)FXL * ((kp.Nxp + FXI - 1) / FXI) * ((kp.Nzp + FXL - 1) / FXL);
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
static dim3 fvm_grid_march(const KP& kp, int D, int L) {
  const int Nt2 = D == 1 ? kp.Nzp : kp.Nyp, Np = D == 1 ? kp.Nyp : kp.Nzp;
  return dim3((kp.Nxp + 63) / 64, (Nt2 + 3) / 4, (Np + L - 1) / L);
}

}  // namespace

HList fvm_heavy_layout(const KP& kp, int es) {
  const int Lx = seg_len(kp.Nxp, kXW - 8), Ly = seg_len(kp.Nyp, fvm_seg(1)), Lz = seg_len(kp.Nzp, fvm_seg(2));
  auto nb = [](const dim3& g) { return (unsigned long long)g.x * g.y * g.z; };
  if (fvm_sweep_xy()) {
    // one wave per z line of a block; per row at most one x face per lane, per closed y face one per lane
    const int Lxy = fvm_xy_seg(kp, es);
    const unsigned long long w = FXL * nb(fvm_grid_xy(kp, Lxy));
    const unsigned long long nw[3] = {w, w, 4 * nb(fvm_grid_march(kp, 2, Lz))};
    const unsigned long long cap[3] = {64ULL * Lxy, 64ULL * (Lxy + 1), 64ULL * (Lz + 1)};
    return make_hlist(nw, cap);
  }
  const unsigned long long nw[3] = {4 * nb(fvm_grid_x(kp, Lx)), 4 * nb(fvm_grid_march(kp, 1, Ly)),
                                    4 * nb(fvm_grid_march(kp, 2, Lz))};
  const unsigned long long cap[3] = {64, 64ULL * (Ly + 1), 64ULL * (Lz + 1)};
  return make_hlist(nw, cap);
}

namespace {

template <typename T>
int fvm_flux(const KP& kp, const WsView<T>& ws, const T* Q, const double* U, const MetricPtrs<T>& m,
             const T* const* geo, const T* phi, hipStream_t s) {
  FvmArgs<T> a{U, {geo[0], geo[1], geo[2]}, ws.R, ws.heavy, ws.hcount};
  a.cls = ws.cls;
  a.pflag = ws.pgen ? ws.pflag : nullptr;
  a.pgen = ws.pgen ? *ws.pgen : 0u;
  const HList hl = fvm_heavy_layout(kp, (int)sizeof(T));
  if (int rc = check_hlist(ws, hl, "fvm sweeps")) return rc;
  if (fvm_sweep_xy()) {
    TimedScope ts("fvm_sweep_xy", s);
    const int L = fvm_xy_seg(kp, (int)sizeof(T));
    hipLaunchKernelGGL(k_fvm_sweep_xy<T>, fvm_grid_xy(kp, L), dim3(FXW, FXL, 1), 0, s, kp, a, L, hl.d[0], hl.d[1]);
  } else {
    {
      TimedScope ts("fvm_sweep_x", s);
      const int L = seg_len(kp.Nxp, kXW - 8);
      hipLaunchKernelGGL(k_fvm_sweep_x<T>, fvm_grid_x(kp, L), dim3(kXW, 2, 1), 0, s, kp, a, L, hl.d[0]);
    }
    {
      TimedScope ts("fvm_sweep_y", s);
      const int L = seg_len(kp.Nyp, fvm_seg(1));
      hipLaunchKernelGGL((k_fvm_march<T, 1>), fvm_grid_march(kp, 1, L), dim3(64, 4, 1), 0, s, kp, a, L, hl.d[1]);
    }
  }
  {
    TimedScope ts("fvm_sweep_z", s);
    const int L = seg_len(kp.Nzp, fvm_seg(2));
    hipLaunchKernelGGL((k_fvm_march<T, 2>), fvm_grid_march(kp, 2, L), dim3(64, 4, 1), 0, s, kp, a, L, hl.d[2]);
  }
  {
    TimedScope ts("fvm_heavy", s);
    const unsigned int nb = hl.items() < 8192u ? (hl.items() ? hl.items() : 1u) : 8192u;
    const HMark hm = hmarks(ws.hidx);
    if (kp.fvm == F3D_FVM_EIGEN)
      hipLaunchKernelGGL((k_fvm_heavy<T, true>), dim3(nb), dim3(64), 0, s, kp, a, Q, m, hl, ws.hflux, hm);
    else
      hipLaunchKernelGGL((k_fvm_heavy<T, false>), dim3(nb), dim3(64), 0, s, kp, a, Q, m, hl, ws.hflux, hm);
    launch_heavy_apply<T>(kp, ws.R, hm, ws.hflux, s, a.pflag, a.pgen);
  }
  return check_launch("fvm sweeps");
}

}  // namespace

template <typename T>
int launch_fvm_face(const KP& kp, int d, const T* Q, const double* U, const T* phi, const T* S, const T* Ax,
                    const T* Ay, const T* Az, T* F, const double* x, const double* y, const double* z, const T* geo,
                    hipStream_t s) {
  TimedScope ts_("fvm_reconstruct", s);
  const int ni = kp.Nxp + (d == 0), nj = kp.Nyp + (d == 1), nk = kp.Nzp + (d == 2);
  const dim3 g((ni + 63) / 64, (nj + 3) / 4, nk), b(64, 4, 1);
  const bool eig = kp.fvm == F3D_FVM_EIGEN;
#define F3D_FVMF(D, E)                                                                                          \
  do {                                                                                                          \
    if (S)                                                                                                      \
      hipLaunchKernelGGL((k_fvm_face<T, D, E, false>), g, b, 0, s, kp, ni, nj, nk, Q, U, phi, S, Ax, Ay, Az, F, \
                         x, y, z, geo);                                                                         \
    else                                                                                                        \
      hipLaunchKernelGGL((k_fvm_face<T, D, E, true>), g, b, 0, s, kp, ni, nj, nk, Q, U, phi, S, Ax, Ay, Az, F,  \
                         x, y, z, geo);                                                                         \
  } while (0)
  if (d == 0) {
    if (eig) F3D_FVMF(0, true);
    else F3D_FVMF(0, false);
  } else if (d == 1) {
    if (eig) F3D_FVMF(1, true);
    else F3D_FVMF(1, false);
  } else {
    if (eig) F3D_FVMF(2, true);
    else F3D_FVMF(2, false);
  }
#undef F3D_FVMF
  return check_launch("fvm_reconstruct");
}

template <typename T>
int launch_fvm_geometry(const KP& kp, const double* x, const double* y, const double* z, T* const* geo, hipStream_t s) {
  TimedScope ts_("fvm_geometry", s);
  for (int d = 0; d < 3; ++d) {
    const int ni = kp.Nxp + (d == 0), nj = kp.Nyp + (d == 1), nk = kp.Nzp + (d == 2);
    const dim3 g((ni + 63) / 64, (nj + 3) / 4, nk), b(64, 4, 1);
    if (d == 0) hipLaunchKernelGGL((k_fvm_geometry<T, 0>), g, b, 0, s, kp, x, y, z, geo[0]);
    if (d == 1) hipLaunchKernelGGL((k_fvm_geometry<T, 1>), g, b, 0, s, kp, x, y, z, geo[1]);
    if (d == 2) hipLaunchKernelGGL((k_fvm_geometry<T, 2>), g, b, 0, s, kp, x, y, z, geo[2]);
  }
  return check_launch("fvm_geometry");
}

template <typename T>
int fvm_fused_rhs(const KP& kp, const WsView<T>& ws, const T* Q, const double* U, const MetricPtrs<T>& m,
                  const double* J, const T* const* geo, T* phi, double* R, hipStream_t s) {
  fused_sensor<T>(kp, ws, Q, phi, s);
  int rc = fvm_flux<T>(kp, ws, Q, U, m, geo, phi, s);
  if (rc) return rc;
  return fused_update<T>(kp, ws, 0, nullptr, nullptr, Q, nullptr, m, J, phi, R, 0.0, s);
}

template <typename T>
int fvm_fused_stage(const KP& kp, const WsView<T>& ws, int KRK, double* U, double* Un, const T* Qin, T* Qout,
                    const MetricPtrs<T>& m, const double* J, const T* const* geo, double dt, hipStream_t s) {
  if (Qin == Qout && stage_needs_separate_q(kp)) {
    set_error("fused stage: the output state must not alias the input state");
    return F3D_E_ARG;
  }
  fused_sensor<T>(kp, ws, Qin, nullptr, s);  // the sweeps read the sensor classes; phi is not kept
  int rc = fvm_flux<T>(kp, ws, Qin, U, m, geo, ws.phi, s);
  if (rc) return rc;
  return fused_update<T>(kp, ws, KRK, U, Un, Qin, Qout, m, J, ws.phi, nullptr, dt, s);
}

#define F3D_FVM_INST(T)                                                                                           \
  template int launch_fvm_face<T>(const KP&, int, const T*, const double*, const T*, const T*, const T*,         \
                                  const T*, const T*, T*, const double*, const double*, const double*, const T*,  \
                                  hipStream_t);                                                                   \
  template int launch_fvm_geometry<T>(const KP&, const double*, const double*, const double*, T* const*,         \
                                      hipStream_t);                                                               \
  template int fvm_fused_rhs<T>(const KP&, const WsView<T>&, const T*, const double*, const MetricPtrs<T>&,       \
                                const double*, const T* const*, T*, double*, hipStream_t);                        \
  template int fvm_fused_stage<T>(const KP&, const WsView<T>&, int, double*, double*, const T*, T*,              \
                                  const MetricPtrs<T>&, const double*, const T* const*, double, hipStream_t);
F3D_FVM_INST(double)
F3D_FVM_INST(float)

}  // namespace f3d
