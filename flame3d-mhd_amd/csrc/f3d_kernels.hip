// Kernel-by-kernel drop-ins for the reference's per-stage launches (solver.jl:416-444,
// 133-188): one thread per cell or face, 64x4 blocks along (i, j) so each wave reads 64
// consecutive i (coalesced rows of the column-major arrays).
#include "f3d_internal.hpp"

namespace f3d {

struct Range3 {
  int i0, j0, k0, ni, nj, nk;
};

#define RANGE_IDX(r)                                       \
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;    \
  const int jj = blockIdx.y * blockDim.y + threadIdx.y;    \
  const int kk = blockIdx.z;                               \
  if (ii >= (r).ni || jj >= (r).nj || kk >= (r).nk) return; \
  const int i = (r).i0 + ii, j = (r).j0 + jj, k = (r).k0 + kk;

static inline dim3 range_grid(const Range3& r) {
  return dim3((unsigned)((r.ni + 63) / 64), (unsigned)((r.nj + 3) / 4), (unsigned)r.nk);
}
static const dim3 kBlock(64, 4, 1);
// x ghost slabs are 2*NG = 8 cells wide along i: 8 x 32 blocks keep every lane busy
static inline dim3 slab_grid_x(const Range3& r) {
  return dim3((unsigned)((r.ni + 7) / 8), (unsigned)((r.nj + 31) / 32), (unsigned)r.nk);
}
static const dim3 kBlockX(8, 32, 1);

// ------------------------------------------------------------------ shockSensor  schemes.jl:1-24
template <typename T>
__global__ __launch_bounds__(256) void k_shock_sensor(KP kp, Range3 r, T* __restrict__ phi,
                                                      const T* __restrict__ Q) {
  RANGE_IDX(r);
  const T* P = Q + 4 * kp.nc;
  const long long o = cidx(kp, i, j, k), sy = kp.Nxt, sz = (long long)kp.Nxt * kp.Nyt;
  phi[o] = sensor_cell<T>(P[o - 1], P[o], P[o + 1], P[o - sy], P[o + sy], P[o - sz], P[o + sz]);
}

// ------------------------------------------------------------------ scheme mask (per face)
template <typename T, int D>
__global__ __launch_bounds__(256) void k_scheme_mask(KP kp, Range3 r, uint8_t* __restrict__ mask,
                                                     const T* __restrict__ phi) {
  RANGE_IDX(r);
  const long long st = dstride(kp, D), o = cidx(kp, i, j, k);
  T m = phi[o - 3 * st];
#pragma unroll
  for (int s = -2; s <= 2; ++s) m = jmax(m, phi[o + s * st]);
  mask[ii + (long long)r.ni * (jj + (long long)r.nj * kk)] = (uint8_t)scheme_of<T>(kp, m);
}

// ------------------------------------------------------------------ fluxSplit_SW / LF / VL / AUSM  split.jl
// SFROM_A: S recomputed as |A| instead of read from an array; SPLIT: F3D_SPLIT_*.
template <typename T, bool SFROM_A, int SPLIT>
__global__ __launch_bounds__(256) void k_flux_split(KP kp, Range3 r, const T* __restrict__ Q, T* __restrict__ Fp,
                                                    T* __restrict__ Fm, const T* __restrict__ S,
                                                    const T* __restrict__ Ax, const T* __restrict__ Ay,
                                                    const T* __restrict__ Az) {
  RANGE_IDX(r);
  const long long o = cidx(kp, i, j, k), nc = kp.nc;
  const T A1 = Ax[o], A2 = Ay[o], A3 = Az[o];
  const T ss = SFROM_A ? sqrt(A1 * A1 + A2 * A2 + A3 * A3) : S[o];
  T fp[5], fm[5];
  split_cell<T, SPLIT>(kp, Q[o], Q[o + nc], Q[o + 2 * nc], Q[o + 3 * nc], Q[o + split_qvar<SPLIT>() * nc], A1, A2, A3,
                       ss, fp, fm);
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    Fp[o + n * nc] = fp[n];
    Fm[o + n * nc] = fm[n];
  }
}

// ------------------------------------------------------------------ advect_{x,y,z}c / advect_{x,y,z}
// CHAR: characteristic advect_*c (schemes.jl:700-1643); otherwise the component-wise advect_*
// (schemes.jl:86-698), which needs neither Q nor the metric rows (S only) and writes the first NV fluxes.
template <typename T, int D, bool SFROM_A, bool CHAR>
__global__ __launch_bounds__(256) void k_advect(KP kp, Range3 r, T* __restrict__ F, const T* __restrict__ phi,
                                                const T* __restrict__ S, const T* __restrict__ Fp,
                                                const T* __restrict__ Fm, const T* __restrict__ Q,
                                                const T* __restrict__ Ax, const T* __restrict__ Ay,
                                                const T* __restrict__ Az, int NV) {
  RANGE_IDX(r);
  const long long st = dstride(kp, D), o = cidx(kp, i, j, k), nc = kp.nc;
  T m = phi[o - 3 * st];
#pragma unroll
  for (int s = -2; s <= 2; ++s) m = jmax(m, phi[o + s * st]);
  const int sch = scheme_of<T>(kp, m);
  auto FP = [&](int s, int n) { return Fp[o + s * st + n * nc]; };
  auto FM = [&](int s, int n) { return Fm[o + s * st + n * nc]; };
  const long long oL = o - st;
  T out[5];
  if (CHAR) {
    const T qL[4] = {Q[oL + nc], Q[oL + 2 * nc], Q[oL + 3 * nc], Q[oL + 5 * nc]};
    const T qR[4] = {Q[o + nc], Q[o + 2 * nc], Q[o + 3 * nc], Q[o + 5 * nc]};
    const T AL[3] = {Ax[oL], Ay[oL], Az[oL]};
    const T AR[3] = {Ax[o], Ay[o], Az[o]};
    const T SL = SFROM_A ? sqrt(AL[0] * AL[0] + AL[1] * AL[1] + AL[2] * AL[2]) : S[oL];
    const T SR = SFROM_A ? sqrt(AR[0] * AR[0] + AR[1] * AR[1] + AR[2] * AR[2]) : S[o];
    face_flux<T, true>(kp, sch, FP, FM, qL, qR, AL, AR, SL, SR, out);
  } else {
    const T SL = SFROM_A ? sqrt(Ax[oL] * Ax[oL] + Ay[oL] * Ay[oL] + Az[oL] * Az[oL]) : S[oL];
    const T SR = SFROM_A ? sqrt(Ax[o] * Ax[o] + Ay[o] * Ay[o] + Az[o] * Az[o]) : S[o];
    face_flux<T, false>(kp, sch, FP, FM, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr,
                        SL, SR, out);
  }
  const long long fo = ii + (long long)r.ni * (jj + (long long)r.nj * kk), fn = (long long)r.ni * r.nj * r.nk;
#pragma unroll
  for (int n = 0; n < 5; ++n)
    if (n < NV) F[fo + n * fn] = out[n];
}

// ------------------------------------------------------------------ viscousFlux  viscous.jl:1-134
template <typename T, int LES>
__global__ __launch_bounds__(256) void k_viscous(KP kp, Range3 r, T* __restrict__ Fvx, T* __restrict__ Fvy,
                                                 T* __restrict__ Fvz, const T* __restrict__ Q, MetricPtrs<T> mp,
                                                 const double* __restrict__ J) {
  RANGE_IDX(r);
  const long long o = cidx(kp, i, j, k), nc = kp.nc;
  const long long st[3] = {1, kp.Nxt, (long long)kp.Nxt * kp.Nyt};
  const T c1 = T(1) / T(12), c2 = T(-2) / T(3);
  T g[4][3];
  const int var[4] = {1, 2, 3, 5};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const T* f = Q + var[q] * nc + o;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      g[q][d] = c1 * (f[-2 * st[d]] - f[2 * st[d]]) + c2 * (f[-st[d]] - f[st[d]]);
  }
  T m[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) m[q] = mp.m[q][o];
  T out[3][4];
  viscous_cell<T, LES>(kp, Q[o + nc], Q[o + 2 * nc], Q[o + 3 * nc], Q[o + 5 * nc], g, m, (T)J[o], out,
                       LES != F3D_LES_NONE ? Q[o] : T(0));
  const int Nv1 = kp.Nxp + kp.NG, Nv2 = kp.Nyp + kp.NG;
  const long long vn = (long long)Nv1 * Nv2 * (kp.Nzp + kp.NG);
  const long long vo = (i - 2) + (long long)Nv1 * ((j - 2) + (long long)Nv2 * (k - 2));
  T* Fo[3] = {Fvx, Fvy, Fvz};
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int n = 0; n < 4; ++n) Fo[d][vo + n * vn] = out[d][n];
}

// ------------------------------------------------------------------ div  div.jl:1-50
// Range is over interior cells (0-based i in [0, Nxp)).
template <typename T>
__global__ __launch_bounds__(256) void k_div(KP kp, Range3 r, double* __restrict__ U, const T* __restrict__ Fx,
                                             const T* __restrict__ Fy, const T* __restrict__ Fz,
                                             const T* __restrict__ Fvx, const T* __restrict__ Fvy,
                                             const T* __restrict__ Fvz, double dt, const double* __restrict__ J,
                                             const double* __restrict__ ldt) {
  RANGE_IDX(r);
  const int NG = kp.NG, Nxp = kp.Nxp, Nyp = kp.Nyp, Nzp = kp.Nzp;
  const long long o = cidx(kp, i + NG, j + NG, k + NG), nc = kp.nc;
  const double Jact = J[o] * (ldt ? ldt[o] : dt);  // div_LTS (div.jl:63) reads the cell's own dt
  const T c1 = T(1) / T(12), c2 = T(-2) / T(3);
  const int Nv1 = Nxp + NG, Nv2 = Nyp + NG;
  const long long vn = (long long)Nv1 * Nv2 * (Nzp + NG);
  const long long vo = (i + 2) + (long long)Nv1 * ((j + 2) + (long long)Nv2 * (k + 2));
  const long long vs[3] = {1, Nv1, (long long)Nv1 * Nv2};
  const T* Fv[3] = {Fvx, Fvy, Fvz};
  const long long fxn = (long long)(Nxp + 1) * Nyp * Nzp, fyn = (long long)Nxp * (Nyp + 1) * Nzp,
                  fzn = (long long)Nxp * Nyp * (Nzp + 1);
  const long long ox = i + (long long)(Nxp + 1) * (j + (long long)Nyp * k);
  const long long oy = i + (long long)Nxp * (j + (long long)(Nyp + 1) * k);
  const long long oz = i + (long long)Nxp * (j + (long long)Nyp * k);
  const long long szf = (long long)Nxp * Nyp;
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const T s = Fx[ox + n * fxn] - Fx[ox + 1 + n * fxn] + Fy[oy + n * fyn] - Fy[oy + Nxp + n * fyn] +
                Fz[oz + n * fzn] - Fz[oz + szf + n * fzn];
    U[o + n * nc] += (double)s * Jact;
  }
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    T s = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T* f = Fv[d] + vo + n * vn;
      s += c1 * (f[-2 * vs[d]] - f[2 * vs[d]]) + c2 * (f[-vs[d]] - f[vs[d]]);
    }
    U[o + (n + 1) * nc] += (double)s * Jact;
  }
}

// ------------------------------------------------------------------ compute_dt  utils.jl:73-100
// Local time step of every interior cell: lts_cfl * min(dx/(u+c), dy/(v+c), dz/(w+c)) with
// dx = (1/J)/S1 etc. (u+c as written, not |u|+c), in double from the bulk-type inputs.  SFROM_A: the
// sizes from the metric rows (|A_d|, the fused stage) instead of the S arrays.
template <typename T, bool SFROM_A>
__global__ __launch_bounds__(256) void k_compute_dt(KP kp, Range3 r, double* __restrict__ dt, const T* __restrict__ Q,
                                                    const double* __restrict__ J, const T* __restrict__ S1,
                                                    const T* __restrict__ S2, const T* __restrict__ S3,
                                                    MetricPtrs<T> mp) {
  RANGE_IDX(r);
  const long long o = cidx(kp, i, j, k), nc = kp.nc;
  const double u = Q[o + nc], v = Q[o + 2 * nc], w = Q[o + 3 * nc], Tt = Q[o + 5 * nc];
  const double V = 1.0 / J[o];
  const double c = sqrt(kp.gamma * kp.Rg * Tt);
  double S[3];
  if (SFROM_A) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T a = mp.m[3 * d][o], b = mp.m[3 * d + 1][o], e = mp.m[3 * d + 2][o];
      S[d] = (double)sqrt(a * a + b * b + e * e);
    }
  } else {
    S[0] = S1[o];
    S[1] = S2[o];
    S[2] = S3[o];
  }
  const double dtx = (V / S[0]) / (u + c), dty = (V / S[1]) / (v + c), dtz = (V / S[2]) / (w + c);
  dt[o] = jmin<double>(jmin<double>(dtx, dty), dtz) * kp.lts_cfl;
}

// ------------------------------------------------------------------ filters  utils.jl:102-316
// pre_x/_y/_z (102-169): the nonlinear filter's sensor along D from the pressure, written on the
// interior and the first ghost layer each side along D (the filter's boundary faces read those), in
// double from the bulk-type pressure.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_pre_sensor(KP kp, Range3 r, double* __restrict__ sc,
                                                    const T* __restrict__ Q, double rth) {
  RANGE_IDX(r);
  const long long st = dstride(kp, D), o = cidx(kp, i, j, k);
  const T* pq = Q + 4 * kp.nc + o;
  const double p1 = pq[-2 * st], p2 = pq[-st], p3 = pq[0], p4 = pq[st], p5 = pq[2 * st];
  const double dp0 = 0.25 * (-p4 + 2 * p3 - p2);
  const double dp1 = 0.25 * (-p5 + 2 * p4 - p3);
  const double dp2 = 0.25 * (-p3 + 2 * p2 - p1);
  const double ri = 0.5 * ((dp0 - dp1) * (dp0 - dp1) + (dp0 - dp2) * (dp0 - dp2)) / (p3 * p3) + 1e-16;
  const double t = 1.0 - rth / ri;
  sc[o] = 0.5 * (t + fabs(t));
}

// filter_x/_y/_z (171-241): U = Un - s0 (sc+ D+ - sc- D-) on the interior, 5 conserved variables.
template <int D>
__global__ __launch_bounds__(256) void k_filter_nl(KP kp, Range3 r, double* __restrict__ U,
                                                   const double* __restrict__ Un, const double* __restrict__ sc,
                                                   double s0) {
  RANGE_IDX(r);
  const long long st = dstride(kp, D), o = cidx(kp, i, j, k), nc = kp.nc;
  const double c1 = -0.210383, c2 = 0.039617;
  const double sc1 = 0.5 * (sc[o] + sc[o + st]);
  const double sc2 = 0.5 * (sc[o] + sc[o - st]);
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const double* v = Un + n * nc + o;
    const double fp = c1 * (v[st] - v[0]) + c2 * (v[2 * st] - v[-st]);
    const double fm = c1 * (v[0] - v[-st]) + c2 * (v[st] - v[-2 * st]);
    U[o + n * nc] = v[0] - s0 * (sc1 * fp - sc2 * fm);
  }
}

// linearFilter_x/_y/_z (243-316): 9-point linear filter of the interior.
template <int D>
__global__ __launch_bounds__(256) void k_filter_lin(KP kp, Range3 r, double* __restrict__ U,
                                                    const double* __restrict__ Un, double s0) {
  RANGE_IDX(r);
  const long long st = dstride(kp, D), o = cidx(kp, i, j, k), nc = kp.nc;
  const double d0 = 0.243527493120, d1 = -0.204788880640, d2 = 0.120007591680, d3 = -0.045211119360,
               d4 = 0.008228661760;
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const double* v = Un + n * nc + o;
    const double f = d0 * v[0] + d1 * (v[-st] + v[st]) + d2 * (v[-2 * st] + v[2 * st]) +
                     d3 * (v[-3 * st] + v[3 * st]) + d4 * (v[-4 * st] + v[4 * st]);
    U[o + n * nc] = v[0] - s0 * f;
  }
}

// ------------------------------------------------------------------ linComb, c2Prim, prim2c
__global__ __launch_bounds__(256) void k_lin_comb(KP kp, Range3 r, double* __restrict__ U,
                                                  const double* __restrict__ Un, int NV, double a, double b) {
  RANGE_IDX(r);
  const long long o = cidx(kp, i, j, k);
  for (int n = 0; n < NV; ++n) U[o + n * kp.nc] = U[o + n * kp.nc] * a + Un[o + n * kp.nc] * b;
}

template <typename T>
__device__ __forceinline__ void c2prim_cell(const KP& kp, const double* __restrict__ U, T* __restrict__ Q,
                                            long long o) {
  const long long nc = kp.nc;
  const double eps = 2.220446049250313e-16;  // machine epsilon of double, the type c2Prim runs in
  const double rho = jmax(U[o], eps);
  const double ri = 1.0 / rho;
  const double u = U[o + nc] * ri, v = U[o + 2 * nc] * ri, w = U[o + 3 * nc] * ri;
  const double ei = jmax(U[o + 4 * nc] - 0.5 * rho * (u * u + v * v + w * w), eps);
  const double p = (kp.gamma - 1) * ei;
  const double Tt = p / (rho * kp.Rg);
  Q[o] = (T)rho;
  Q[o + nc] = (T)u;
  Q[o + 2 * nc] = (T)v;
  Q[o + 3 * nc] = (T)w;
  Q[o + 4 * nc] = (T)p;
  Q[o + 5 * nc] = (T)Tt;
}

template <typename T>
__global__ __launch_bounds__(256) void k_c2prim(KP kp, Range3 r, const double* __restrict__ U, T* __restrict__ Q) {
  RANGE_IDX(r);
  c2prim_cell<T>(kp, U, Q, cidx(kp, i, j, k));
}

template <typename T>
__global__ __launch_bounds__(256) void k_prim2c(KP kp, Range3 r, double* __restrict__ U, const T* __restrict__ Q) {
  RANGE_IDX(r);
  const long long o = cidx(kp, i, j, k), nc = kp.nc;
  const double rho = Q[o], u = Q[o + nc], v = Q[o + 2 * nc], w = Q[o + 3 * nc];
  const double ei = (double)Q[o + 4 * nc] / (kp.gamma - 1);
  U[o] = rho;
  U[o + nc] = u * rho;
  U[o + 2 * nc] = v * rho;
  U[o + 3 * nc] = w * rho;
  U[o + 4 * nc] = ei + 0.5 * rho * (u * u + v * v + w * w);
}

// ------------------------------------------------------------------ ghost cells
// Periodic exchange of one dimension D (mpi.jl:2-103 with one rank): range covers both NG-thick
// ghost slabs (2*NG layers along D) and the full extent of the other two dimensions.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_exchange(KP kp, Range3 r, T* __restrict__ Q, int NV) {
  RANGE_IDX(r);
  const int Np = D == 0 ? kp.Nxp : (D == 1 ? kp.Nyp : kp.Nzp);
  int idx = D == 0 ? i : (D == 1 ? j : k);
  // map the packed range [0, 2NG) onto ghost layers [0,NG) and [Np+NG, Np+2NG)
  const int g = idx >= kp.NG ? idx - kp.NG + Np + kp.NG : idx;
  const int src = g < kp.NG ? g + Np : g - Np;
  const int gi = D == 0 ? g : i, gj = D == 1 ? g : j, gk = D == 2 ? g : k;
  const int si = D == 0 ? src : i, sj = D == 1 ? src : j, sk = D == 2 ? src : k;
  const long long o = cidx(kp, gi, gj, gk), s = cidx(kp, si, sj, sk);
  for (int n = 0; n < NV; ++n) Q[o + n * kp.nc] = Q[s + n * kp.nc];
}

// Ghost fill of dimension D for both faces (fill_x/fill_y/fill_z, boundary.jl:1-148 with the kinds
// of f3d.h); same packed range as k_exchange.  U may be null.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_fill(KP kp, Range3 r, T* __restrict__ Q, double* __restrict__ U,
                                              const T* __restrict__ inflow) {
  RANGE_IDX(r);
  const int NG = kp.NG, Np = D == 0 ? kp.Nxp : (D == 1 ? kp.Nyp : kp.Nzp);
  const int idx = D == 0 ? i : (D == 1 ? j : k);
  const int side = idx >= NG ? 1 : 0;
  const int g = side ? idx - NG + Np + NG : idx;
  const int bc = kp.bc[2 * D + side];
  const int gi = D == 0 ? g : i, gj = D == 1 ? g : j, gk = D == 2 ? g : k;
  const long long o = cidx(kp, gi, gj, gk), nc = kp.nc;
  auto at = [&](int t) { return cidx(kp, D == 0 ? t : i, D == 1 ? t : j, D == 2 ? t : k); };
  double q[6];
  if (bc == F3D_BC_PERIODIC || bc == F3D_BC_ZEROGRAD) {
    const int t = bc == F3D_BC_PERIODIC ? (side ? g - Np : g + Np) : (side ? Np + NG - 1 : NG);
    const long long s = at(t);
#pragma unroll
    for (int n = 0; n < 6; ++n) Q[o + n * nc] = Q[s + n * nc];
    if (U)
#pragma unroll
      for (int n = 0; n < 5; ++n) U[o + n * nc] = U[s + n * nc];
    return;
  } else if (bc == F3D_BC_DIRICHLET) {
#pragma unroll
    for (int n = 0; n < 6; ++n) q[n] = kp.qinf[n];
  } else if (bc == F3D_BC_INFLOW) {
    const long long io = j + (long long)kp.Nyt * k, in = (long long)kp.Nyt * kp.Nzt;
#pragma unroll
    for (int n = 0; n < 6; ++n) q[n] = inflow[io + n * in];
  } else {  // walls: mirror layer (boundary.jl:69), 0-based 2NG-1-g low, 2(Np+NG)-1-g high
    const int mi = side ? 2 * (Np + NG) - 1 - g : 2 * NG - 1 - g;
    const long long s = at(mi);
    q[0] = Q[s];
    q[1] = -(double)Q[s + nc];
    q[2] = -(double)Q[s + 2 * nc];
    q[3] = -(double)Q[s + 3 * nc];
    q[4] = Q[s + 4 * nc];
    q[5] = Q[s + 5 * nc];
    if (bc == F3D_BC_WALL_ISOTHERMAL) {
      q[5] = 2 * kp.Twall - q[5];
      q[0] = q[4] / (kp.Rg * q[5]);
    }
  }
  T qs[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    qs[n] = (T)q[n];
    Q[o + n * nc] = qs[n];
  }
  if (U) {
    const double rho = qs[0], u = qs[1], v = qs[2], w = qs[3], p = qs[4];
    U[o] = rho;
    U[o + nc] = rho * u;
    U[o + 2 * nc] = rho * v;
    U[o + 3 * nc] = rho * w;
    U[o + 4 * nc] = p / (kp.gamma - 1) + 0.5 * rho * (u * u + v * v + w * w);
  }
}

// All ghost cells in one launch.  fill_d of x, then y, then z rewrites its whole slab (the other two
// coordinates over the full array, ghosts included), reading only cells of the same line whose coordinate
// along d is interior (periodic and zero-gradient sources, wall mirrors) or nothing (Dirichlet, inflow).  So
// the final value of a ghost cell is: a z ghost, the z map of the final value of its source (whose x/y ghost
// coordinates, if any, are resolved the same way); a y ghost with interior k, the y map of its source's final
// value; an x ghost with interior j and k, the x map of interior cells.  k_fill_all evaluates that
// composition for the disjoint union of the z slab, the y slab at interior k and the x slab at interior j
// and k, one cell per lane: the same values, with the same arithmetic, as the three k_fill launches, without
// the intermediate writes the later fills overwrite.
template <typename T>
struct GhostVal {
  T q[6];
  double u[5];
};

template <typename T>
__device__ __forceinline__ GhostVal<T> ghost_cell(const KP& kp, const T* Q, const double* U, int i, int j, int k) {
  GhostVal<T> v;
  const long long o = cidx(kp, i, j, k), nc = kp.nc;
#pragma unroll
  for (int n = 0; n < 6; ++n) v.q[n] = Q[o + n * nc];
  if (U) {
#pragma unroll
    for (int n = 0; n < 5; ++n) v.u[n] = U[o + n * nc];
  }
  return v;
}

// fill_d of direction D (k_fill's expressions) for the cell whose coordinate along D is the ghost g, with
// (j, k) its other coordinates as k_fill reads them for the inflow plane; src(t) = final values of the cell
// at coordinate t along D on the same line.
template <typename T, int D, typename Src>
__device__ __forceinline__ GhostVal<T> ghost_map(const KP& kp, const T* __restrict__ inflow, bool withU, int g, int j,
                                                 int k, Src&& src) {
  const int NG = kp.NG, Np = D == 0 ? kp.Nxp : (D == 1 ? kp.Nyp : kp.Nzp);
  const int side = g >= NG ? 1 : 0;
  const int bc = kp.bc[2 * D + side];
  double q[6];
  if (bc == F3D_BC_PERIODIC || bc == F3D_BC_ZEROGRAD) {
    return src(bc == F3D_BC_PERIODIC ? (side ? g - Np : g + Np) : (side ? Np + NG - 1 : NG));
  } else if (bc == F3D_BC_DIRICHLET) {
#pragma unroll
    for (int n = 0; n < 6; ++n) q[n] = kp.qinf[n];
  } else if (bc == F3D_BC_INFLOW) {
    // k_fill's slab index along D stands in the inflow plane index for y and z boundaries
    const int jj = D == 1 ? (side ? g - Np : g) : j, kk = D == 2 ? (side ? g - Np : g) : k;
    const long long io = jj + (long long)kp.Nyt * kk, in = (long long)kp.Nyt * kp.Nzt;
#pragma unroll
    for (int n = 0; n < 6; ++n) q[n] = inflow[io + n * in];
  } else {
    const GhostVal<T> sv = src(side ? 2 * (Np + NG) - 1 - g : 2 * NG - 1 - g);
    q[0] = sv.q[0];
    q[1] = -(double)sv.q[1];
    q[2] = -(double)sv.q[2];
    q[3] = -(double)sv.q[3];
    q[4] = sv.q[4];
    q[5] = sv.q[5];
    if (bc == F3D_BC_WALL_ISOTHERMAL) {
      q[5] = 2 * kp.Twall - q[5];
      q[0] = q[4] / (kp.Rg * q[5]);
    }
  }
  GhostVal<T> v;
#pragma unroll
  for (int n = 0; n < 6; ++n) v.q[n] = (T)q[n];
  if (withU) {
    const double rho = v.q[0], u = v.q[1], w1 = v.q[2], w2 = v.q[3], p = v.q[4];
    v.u[0] = rho;
    v.u[1] = rho * u;
    v.u[2] = rho * w1;
    v.u[3] = rho * w2;
    v.u[4] = p / (kp.gamma - 1) + 0.5 * rho * (u * u + w1 * w1 + w2 * w2);
  }
  return v;
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill_all(KP kp, T* __restrict__ Q, double* __restrict__ U,
                                                  const T* __restrict__ inflow, long long nz, long long ny,
                                                  long long nx) {
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  const int NG = kp.NG, G2 = 2 * NG;
  int i, j, k;
  if (id < nz) {  // z slab: every (i, j)
    const long long r = id / kp.Nxt;
    const int kk = (int)(r / kp.Nyt);
    i = (int)(id % kp.Nxt);
    j = (int)(r % kp.Nyt);
    k = kk < NG ? kk : kk - NG + kp.Nzp + NG;
  } else if (id < nz + ny) {  // y slab at interior k
    const long long e = id - nz, r = e / kp.Nxt;
    const int jj = (int)(r % G2);
    i = (int)(e % kp.Nxt);
    j = jj < NG ? jj : jj - NG + kp.Nyp + NG;
    k = NG + (int)(r / G2);
  } else {  // x slab at interior j and k
    const long long e = id - nz - ny;
    if (e >= nx) return;
    const long long r = e / G2;
    const int ii = (int)(e % G2);
    i = ii < NG ? ii : ii - NG + kp.Nxp + NG;
    j = NG + (int)(r % kp.Nyp);
    k = NG + (int)(r / kp.Nyp);
  }
  const bool withU = U != nullptr;
  auto xg = [&](int c) { return c < NG || c >= NG + kp.Nxp; };
  auto yg = [&](int c) { return c < NG || c >= NG + kp.Nyp; };
  auto fx = [&](int ci, int cj, int ck) {  // x ghost ci, interior cj and ck
    return ghost_map<T, 0>(kp, inflow, withU, ci, cj, ck, [&](int t) { return ghost_cell<T>(kp, Q, U, t, cj, ck); });
  };
  auto fy = [&](int ci, int cj, int ck) {  // y ghost cj, interior ck
    return ghost_map<T, 1>(kp, inflow, withU, cj, cj, ck, [&](int t) {
      return xg(ci) ? fx(ci, t, ck) : ghost_cell<T>(kp, Q, U, ci, t, ck);
    });
  };
  GhostVal<T> v;
  if (k < NG || k >= NG + kp.Nzp) {
    v = ghost_map<T, 2>(kp, inflow, withU, k, j, k, [&](int t) {
      return yg(j) ? fy(i, j, t) : (xg(i) ? fx(i, j, t) : ghost_cell<T>(kp, Q, U, i, j, t));
    });
  } else if (yg(j)) {
    v = fy(i, j, k);
  } else {
    v = fx(i, j, k);
  }
  const long long o = cidx(kp, i, j, k), nc = kp.nc;
#pragma unroll
  for (int n = 0; n < 6; ++n) Q[o + n * nc] = v.q[n];
  if (withU) {
#pragma unroll
    for (int n = 0; n < 5; ++n) U[o + n * nc] = v.u[n];
  }
}

// 0 when F3D_FILL_SPLIT=1: the three per-direction fills instead of k_fill_all (A/B measurement).
static bool fill_all() {
  static const bool v = [] {
    const char* e = std::getenv("F3D_FILL_SPLIT");
    return !(e && std::atoi(e) == 1);
  }();
  return v;
}

// ------------------------------------------------------------------ host launchers
template <typename T>
int launch_shock_sensor(const KP& kp, T* phi, const T* Q, hipStream_t s) {
  TimedScope ts_("shock_sensor", s);
  Range3 r{1, 1, 1, kp.Nxt - 2, kp.Nyt - 2, kp.Nzt - 2};
  hipLaunchKernelGGL(k_shock_sensor<T>, range_grid(r), kBlock, 0, s, kp, r, phi, Q);
  return check_launch("shock_sensor");
}

static Range3 face_range(const KP& kp, int d) {
  return Range3{kp.NG, kp.NG, kp.NG, kp.Nxp + (d == 0), kp.Nyp + (d == 1), kp.Nzp + (d == 2)};
}

template <typename T>
int launch_scheme_mask(const KP& kp, uint8_t* mx, uint8_t* my, uint8_t* mz, const T* phi, hipStream_t s) {
  TimedScope ts_("scheme_mask", s);
  if (mx) {
    Range3 r = face_range(kp, 0);
    hipLaunchKernelGGL((k_scheme_mask<T, 0>), range_grid(r), kBlock, 0, s, kp, r, mx, phi);
  }
  if (my) {
    Range3 r = face_range(kp, 1);
    hipLaunchKernelGGL((k_scheme_mask<T, 1>), range_grid(r), kBlock, 0, s, kp, r, my, phi);
  }
  if (mz) {
    Range3 r = face_range(kp, 2);
    hipLaunchKernelGGL((k_scheme_mask<T, 2>), range_grid(r), kBlock, 0, s, kp, r, mz, phi);
  }
  return check_launch("scheme_mask");
}

template <typename T>
int launch_flux_split(const KP& kp, const T* Q, T* Fp, T* Fm, const T* S, const T* Ax, const T* Ay, const T* Az,
                      hipStream_t s) {
  static const char* names[4] = {"flux_split_sw", "flux_split_lf", "flux_split_vl", "flux_split_ausm"};
  TimedScope ts_(names[kp.split & 3], s);
  Range3 r{0, 0, 0, kp.Nxt, kp.Nyt, kp.Nzt};
#define F3D_SPL(M)                                                                                                 \
  do {                                                                                                             \
    if (S)                                                                                                         \
      hipLaunchKernelGGL((k_flux_split<T, false, M>), range_grid(r), kBlock, 0, s, kp, r, Q, Fp, Fm, S, Ax, Ay, Az); \
    else                                                                                                           \
      hipLaunchKernelGGL((k_flux_split<T, true, M>), range_grid(r), kBlock, 0, s, kp, r, Q, Fp, Fm, S, Ax, Ay, Az);  \
  } while (0)
  switch (kp.split) {
    case F3D_SPLIT_LF: F3D_SPL(F3D_SPLIT_LF); break;
    case F3D_SPLIT_VL: F3D_SPL(F3D_SPLIT_VL); break;
    case F3D_SPLIT_AUSM: F3D_SPL(F3D_SPLIT_AUSM); break;
    default: F3D_SPL(F3D_SPLIT_SW); break;
  }
#undef F3D_SPL
  return check_launch(names[kp.split & 3]);
}

template <typename T>
int launch_advect(const KP& kp, int d, T* F, const T* phi, const T* S, const T* Fp, const T* Fm, const T* Q,
                  const T* Ax, const T* Ay, const T* Az, hipStream_t s, int NV) {
  TimedScope ts_(kp.character ? "advect" : "advect_component", s);
  Range3 r = face_range(kp, d);
#define F3D_ADV(D, C)                                                                                         \
  do {                                                                                                        \
    if (S)                                                                                                    \
      hipLaunchKernelGGL((k_advect<T, D, false, C>), range_grid(r), kBlock, 0, s, kp, r, F, phi, S, Fp, Fm, Q, \
                         Ax, Ay, Az, NV);                                                                     \
    else                                                                                                      \
      hipLaunchKernelGGL((k_advect<T, D, true, C>), range_grid(r), kBlock, 0, s, kp, r, F, phi, S, Fp, Fm, Q,  \
                         Ax, Ay, Az, NV);                                                                     \
  } while (0)
  if (kp.character) {
    if (d == 0) F3D_ADV(0, true);
    else if (d == 1) F3D_ADV(1, true);
    else F3D_ADV(2, true);
  } else {
    if (d == 0) F3D_ADV(0, false);
    else if (d == 1) F3D_ADV(1, false);
    else F3D_ADV(2, false);
  }
#undef F3D_ADV
  return check_launch("advect");
}

template <typename T>
int launch_viscous(const KP& kp, T* Fvx, T* Fvy, T* Fvz, const T* Q, const MetricPtrs<T>& m, const double* J,
                   hipStream_t s) {
  TimedScope ts_("viscous_flux", s);
  Range3 r{2, 2, 2, kp.Nxt - 4, kp.Nyt - 4, kp.Nzt - 4};
  if (kp.les == F3D_LES_SMAG)
    hipLaunchKernelGGL((k_viscous<T, F3D_LES_SMAG>), range_grid(r), kBlock, 0, s, kp, r, Fvx, Fvy, Fvz, Q, m, J);
  else if (kp.les == F3D_LES_WALE)
    hipLaunchKernelGGL((k_viscous<T, F3D_LES_WALE>), range_grid(r), kBlock, 0, s, kp, r, Fvx, Fvy, Fvz, Q, m, J);
  else
    hipLaunchKernelGGL((k_viscous<T, F3D_LES_NONE>), range_grid(r), kBlock, 0, s, kp, r, Fvx, Fvy, Fvz, Q, m, J);
  return check_launch("viscous_flux");
}

template <typename T>
int launch_div(const KP& kp, double* U, const T* Fx, const T* Fy, const T* Fz, const T* Fvx, const T* Fvy,
               const T* Fvz, double dt, const double* J, hipStream_t s, const double* ldt) {
  TimedScope ts_("div", s);
  Range3 r{0, 0, 0, kp.Nxp, kp.Nyp, kp.Nzp};
  hipLaunchKernelGGL(k_div<T>, range_grid(r), kBlock, 0, s, kp, r, U, Fx, Fy, Fz, Fvx, Fvy, Fvz, dt, J, ldt);
  return check_launch("div");
}

template <typename T>
int launch_compute_dt(const KP& kp, double* dt, const T* Q, const double* J, const T* S1, const T* S2, const T* S3,
                      const MetricPtrs<T>* m, hipStream_t s) {
  TimedScope ts_("compute_dt", s);
  Range3 r{kp.NG, kp.NG, kp.NG, kp.Nxp, kp.Nyp, kp.Nzp};
  MetricPtrs<T> mp{};
  if (m) mp = *m;
  if (m)
    hipLaunchKernelGGL((k_compute_dt<T, true>), range_grid(r), kBlock, 0, s, kp, r, dt, Q, J, S1, S2, S3, mp);
  else
    hipLaunchKernelGGL((k_compute_dt<T, false>), range_grid(r), kBlock, 0, s, kp, r, dt, Q, J, S1, S2, S3, mp);
  return check_launch("compute_dt");
}

static Range3 interior_range(const KP& kp) { return Range3{kp.NG, kp.NG, kp.NG, kp.Nxp, kp.Nyp, kp.Nzp}; }

template <typename T>
int launch_pre_sensor(const KP& kp, int d, double* sc, const T* Q, double rth, hipStream_t s) {
  TimedScope ts_("filter_sensor", s);
  Range3 r{kp.NG - (d == 0), kp.NG - (d == 1), kp.NG - (d == 2), kp.Nxp + 2 * (d == 0), kp.Nyp + 2 * (d == 1),
           kp.Nzp + 2 * (d == 2)};
  if (d == 0)
    hipLaunchKernelGGL((k_pre_sensor<T, 0>), range_grid(r), kBlock, 0, s, kp, r, sc, Q, rth);
  else if (d == 1)
    hipLaunchKernelGGL((k_pre_sensor<T, 1>), range_grid(r), kBlock, 0, s, kp, r, sc, Q, rth);
  else
    hipLaunchKernelGGL((k_pre_sensor<T, 2>), range_grid(r), kBlock, 0, s, kp, r, sc, Q, rth);
  return check_launch("filter sensor");
}

int launch_filter(const KP& kp, int d, double* U, const double* Un, const double* sc, double s0, hipStream_t s) {
  TimedScope ts_("filter", s);
  const Range3 r = interior_range(kp);
  if (sc) {
    if (d == 0)
      hipLaunchKernelGGL((k_filter_nl<0>), range_grid(r), kBlock, 0, s, kp, r, U, Un, sc, s0);
    else if (d == 1)
      hipLaunchKernelGGL((k_filter_nl<1>), range_grid(r), kBlock, 0, s, kp, r, U, Un, sc, s0);
    else
      hipLaunchKernelGGL((k_filter_nl<2>), range_grid(r), kBlock, 0, s, kp, r, U, Un, sc, s0);
  } else {
    if (d == 0)
      hipLaunchKernelGGL((k_filter_lin<0>), range_grid(r), kBlock, 0, s, kp, r, U, Un, s0);
    else if (d == 1)
      hipLaunchKernelGGL((k_filter_lin<1>), range_grid(r), kBlock, 0, s, kp, r, U, Un, s0);
    else
      hipLaunchKernelGGL((k_filter_lin<2>), range_grid(r), kBlock, 0, s, kp, r, U, Un, s0);
  }
  return check_launch("filter");
}

template <typename T>
int launch_filter_block(const KP& kp, double* U, double* Un, const T* Q, double* sc, int nonlinear, double rth,
                        double s0, hipStream_t s) {
  for (int d = 0; d < 3; ++d) {
    F3D_HIP(hipMemcpyAsync(Un, U, sizeof(double) * 5 * kp.nc, hipMemcpyDeviceToDevice, s));  // solver.jl:447
    int rc = nonlinear ? launch_pre_sensor<T>(kp, d, sc, Q, rth, s) : F3D_OK;
    if (!rc) rc = launch_filter(kp, d, U, Un, nonlinear ? sc : nullptr, s0, s);
    if (rc) return rc;
  }
  return F3D_OK;
}

int launch_lin_comb(const KP& kp, double* U, const double* Un, int NV, double a, double b, hipStream_t s) {
  TimedScope ts_("lin_comb", s);
  Range3 r{kp.NG, kp.NG, kp.NG, kp.Nxp, kp.Nyp, kp.Nzp};
  hipLaunchKernelGGL(k_lin_comb, range_grid(r), kBlock, 0, s, kp, r, U, Un, NV, a, b);
  return check_launch("lin_comb");
}

template <typename T>
int launch_c2prim(const KP& kp, const double* U, T* Q, hipStream_t s) {
  TimedScope ts_("c2prim", s);
  Range3 r{kp.NG, kp.NG, kp.NG, kp.Nxp, kp.Nyp, kp.Nzp};
  hipLaunchKernelGGL(k_c2prim<T>, range_grid(r), kBlock, 0, s, kp, r, U, Q);
  return check_launch("c2prim");
}

template <typename T>
int launch_prim2c(const KP& kp, double* U, const T* Q, hipStream_t s) {
  TimedScope ts_("prim2c", s);
  Range3 r{0, 0, 0, kp.Nxt, kp.Nyt, kp.Nzt};
  hipLaunchKernelGGL(k_prim2c<T>, range_grid(r), kBlock, 0, s, kp, r, U, Q);
  return check_launch("prim2c");
}

static Range3 slab_range(const KP& kp, int d) {
  return Range3{0, 0, 0, d == 0 ? 2 * kp.NG : kp.Nxt, d == 1 ? 2 * kp.NG : kp.Nyt, d == 2 ? 2 * kp.NG : kp.Nzt};
}

template <typename T>
int launch_exchange(const KP& kp, T* Q, int NV, hipStream_t s) {
  TimedScope ts_("exchange_ghost", s);
  for (int d = 0; d < 3; ++d) {
    if (kp.bc[2 * d] != F3D_BC_PERIODIC || kp.bc[2 * d + 1] != F3D_BC_PERIODIC) continue;
    Range3 r = slab_range(kp, d);
    if (d == 0) hipLaunchKernelGGL((k_exchange<T, 0>), slab_grid_x(r), kBlockX, 0, s, kp, r, Q, NV);
    if (d == 1) hipLaunchKernelGGL((k_exchange<T, 1>), range_grid(r), kBlock, 0, s, kp, r, Q, NV);
    if (d == 2) hipLaunchKernelGGL((k_exchange<T, 2>), range_grid(r), kBlock, 0, s, kp, r, Q, NV);
  }
  return check_launch("exchange_ghost");
}

template <typename T>
int launch_fill(const KP& kp, T* Q, double* U, const T* inflow, hipStream_t s) {
  TimedScope ts_("fill_ghost", s);
  if (fill_all()) {
    const long long G2 = 2 * kp.NG;
    const long long nz = (long long)kp.Nxt * kp.Nyt * G2, ny = (long long)kp.Nxt * G2 * kp.Nzp,
                    nx = G2 * kp.Nyp * kp.Nzp;
    const long long nb = (nz + ny + nx + 255) / 256;
    hipLaunchKernelGGL(k_fill_all<T>, dim3((unsigned int)nb), dim3(256), 0, s, kp, Q, U, inflow, nz, ny, nx);
    return check_launch("fill_ghost");
  }
  for (int d = 0; d < 3; ++d) {
    Range3 r = slab_range(kp, d);
    if (d == 0) hipLaunchKernelGGL((k_fill<T, 0>), slab_grid_x(r), kBlockX, 0, s, kp, r, Q, U, inflow);
    if (d == 1) hipLaunchKernelGGL((k_fill<T, 1>), range_grid(r), kBlock, 0, s, kp, r, Q, U, inflow);
    if (d == 2) hipLaunchKernelGGL((k_fill<T, 2>), range_grid(r), kBlock, 0, s, kp, r, Q, U, inflow);
  }
  return check_launch("fill_ghost");
}

#define F3D_INST(T)                                                                                              \
  template int launch_shock_sensor<T>(const KP&, T*, const T*, hipStream_t);                                    \
  template int launch_scheme_mask<T>(const KP&, uint8_t*, uint8_t*, uint8_t*, const T*, hipStream_t);           \
  template int launch_flux_split<T>(const KP&, const T*, T*, T*, const T*, const T*, const T*, const T*,        \
                                    hipStream_t);                                                               \
  template int launch_advect<T>(const KP&, int, T*, const T*, const T*, const T*, const T*, const T*, const T*, \
                                const T*, const T*, hipStream_t, int);                                          \
  template int launch_viscous<T>(const KP&, T*, T*, T*, const T*, const MetricPtrs<T>&, const double*,          \
                                 hipStream_t);                                                                  \
  template int launch_div<T>(const KP&, double*, const T*, const T*, const T*, const T*, const T*, const T*,    \
                             double, const double*, hipStream_t, const double*);                                \
  template int launch_pre_sensor<T>(const KP&, int, double*, const T*, double, hipStream_t);                     \
  template int launch_filter_block<T>(const KP&, double*, double*, const T*, double*, int, double, double,       \
                                      hipStream_t);                                                             \
  template int launch_compute_dt<T>(const KP&, double*, const T*, const double*, const T*, const T*, const T*,  \
                                    const MetricPtrs<T>*, hipStream_t);                                         \
  template int launch_c2prim<T>(const KP&, const double*, T*, hipStream_t);                                     \
  template int launch_prim2c<T>(const KP&, double*, const T*, hipStream_t);                                     \
  template int launch_exchange<T>(const KP&, T*, int, hipStream_t);                                             \
  template int launch_fill<T>(const KP&, T*, double*, const T*, hipStream_t);
F3D_INST(double)
F3D_INST(float)

}  // namespace f3d
