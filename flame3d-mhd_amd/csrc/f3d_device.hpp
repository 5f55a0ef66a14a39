// Device-side building blocks of the Flame3D finite-difference RHS for gfx950 (CDNA4, wave64).
//
// Indexing is 0-based here; cell (i,j,k) of variable n sits at i + Nxt*(j + Nyt*k) + n*ncell, the
// column-major layout of the reference's CuArrays (solver.jl:294-308).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/f3d.h"

// Point-wise math is host+device so that the local derivative pieces can be checked on the CPU.
#define F3D_HD __host__ __device__ __forceinline__

namespace f3d {

// Kernel-side copy of f3d_params (passed by value as a kernel argument).
struct KP {
  int Nxp, Nyp, Nzp, NG;
  int Nxt, Nyt, Nzt, viscous;
  int bc[6];
  int fvm;       // F3D_FDM / F3D_FVM_EIGEN / F3D_FVM_CONSER
  int split, character, les, lts;  // F3D_SPLIT_*, characteristic flag, F3D_LES_*, local time stepping
  long long nc;  // cells per variable, Nxt*Nyt*Nzt
  double gamma, Rg, Cp, Cs, Ts, Pr;
  double phi1, phi2, phi3, eps7, eps5, Twall;
  double lts_cfl;  // LTS_CFL of compute_dt
  double lin[7];
  double qinf[6];
  // differenced UP7 weights: F(i) - F(i+1) = sum_r da[r+4] Fp[i+r] (r=-4..3) + sum_r db[r+3] Fm[i+r] (r=-3..4)
  double da[8], db[8];
  float linf[7], daf[8], dbf[8];  // float copies (uniform values stay in scalar registers)
  // Steger-Warming constants, evaluated once on the host: gamma*Rg, gamma-1, (3-gamma)/(2(gamma-1)),
  // 1/(2 gamma); [0] double, [1] float (stored as double-width slots for alignment)
  double gRg, gm1, w2c, inv2g;
  float gRgf, gm1f, w2cf, inv2gf, gamf;
};

// Scheme coefficients in the bulk type, read straight from the kernel argument.
template <typename T>
struct Coef;
template <>
struct Coef<double> {
  static F3D_HD double lin(const KP& kp, int q) { return kp.lin[q]; }
  static F3D_HD double da(const KP& kp, int q) { return kp.da[q]; }
  static F3D_HD double db(const KP& kp, int q) { return kp.db[q]; }
  static F3D_HD double gRg(const KP& kp) { return kp.gRg; }
  static F3D_HD double gm1(const KP& kp) { return kp.gm1; }
  static F3D_HD double w2c(const KP& kp) { return kp.w2c; }
  static F3D_HD double inv2g(const KP& kp) { return kp.inv2g; }
};
template <>
struct Coef<float> {
  static F3D_HD float lin(const KP& kp, int q) { return kp.linf[q]; }
  static F3D_HD float da(const KP& kp, int q) { return kp.daf[q]; }
  static F3D_HD float db(const KP& kp, int q) { return kp.dbf[q]; }
  static F3D_HD float gRg(const KP& kp) { return kp.gRgf; }
  static F3D_HD float gm1(const KP& kp) { return kp.gm1f; }
  static F3D_HD float w2c(const KP& kp) { return kp.w2cf; }
  static F3D_HD float inv2g(const KP& kp) { return kp.inv2gf; }
};

F3D_HD long long cidx(const KP& kp, int i, int j, int k) {
  return (long long)i + (long long)kp.Nxt * ((long long)j + (long long)kp.Nyt * k);
}

// Stride of one cell step along direction d (0:i, 1:j, 2:k).
F3D_HD long long dstride(const KP& kp, int d) {
  return d == 0 ? 1LL : (d == 1 ? (long long)kp.Nxt : (long long)kp.Nxt * kp.Nyt);
}

// NaN-propagating max (Julia Base.max semantics on floats; schemes.jl:724-729 uses it).
template <typename T>
F3D_HD T jmax(T a, T b) {
  return (a != a) ? a : ((b != b) ? b : ((b > a) ? b : a));
}

// NaN-propagating min (Julia Base.min on floats).
template <typename T>
F3D_HD T jmin(T a, T b) {
  return (a != a) ? a : ((b != b) ? b : ((b < a) ? b : a));
}

// Scheme index from the 6-cell sensor maximum (schemes.jl:731, 813, 931, else NND2 at 1000).
template <typename T>
F3D_HD int scheme_of(const KP& kp, T phix) {
  return phix < (T)kp.phi1 ? F3D_SCHEME_UP7
                           : (phix < (T)kp.phi2 ? F3D_SCHEME_WENO7
                                                : (phix < (T)kp.phi3 ? F3D_SCHEME_WENO5 : F3D_SCHEME_NND2));
}

// Jameson-type pressure sensor of one cell (schemes.jl:19-22), contraction off so that the mask
// derived from it is reproducible bit for bit against a non-contracted CPU evaluation.
template <typename T>
F3D_HD T sensor_cell(T px1, T p, T px3, T py1, T py3, T pz1, T pz3) {
#pragma clang fp contract(off)
  T a = fabs(-px1 + 2 * p - px3) / (px1 + 2 * p + px3);
  T b = fabs(-py1 + 2 * p - py3) / (py1 + 2 * p + py3);
  T c = fabs(-pz1 + 2 * p - pz3) / (pz1 + 2 * p + pz3);
  return a + b + c;
}

// ------------------------------------------------------------------ Steger-Warming splitting
// split.jl:2-68 for one cell: Fp, Fm (5 each) from rho,u,v,w,T and the metric row A (J^-1
// scaled) with S = |A|.
template <typename T>
F3D_HD void sw_split(const KP& kp, T rho, T u, T v, T w, T Tt, T A1, T A2, T A3,
                                         T ss, T* __restrict__ fp, T* __restrict__ fm) {
  const T c = sqrt(Coef<T>::gRg(kp) * Tt);
  const T E1 = A1 * u + A2 * v + A3 * w;
  const T cs = c * ss;
  const T E2 = E1 - cs, E3 = E1 + cs;
  const T inv = T(1) / ss;
  const T a1 = A1 * inv, a2 = A2 * inv, a3 = A3 * inv;
  const T E1P = (E1 + fabs(E1)) * T(0.5), E2P = (E2 + fabs(E2)) * T(0.5), E3P = (E3 + fabs(E3)) * T(0.5);
  const T E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
  const T uc1 = u - c * a1, uc2 = u + c * a1;
  const T vc1 = v - c * a2, vc2 = v + c * a2;
  const T wc1 = w - c * a3, wc2 = w + c * a3;
  const T vvc1 = (uc1 * uc1 + vc1 * vc1 + wc1 * wc1) * T(0.5);
  const T vvc2 = (uc2 * uc2 + vc2 * vc2 + wc2 * wc2) * T(0.5);
  const T vv = Coef<T>::gm1(kp) * (u * u + v * v + w * w);
  const T W2 = Coef<T>::w2c(kp) * (c * c);
  const T t1 = rho * Coef<T>::inv2g(kp);
  const T t2 = 2 * Coef<T>::gm1(kp);
  const T p1 = t2 * E1P, m1 = t2 * E1M;
  fp[0] = t1 * (p1 + E2P + E3P);
  fp[1] = t1 * (p1 * u + E2P * uc1 + E3P * uc2);
  fp[2] = t1 * (p1 * v + E2P * vc1 + E3P * vc2);
  fp[3] = t1 * (p1 * w + E2P * wc1 + E3P * wc2);
  fp[4] = t1 * (E1P * vv + E2P * vvc1 + E3P * vvc2 + W2 * (E2P + E3P));
  fm[0] = t1 * (m1 + E2M + E3M);
  fm[1] = t1 * (m1 * u + E2M * uc1 + E3M * uc2);
  fm[2] = t1 * (m1 * v + E2M * vc1 + E3M * vc2);
  fm[3] = t1 * (m1 * w + E2M * wc1 + E3M * wc2);
  fm[4] = t1 * (E1M * vv + E2M * vvc1 + E3M * vvc2 + W2 * (E2M + E3M));
}

// The three other splittings of split.jl (LF 70-108, van Leer 110-183, AUSM 185-255) for one cell from
// rho, u, v, w, p and the metric row A with S = |A|; c = sqrt(gamma p / rho).
template <typename T>
F3D_HD void other_split(const KP& kp, int method, T rho, T u, T v, T w, T pr, T A1, T A2, T A3, T ss,
                        T* __restrict__ fp, T* __restrict__ fm) {
  const T gam = (T)kp.gamma;
  const T c = sqrt(gam * pr / rho);
  T un = A1 * u + A2 * v + A3 * w;
  const T q2 = u * u + v * v + w * w;
  if (method == F3D_SPLIT_LF) {
    const T l0 = fabs(un) + c * ss;
    const T E0 = pr / (gam - 1) + T(0.5) * rho * q2;
    const T f[5] = {rho * un, rho * u * un + A1 * pr, rho * v * un + A2 * pr, rho * w * un + A3 * pr, (E0 + pr) * un};
    const T U[5] = {rho, rho * u, rho * v, rho * w, E0};
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      fp[q] = T(0.5) * (f[q] + l0 * U[q]);
      fm[q] = T(0.5) * (f[q] - l0 * U[q]);
    }
    return;
  }
  const T M = un / (c * ss);
  const T E0 = pr / (gam - 1) + T(0.5) * rho * q2;
  const T h = gam * pr / (rho * (gam - 1)) + T(0.5) * q2;
  if (M <= T(-1) || M >= T(1)) {
    const T f[5] = {rho * un, rho * u * un + A1 * pr, rho * v * un + A2 * pr, rho * w * un + A3 * pr,
                    method == F3D_SPLIT_VL ? (E0 + pr) * un : rho * h * un};
    const bool right = M >= T(1);
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      fp[q] = right ? f[q] : T(0);
      fm[q] = right ? T(0) : f[q];
    }
    return;
  }
  const T f1p = T(0.25) * ((M + 1) * (M + 1)) * rho * c * ss;
  const T f1m = -T(0.25) * ((M - 1) * (M - 1)) * rho * c * ss;
  if (method == F3D_SPLIT_VL) {
    un = un / ss;
    const T t1 = (-un + 2 * c) / (gam * ss), t2 = (-un - 2 * c) / (gam * ss);
    const T t3 = T(1) / (2 * (gam * gam - 1));
    const T t4 = T(0.5) * (q2 - un * un);
    const T ep = (gam - 1) * un + 2 * c, em = (gam - 1) * un - 2 * c;
    fp[0] = f1p;
    fp[1] = f1p * (A1 * t1 + u);
    fp[2] = f1p * (A2 * t1 + v);
    fp[3] = f1p * (A3 * t1 + w);
    fp[4] = f1p * (ep * ep * t3 + t4);
    fm[0] = f1m;
    fm[1] = f1m * (A1 * t2 + u);
    fm[2] = f1m * (A2 * t2 + v);
    fm[3] = f1m * (A3 * t2 + w);
    fm[4] = f1m * (em * em * t3 + t4);
    return;
  }
  const T Pp = pr * (T(0.25) * ((M + 1) * (M + 1)) * (2 - M));
  const T Pm = pr * (T(0.25) * ((M - 1) * (M - 1)) * (2 + M));
  fp[0] = f1p;
  fp[1] = f1p * u + A1 * Pp;
  fp[2] = f1p * v + A2 * Pp;
  fp[3] = f1p * w + A3 * Pp;
  fp[4] = f1p * h;
  fm[0] = f1m;
  fm[1] = f1m * u + A1 * Pm;
  fm[2] = f1m * v + A2 * Pm;
  fm[3] = f1m * w + A3 * Pm;
  fm[4] = f1m * h;
}

// Split of one cell by the method SPLIT (F3D_SPLIT_*): Steger-Warming reads T (Q index 5), the others p
// (Q index 4); `qe` is that value.
template <typename T, int SPLIT>
F3D_HD void split_cell(const KP& kp, T rho, T u, T v, T w, T qe, T A1, T A2, T A3, T ss, T* __restrict__ fp,
                       T* __restrict__ fm) {
  if (SPLIT == F3D_SPLIT_SW)
    sw_split<T>(kp, rho, u, v, w, qe, A1, A2, A3, ss, fp, fm);
  else
    other_split<T>(kp, SPLIT, rho, u, v, w, qe, A1, A2, A3, ss, fp, fm);
}
template <int SPLIT>
F3D_HD constexpr int split_qvar() {
  return SPLIT == F3D_SPLIT_SW ? 5 : 4;
}

// ------------------------------------------------------------------ WENO kernels
// Nonlinear WENO7 weights (schemes.jl:871-879).  Double: the literal form of the reference, whose
// weights are Float64.  Float: the same normalised weights formed from ratios to the smallest
// |eps + Is*s0| so that 1/(1e-20)^2 cannot overflow float.
template <typename T>
struct W7;
template <>
struct W7<double> {
  static F3D_HD double mix(double eps, double s0, double I1, double I2, double I3,
                                               double I4, double q1, double q2, double q3, double q4) {
    const double b1 = eps + I1 * s0, b2 = eps + I2 * s0, b3 = eps + I3 * s0, b4 = eps + I4 * s0;
    const double a1 = 1.0 / (b1 * b1), a2 = 12.0 / (b2 * b2), a3 = 18.0 / (b3 * b3), a4 = 4.0 / (b4 * b4);
    const double inv = 1.0 / (a1 + a2 + a3 + a4);
    return inv * (a1 * q1 + a2 * q2 + a3 * q3 + a4 * q4);
  }
};
// Ratio bm/|b| of the range-safe float weights, bm the smallest |b| of the stencil (so the ratio lies
// in [0, 1] even where rounding makes a smoothness indicator slightly negative); a vanishing |b|, whose
// weight is unbounded, gets 1 and every other stencil 0.
F3D_HD float wratio(float bm, float b) {
  const float ab = fabsf(b);
  return ab > 0.f ? bm / ab : 1.f;
}

template <>
struct W7<float> {
  static F3D_HD float mix(float eps, float s0, float I1, float I2, float I3, float I4,
                                              float q1, float q2, float q3, float q4) {
    const float b1 = eps + I1 * s0, b2 = eps + I2 * s0, b3 = eps + I3 * s0, b4 = eps + I4 * s0;
    const float bm = fminf(fminf(fabsf(b1), fabsf(b2)), fminf(fabsf(b3), fabsf(b4)));
    const float r1 = wratio(bm, b1), r2 = wratio(bm, b2), r3 = wratio(bm, b3), r4 = wratio(bm, b4);
    const float a1 = r1 * r1, a2 = 12.f * (r2 * r2), a3 = 18.f * (r3 * r3), a4 = 4.f * (r4 * r4);
    return (a1 * q1 + a2 * q2 + a3 * q3 + a4 * q4) / (a1 + a2 + a3 + a4);
  }
};

// WENO7 smoothness indicators (schemes.jl:848-869) in the float form: the same quadratic forms
// regrouped in the first differences a_k = V_{k+1} - V_k (each indicator is invariant to adding a
// constant to V; identical polynomials in rational arithmetic).  In float the V form loses smooth
// stencils' indicators to the cancellation of O(17246 V^2) terms and the weights follow rounding noise.
template <typename T>
F3D_HD void weno7_is_diff(const T* V, T& I1, T& I2, T& I3, T& I4) {
  const T a1 = V[1] - V[0], a2 = V[2] - V[1], a3 = V[3] - V[2], a4 = V[4] - V[3], a5 = V[5] - V[4],
          a6 = V[6] - V[5];
  I1 = a1 * (547 * a1 - 2788 * a2 + 1854 * a3) + a2 * (3708 * a2 - 5188 * a3) + a3 * (2107 * a3);
  I2 = a2 * (267 * a2 - 1108 * a3 + 494 * a4) + a3 * (1468 * a3 - 1428 * a4) + a4 * (547 * a4);
  I3 = a3 * (547 * a3 - 1428 * a4 + 494 * a5) + a4 * (1468 * a4 - 1108 * a5) + a5 * (267 * a5);
  I4 = a4 * (2107 * a4 - 5188 * a5 + 1854 * a6) + a5 * (3708 * a5 - 2788 * a6) + a6 * (547 * a6);
}

// WENO7 of a 7-point stencil (schemes.jl:847-879), unscaled (the caller multiplies by 1/12).
// FP64: the reference's indicators in V; float: weno7_is_diff.
template <typename T>
F3D_HD T weno7(const T* V, T s0, T eps) {
  const T V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4], V6 = V[5], V7 = V[6];
  const T q1 = -3 * V1 + 13 * V2 - 23 * V3 + 25 * V4;
  const T q2 = V2 - 5 * V3 + 13 * V4 + 3 * V5;
  const T q3 = -V3 + 7 * V4 + 7 * V5 - V6;
  const T q4 = 3 * V4 + 13 * V5 - 5 * V6 + V7;
  if (sizeof(T) == 4) {
    T J1, J2, J3, J4;
    weno7_is_diff(V, J1, J2, J3, J4);
    return W7<T>::mix(eps, s0, J1, J2, J3, J4, q1, q2, q3, q4);
  }
  const T I1 = V1 * (547 * V1 - 3882 * V2 + 4642 * V3 - 1854 * V4) + V2 * (7043 * V2 - 17246 * V3 + 7042 * V4) +
               V3 * (11003 * V3 - 9402 * V4) + V4 * (2107 * V4);
  const T I2 = V2 * (267 * V2 - 1642 * V3 + 1602 * V4 - 494 * V5) + V3 * (2843 * V3 - 5966 * V4 + 1922 * V5) +
               V4 * (3443 * V4 - 2522 * V5) + V5 * (547 * V5);
  const T I3 = V3 * (547 * V3 - 2522 * V4 + 1922 * V5 - 494 * V6) + V4 * (3443 * V4 - 5966 * V5 + 1602 * V6) +
               V5 * (2843 * V5 - 1642 * V6) + V6 * (267 * V6);
  const T I4 = V4 * (2107 * V4 - 9402 * V5 + 7042 * V6 - 1854 * V7) + V5 * (11003 * V5 - 17246 * V6 + 4642 * V7) +
               V6 * (7043 * V6 - 3882 * V7) + V7 * (547 * V7);
  return W7<T>::mix(eps, s0, I1, I2, I3, I4, q1, q2, q3, q4);
}

// WENO5 of a 5-point stencil (schemes.jl:958-972), unscaled (the caller multiplies by 1/6).
template <typename T>
F3D_HD T weno5(const T* V, T s0, T eps) {
  const T V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4];
  const T a1 = V1 - 2 * V2 + V3, c1 = V1 - 4 * V2 + 3 * V3;
  const T a2 = V2 - 2 * V3 + V4, c2 = V2 - V4;
  const T a3 = V3 - 2 * V4 + V5, c3 = 3 * V3 - 4 * V4 + V5;
  const T b1 = eps + (13 * (a1 * a1) + 3 * (c1 * c1)) * s0;
  const T b2 = eps + (13 * (a2 * a2) + 3 * (c2 * c2)) * s0;
  const T b3 = eps + (13 * (a3 * a3) + 3 * (c3 * c3)) * s0;
  const T w1 = T(1) / (b1 * b1), w2 = T(6) / (b2 * b2), w3 = T(3) / (b3 * b3);
  const T inv = T(1) / (w1 + w2 + w3);
  const T v1 = 2 * V1 - 7 * V2 + 11 * V3;
  const T v2 = -V2 + 5 * V3 + 2 * V4;
  const T v3 = 2 * V3 + 5 * V4 - V5;
  return inv * (w1 * v1 + w2 * v2 + w3 * v3);
}

template <typename T>
F3D_HD T minmod(T a, T b) {
  return (a * b > 0) ? ((fabs(a) > fabs(b)) ? b : a) : T(0);
}

// ------------------------------------------------------------------ one face
// Flux through the face between stencil cells c-1 and c (advect_xc, schemes.jl:731-1009).
//   FP(s,m), FM(s,m): split fluxes of variable m at stencil offset s in [-4, 3] from cell c.
//   qL/qR: (u,v,w,T) of cells c-1 / c; AL/AR: metric row; SL/SR: |A|.
// CHAR = false: the component-wise advect_* (schemes.jl:86-698): the same schemes on each split-flux
// component, no characteristic projection (qL, qR, AL, AR unused).
template <typename T, bool CHAR = true, class FPL, class FML>
F3D_HD void face_flux(const KP& kp, int sch, const FPL& FP, const FML& FM,
                                          const T* qL, const T* qR, const T* AL, const T* AR, T SL,
                                          T SR, T* __restrict__ out) {
  if (sch == F3D_SCHEME_UP7) {  // schemes.jl:731-754
    const T c0 = Coef<T>::lin(kp, 0), c1 = Coef<T>::lin(kp, 1), c2 = Coef<T>::lin(kp, 2), c3 = Coef<T>::lin(kp, 3),
            c4 = Coef<T>::lin(kp, 4), c5 = Coef<T>::lin(kp, 5), c6 = Coef<T>::lin(kp, 6);
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      const T fp = c0 * FP(-4, n) + c1 * FP(-3, n) + c2 * FP(-2, n) + c3 * FP(-1, n) + c4 * FP(0, n) +
                   c5 * FP(1, n) + c6 * FP(2, n);
      const T fm = c0 * FM(3, n) + c1 * FM(2, n) + c2 * FM(1, n) + c3 * FM(0, n) + c4 * FM(-1, n) +
                   c5 * FM(-2, n) + c6 * FM(-3, n);
      out[n] = fp + fm;
    }
    return;
  }
  if (sch == F3D_SCHEME_NND2) {  // schemes.jl:1000-1009
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      const T fp = FP(-1, n) + T(0.5) * minmod(FP(0, n) - FP(-1, n), FP(-1, n) - FP(-2, n));
      const T fm = FM(0, n) - T(0.5) * minmod(FM(1, n) - FM(0, n), FM(0, n) - FM(-1, n));
      out[n] = fp + fm;
    }
    return;
  }
  if (!CHAR) {  // component-wise WENO7 / WENO5  schemes.jl:144-276
    const T s0 = T(2) / (SL + SR);
    if (sch == F3D_SCHEME_WENO7) {
      const T eps = (T)kp.eps7;
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        T vp[7], vm[7];
#pragma unroll
        for (int s = 0; s < 7; ++s) {
          vp[s] = FP(s - 4, n);
          vm[s] = FM(3 - s, n);
        }
        out[n] = (weno7(vp, s0, eps) + weno7(vm, s0, eps)) * (T(1) / T(12));
      }
    } else {
      const T eps = (T)kp.eps5;
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        T vp[5], vm[5];
#pragma unroll
        for (int s = 0; s < 5; ++s) {
          vp[s] = FP(s - 3, n);
          vm[s] = FM(2 - s, n);
        }
        out[n] = (weno5(vp, s0, eps) + weno5(vm, s0, eps)) * (T(1) / T(6));
      }
    }
    return;
  }
  // characteristic frame  schemes.jl:756-811
  const T s0 = T(2) / (SL + SR);
  const T u = T(0.5) * (qR[0] + qL[0]), v = T(0.5) * (qR[1] + qL[1]), w = T(0.5) * (qR[2] + qL[2]);
  const T Tt = T(0.5) * (qR[3] + qL[3]);
  const T nx = (AR[0] + AL[0]) * T(0.5) * s0, ny = (AR[1] + AL[1]) * T(0.5) * s0, nz = (AR[2] + AL[2]) * T(0.5) * s0;
  T lx, ly, lz;
  if (fabs(nz) <= fabs(ny)) {
    const T ss = T(1) / sqrt(nx * nx + ny * ny);
    lx = -ny * ss; ly = nx * ss; lz = 0;
  } else {
    const T ss = T(1) / sqrt(nx * nx + nz * nz);
    lx = -nz * ss; ly = 0; lz = nx * ss;
  }
  const T mx = ny * lz - nz * ly, my = nz * lx - nx * lz, mz = nx * ly - ny * lx;
  const T qn = u * nx + v * ny + w * nz, ql = u * lx + v * ly + w * lz, qm = u * mx + v * my + w * mz;
  const T q2 = T(0.5) * (u * u + v * v + w * w);
  const T c = sqrt((T)kp.gamma * (T)kp.Rg * Tt);
  const T ic2 = T(1) / (2 * c);
  const T K = ((T)kp.gamma - 1) / (c * c);
  const T K2 = K * T(0.5);
  const T H = T(1) / K + q2;
  // left eigenvectors, rows n = 0..4
  const T L[5][5] = {{K2 * q2 + qn * ic2, -(K2 * u + nx * ic2), -(K2 * v + ny * ic2), -(K2 * w + nz * ic2), K2},
                     {1 - K * q2, K * u, K * v, K * w, -K},
                     {K2 * q2 - qn * ic2, -(K2 * u - nx * ic2), -(K2 * v - ny * ic2), -(K2 * w - nz * ic2), K2},
                     {-ql, lx, ly, lz, T(0)},
                     {-qm, mx, my, mz, T(0)}};
  T flux[5];
  if (sch == F3D_SCHEME_WENO7) {  // schemes.jl:813-930
    const T eps = (T)kp.eps7;
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      T vp[7], vm[7];
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        vp[s] = FP(s - 4, 0) * L[n][0] + FP(s - 4, 1) * L[n][1] + FP(s - 4, 2) * L[n][2] +
                FP(s - 4, 3) * L[n][3] + FP(s - 4, 4) * L[n][4];
        vm[s] = FM(3 - s, 0) * L[n][0] + FM(3 - s, 1) * L[n][1] + FM(3 - s, 2) * L[n][2] +
                FM(3 - s, 3) * L[n][3] + FM(3 - s, 4) * L[n][4];
      }
      flux[n] = (weno7(vp, s0, eps) + weno7(vm, s0, eps)) * (T(1) / T(12));
    }
  } else {  // WENO5  schemes.jl:931-999
    const T eps = (T)kp.eps5;
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      T vp[5], vm[5];
#pragma unroll
      for (int s = 0; s < 5; ++s) {
        vp[s] = FP(s - 3, 0) * L[n][0] + FP(s - 3, 1) * L[n][1] + FP(s - 3, 2) * L[n][2] +
                FP(s - 3, 3) * L[n][3] + FP(s - 3, 4) * L[n][4];
        vm[s] = FM(2 - s, 0) * L[n][0] + FM(2 - s, 1) * L[n][1] + FM(2 - s, 2) * L[n][2] +
                FM(2 - s, 3) * L[n][3] + FM(2 - s, 4) * L[n][4];
      }
      flux[n] = (weno5(vp, s0, eps) + weno5(vm, s0, eps)) * (T(1) / T(6));
    }
  }
  // back-projection with the right eigenvectors (columns of R, schemes.jl:807-811, 926-930)
  out[0] = flux[0] + flux[1] + flux[2];
  out[1] = flux[0] * (u - c * nx) + flux[1] * u + flux[2] * (u + c * nx) + flux[3] * lx + flux[4] * mx;
  out[2] = flux[0] * (v - c * ny) + flux[1] * v + flux[2] * (v + c * ny) + flux[3] * ly + flux[4] * my;
  out[3] = flux[0] * (w - c * nz) + flux[1] * w + flux[2] * (w + c * nz) + flux[3] * lz + flux[4] * mz;
  out[4] = flux[0] * (H - qn * c) + flux[1] * q2 + flux[2] * (H + qn * c) + flux[3] * ql + flux[4] * qm;
}

// Characteristic frame of one face (schemes.jl:756-811): arithmetic means of (u,v,w,T) of the two cells,
// the averaged metric direction n, the tangents l and m, and the quantities the left and right eigenvectors
// are formed from; face_flux's arithmetic, for kernels that form the rows of L one at a time.
template <typename T>
struct CharFrame {
  T u, v, w, nx, ny, nz, lx, ly, lz, mx, my, mz, qn, ql, qm, q2, c, ic2, K, K2;
  F3D_HD CharFrame(const KP& kp, const T* qL, const T* qR, const T* AL, const T* AR, T SL, T SR) {
    const T s0 = T(2) / (SL + SR);
    u = T(0.5) * (qR[0] + qL[0]);
    v = T(0.5) * (qR[1] + qL[1]);
    w = T(0.5) * (qR[2] + qL[2]);
    const T Tt = T(0.5) * (qR[3] + qL[3]);
    nx = (AR[0] + AL[0]) * T(0.5) * s0;
    ny = (AR[1] + AL[1]) * T(0.5) * s0;
    nz = (AR[2] + AL[2]) * T(0.5) * s0;
    if (fabs(nz) <= fabs(ny)) {
      const T ss = T(1) / sqrt(nx * nx + ny * ny);
      lx = -ny * ss; ly = nx * ss; lz = 0;
    } else {
      const T ss = T(1) / sqrt(nx * nx + nz * nz);
      lx = -nz * ss; ly = 0; lz = nx * ss;
    }
    mx = ny * lz - nz * ly; my = nz * lx - nx * lz; mz = nx * ly - ny * lx;
    qn = u * nx + v * ny + w * nz; ql = u * lx + v * ly + w * lz; qm = u * mx + v * my + w * mz;
    q2 = T(0.5) * (u * u + v * v + w * w);
    c = sqrt((T)kp.gamma * (T)kp.Rg * Tt);
    ic2 = T(1) / (2 * c);
    K = ((T)kp.gamma - 1) / (c * c);
    K2 = K * T(0.5);
  }
  // row n of the left eigenvector matrix (schemes.jl:801-806)
  F3D_HD void lrow(int n, T* L) const {
    if (n == 0) {
      L[0] = K2 * q2 + qn * ic2; L[1] = -(K2 * u + nx * ic2); L[2] = -(K2 * v + ny * ic2); L[3] = -(K2 * w + nz * ic2); L[4] = K2;
    } else if (n == 1) {
      L[0] = 1 - K * q2; L[1] = K * u; L[2] = K * v; L[3] = K * w; L[4] = -K;
    } else if (n == 2) {
      L[0] = K2 * q2 - qn * ic2; L[1] = -(K2 * u - nx * ic2); L[2] = -(K2 * v - ny * ic2); L[3] = -(K2 * w - nz * ic2); L[4] = K2;
    } else if (n == 3) {
      L[0] = -ql; L[1] = lx; L[2] = ly; L[3] = lz; L[4] = T(0);
    } else {
      L[0] = -qm; L[1] = mx; L[2] = my; L[3] = mz; L[4] = T(0);
    }
  }
  // back-projection with the right eigenvectors (schemes.jl:807-811, 926-930)
  F3D_HD void project_back(const T* flux, T* out) const {
    const T H = T(1) / K + q2;
    out[0] = flux[0] + flux[1] + flux[2];
    out[1] = flux[0] * (u - c * nx) + flux[1] * u + flux[2] * (u + c * nx) + flux[3] * lx + flux[4] * mx;
    out[2] = flux[0] * (v - c * ny) + flux[1] * v + flux[2] * (v + c * ny) + flux[3] * ly + flux[4] * my;
    out[3] = flux[0] * (w - c * nz) + flux[1] * w + flux[2] * (w + c * nz) + flux[3] * lz + flux[4] * mz;
    out[4] = flux[0] * (H - qn * c) + flux[1] * q2 + flux[2] * (H + qn * c) + flux[3] * ql + flux[4] * qm;
  }
};

// ------------------------------------------------------------------ viscous pieces
// Eddy viscosity of the LES models (viscous.jl:66-106) from the Cartesian velocity gradient
// (dx[q], dy[q], dz[q]: d/dx, d/dy, d/dz of u, v, w for q = 0, 1, 2), added to mu and lam
// (lam_t = Cp mu_t / 0.9), mu_t capped at twice the laminar mu, filter width J^(-1/3).
//   Smagorinsky: mu_t = rho (0.1 Delta)^2 sqrt(2 S_ij S_ij).
//   WALE: mu_t = rho (0.325 Delta)^2 Sd^3 / (S^5 + Sd^(5/2)) with Sd from the traceless symmetric part
//   of the squared gradient; 0 where S = Sd = 0 (the reference's 0/0).  The float form scales the
//   gradient by its largest component first (S^5 leaves float's range beyond ~5e7 1/s).
template <typename T, int LES>
F3D_HD void les_eddy(const KP& kp, T rho, T Jac, const T* dx, const T* dy, const T* dz, T& mu, T& lam) {
  if constexpr (LES != F3D_LES_NONE) {
    const T half = T(0.5);
    T mut;
    if constexpr (LES == F3D_LES_SMAG) {
      const T s12 = half * (dy[0] + dx[1]), s13 = half * (dz[0] + dx[2]), s23 = half * (dz[1] + dy[2]);
      const T SS = dx[0] * dx[0] + dy[1] * dy[1] + dz[2] * dz[2] + 2 * (s12 * s12 + s13 * s13 + s23 * s23);
      const T dl = T(0.1) / cbrt(Jac);
      mut = jmin<T>(rho * (dl * dl) * sqrt(2 * SS), 2 * mu);
    } else {
      T sc = T(1);
      if (sizeof(T) == 4) {
        T gm = T(0);
#pragma unroll
        for (int q = 0; q < 3; ++q) gm = jmax<T>(gm, jmax<T>(fabs(dx[q]), jmax<T>(fabs(dy[q]), fabs(dz[q]))));
        sc = gm > T(0) ? gm : T(1);
      }
      // sc = 1 in double: the divisions below are exact there
      const T ux = dx[0] / sc, uy = dy[0] / sc, uz = dz[0] / sc;
      const T vx = dx[1] / sc, vy = dy[1] / sc, vz = dz[1] / sc;
      const T wx = dx[2] / sc, wy = dy[2] / sc, wz = dz[2] / sc;
      const T t12 = half * (uy + vx), t13 = half * (uz + wx), t23 = half * (vz + wy);
      const T S = sqrt(ux * ux + vy * vy + wz * wz + 2 * (t12 * t12 + t13 * t13 + t23 * t23));
      T Sd11 = ux * ux + uy * vx + uz * wx;
      T Sd22 = vx * uy + vy * vy + vz * wy;
      T Sd33 = wx * uz + wy * vz + wz * wz;
      const T tr = (Sd11 + Sd22 + Sd33) / T(3);
      Sd11 -= tr;
      Sd22 -= tr;
      Sd33 -= tr;
      const T Sd12 = half * (ux * vx + vx * vy + wx * vz + uy * ux + vy * uy + wy * uz);
      const T Sd13 = half * (ux * wx + vx * wy + wx * wz + uz * ux + vz * uy + wz * uz);
      const T Sd23 = half * (uy * wx + vy * wy + wy * wz + uz * vx + vz * vy + wz * vz);
      const T Sd =
          sqrt(Sd11 * Sd11 + Sd22 * Sd22 + Sd33 * Sd33 + 2 * (Sd12 * Sd12 + Sd13 * Sd13 + Sd23 * Sd23));
      const T den = S * S * S * S * S + Sd * Sd * sqrt(Sd);
      const T D = den > T(0) ? sc * (Sd * Sd * Sd / den) : T(0);
      const T dl = T(0.325) / cbrt(Jac);
      mut = jmin<T>(rho * (dl * dl) * D, 2 * mu);
    }
    mu += mut;
    lam += (T)kp.Cp * mut / T(0.9);
  }
}

// Viscous fluxes of one cell in the three directions (viscous.jl:20-65, 108-133), with the eddy
// viscosity of LES (F3D_LES_*; rho is read only then).
//   g[v][d]: 4th-order central derivative of (u,v,w,T)[v] along xi/eta/zeta (not yet divided).
//   m[9]: metric row-major (xi_x, xi_y, xi_z, eta_x, ...); out[d][4].
template <typename T, int LES = F3D_LES_NONE>
F3D_HD void viscous_cell(const KP& kp, T u, T v, T w, T Tt, const T g[4][3],
                                             const T m[9], T Jac, T out[3][4], T rho = T(0)) {
  T mu = (T)kp.Cs * Tt * sqrt(Tt) / (Tt + (T)kp.Ts);
  T lam = (T)kp.Cp * mu / (T)kp.Pr;
  T dx[4], dy[4], dz[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    dx[q] = (g[q][0] * m[0] + g[q][1] * m[3] + g[q][2] * m[6]) * Jac;
    dy[q] = (g[q][0] * m[1] + g[q][1] * m[4] + g[q][2] * m[7]) * Jac;
    dz[q] = (g[q][0] * m[2] + g[q][1] * m[5] + g[q][2] * m[8]) * Jac;
  }
  les_eddy<T, LES>(kp, rho, Jac, dx, dy, dz, mu, lam);
  const T c23 = T(2) / T(3);
  const T dv = dx[0] + dy[1] + dz[2];
  const T t11 = mu * (2 * dx[0] - c23 * dv);
  const T t12 = mu * (dy[0] + dx[1]);
  const T t13 = mu * (dz[0] + dx[2]);
  const T t22 = mu * (2 * dy[1] - c23 * dv);
  const T t23 = mu * (dy[2] + dz[1]);
  const T t33 = mu * (2 * dz[2] - c23 * dv);
  const T E1 = u * t11 + v * t12 + w * t13 + lam * dx[3];
  const T E2 = u * t12 + v * t22 + w * t23 + lam * dy[3];
  const T E3 = u * t13 + v * t23 + w * t33 + lam * dz[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const T a = m[3 * d], b = m[3 * d + 1], e = m[3 * d + 2];
    out[d][0] = a * t11 + b * t12 + e * t13;
    out[d][1] = a * t12 + b * t22 + e * t23;
    out[d][2] = a * t13 + b * t23 + e * t33;
    out[d][3] = a * E1 + b * E2 + e * E3;
  }
}

}  // namespace f3d
