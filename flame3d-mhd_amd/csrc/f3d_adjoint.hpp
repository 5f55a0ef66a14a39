// Local reverse-mode pieces (vector-Jacobian products) of the FP64 RHS, host+device.
//
// Each function takes the primal inputs of one forward building block of f3d_device.hpp and the
// adjoint of its outputs, and returns (or accumulates) the adjoint of its inputs.  The scheme
// mask, the minmod selections and the tangent-vector branch are frozen at the primal point; |x|
// uses the subgradient sign(0) = 0, so d(E+|E|)/2 at E = 0 is 1/2.
#pragma once
#include "f3d_device.hpp"

namespace f3d {

// ------------------------------------------------------------------ scalar forward-mode dual
// Used only for the 4-input Jacobian of the characteristic frame, one tangent at a time.
struct Dv {
  double v, d;
  F3D_HD Dv() : v(0), d(0) {}
  F3D_HD Dv(double x) : v(x), d(0) {}
  F3D_HD Dv(double x, double y) : v(x), d(y) {}
};
F3D_HD Dv operator+(Dv a, Dv b) { return Dv(a.v + b.v, a.d + b.d); }
F3D_HD Dv operator-(Dv a, Dv b) { return Dv(a.v - b.v, a.d - b.d); }
F3D_HD Dv operator-(Dv a) { return Dv(-a.v, -a.d); }
F3D_HD Dv operator*(Dv a, Dv b) { return Dv(a.v * b.v, a.d * b.v + a.v * b.d); }
F3D_HD Dv operator/(Dv a, Dv b) { return Dv(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
F3D_HD Dv sqrt(Dv a) {
  const double s = ::sqrt(a.v);
  return Dv(s, a.d / (2 * s));
}

// ------------------------------------------------------------------ Steger-Warming splitting
// Reverse of sw_split: adjoints gp, gm of Fp, Fm (5 each) -> adjoints of (rho, u, v, w, T).
F3D_HD void sw_split_vjp(const KP& kp, double rho, double u, double v, double w, double Tt, double A1, double A2,
                         double A3, double ss, const double* gp, const double* gm, double* g) {
  const double gRg = Coef<double>::gRg(kp), gm1 = Coef<double>::gm1(kp), w2c = Coef<double>::w2c(kp);
  const double inv2g = Coef<double>::inv2g(kp);
  const double c = ::sqrt(gRg * Tt);
  const double E1 = A1 * u + A2 * v + A3 * w;
  const double cs = c * ss;
  const double E2 = E1 - cs, E3 = E1 + cs;
  const double inv = 1.0 / ss;
  const double a1 = A1 * inv, a2 = A2 * inv, a3 = A3 * inv;
  const double E1P = (E1 + ::fabs(E1)) * 0.5, E2P = (E2 + ::fabs(E2)) * 0.5, E3P = (E3 + ::fabs(E3)) * 0.5;
  const double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
  const double uc1 = u - c * a1, uc2 = u + c * a1;
  const double vc1 = v - c * a2, vc2 = v + c * a2;
  const double wc1 = w - c * a3, wc2 = w + c * a3;
  const double vvc1 = (uc1 * uc1 + vc1 * vc1 + wc1 * wc1) * 0.5;
  const double vvc2 = (uc2 * uc2 + vc2 * vc2 + wc2 * wc2) * 0.5;
  const double vv = gm1 * (u * u + v * v + w * w);
  const double W2 = w2c * (c * c);
  const double t1 = rho * inv2g;
  const double t2 = 2 * gm1;
  const double p1 = t2 * E1P, m1 = t2 * E1M;
  // brackets of fp / fm (f = t1 * bracket)
  const double Bp[5] = {p1 + E2P + E3P, p1 * u + E2P * uc1 + E3P * uc2, p1 * v + E2P * vc1 + E3P * vc2,
                        p1 * w + E2P * wc1 + E3P * wc2, E1P * vv + E2P * vvc1 + E3P * vvc2 + W2 * (E2P + E3P)};
  const double Bm[5] = {m1 + E2M + E3M, m1 * u + E2M * uc1 + E3M * uc2, m1 * v + E2M * vc1 + E3M * vc2,
                        m1 * w + E2M * wc1 + E3M * wc2, E1M * vv + E2M * vvc1 + E3M * vvc2 + W2 * (E2M + E3M)};
  double gt1 = 0;
  double bp[5], bm[5];
  for (int n = 0; n < 5; ++n) {
    gt1 += gp[n] * Bp[n] + gm[n] * Bm[n];
    bp[n] = t1 * gp[n];
    bm[n] = t1 * gm[n];
  }
  const double g_rho = inv2g * gt1;
  const double g_p1 = bp[0] + bp[1] * u + bp[2] * v + bp[3] * w;
  const double g_m1 = bm[0] + bm[1] * u + bm[2] * v + bm[3] * w;
  double gE1P = t2 * g_p1 + bp[4] * vv;
  const double gE1M = t2 * g_m1 + bm[4] * vv;
  double gE2P = bp[0] + bp[1] * uc1 + bp[2] * vc1 + bp[3] * wc1 + bp[4] * (vvc1 + W2);
  double gE3P = bp[0] + bp[1] * uc2 + bp[2] * vc2 + bp[3] * wc2 + bp[4] * (vvc2 + W2);
  const double gE2M = bm[0] + bm[1] * uc1 + bm[2] * vc1 + bm[3] * wc1 + bm[4] * (vvc1 + W2);
  const double gE3M = bm[0] + bm[1] * uc2 + bm[2] * vc2 + bm[3] * wc2 + bm[4] * (vvc2 + W2);
  double gu = bp[1] * p1 + bm[1] * m1, gv = bp[2] * p1 + bm[2] * m1, gw = bp[3] * p1 + bm[3] * m1;
  double guc1 = bp[1] * E2P + bm[1] * E2M, guc2 = bp[1] * E3P + bm[1] * E3M;
  double gvc1 = bp[2] * E2P + bm[2] * E2M, gvc2 = bp[2] * E3P + bm[2] * E3M;
  double gwc1 = bp[3] * E2P + bm[3] * E2M, gwc2 = bp[3] * E3P + bm[3] * E3M;
  const double gvvc1 = bp[4] * E2P + bm[4] * E2M, gvvc2 = bp[4] * E3P + bm[4] * E3M;
  const double gvv = bp[4] * E1P + bm[4] * E1M;
  const double gW2 = bp[4] * (E2P + E3P) + bm[4] * (E2M + E3M);
  guc1 += uc1 * gvvc1; gvc1 += vc1 * gvvc1; gwc1 += wc1 * gvvc1;
  guc2 += uc2 * gvvc2; gvc2 += vc2 * gvvc2; gwc2 += wc2 * gvvc2;
  // E_M = E - E_P;  E_P = H(E) E with H = 1, 0, 1/2 for E >0, <0, =0
  double gE1 = gE1M, gE2 = gE2M, gE3 = gE3M;
  gE1P -= gE1M;
  gE2P -= gE2M;
  gE3P -= gE3M;
  auto H = [](double e) { return e > 0 ? 1.0 : (e < 0 ? 0.0 : 0.5); };
  gE1 += H(E1) * gE1P;
  gE2 += H(E2) * gE2P;
  gE3 += H(E3) * gE3P;
  gE1 += gE2 + gE3;
  const double gcs = gE3 - gE2;
  gu += guc1 + guc2;
  gv += gvc1 + gvc2;
  gw += gwc1 + gwc2;
  double gc = a1 * (guc2 - guc1) + a2 * (gvc2 - gvc1) + a3 * (gwc2 - gwc1);
  gu += 2 * gm1 * u * gvv;
  gv += 2 * gm1 * v * gvv;
  gw += 2 * gm1 * w * gvv;
  gc += 2 * w2c * c * gW2;
  gc += ss * gcs;
  gu += A1 * gE1;
  gv += A2 * gE1;
  gw += A3 * gE1;
  g[0] = g_rho;
  g[1] = gu;
  g[2] = gv;
  g[3] = gw;
  g[4] = gc * gRg / (2 * c);
}

// ------------------------------------------------------------------ WENO reverse
// Derivative convention for the nonlinear weights: a smoothness indicator that is not resolved
// above the rounding of its own evaluation contributes no weight derivative.  With eps = 1e-20
// the weight of a stencil flat to the last bit is 1/(rounding noise)^2 and its derivative is a
// ratio of two noise terms; such a stencil keeps its weight frozen.  Unresolved means
//   WENO7: Is < kIs7Floor * (the same quadratic form on |V| with |coefficients|)
//   WENO5: Is < kIs5Floor * (13 A^2 + 3 C^2), A, C the |V|-sums of its two differences.
constexpr double kIs7Floor = 1e-12;
constexpr double kIs5Floor = 1e-24;

F3D_HD void weno7_abs_is(const double* V, double* Ia) {
  const double V1 = ::fabs(V[0]), V2 = ::fabs(V[1]), V3 = ::fabs(V[2]), V4 = ::fabs(V[3]), V5 = ::fabs(V[4]),
               V6 = ::fabs(V[5]), V7 = ::fabs(V[6]);
  Ia[0] = V1 * (547 * V1 + 3882 * V2 + 4642 * V3 + 1854 * V4) + V2 * (7043 * V2 + 17246 * V3 + 7042 * V4) +
          V3 * (11003 * V3 + 9402 * V4) + V4 * (2107 * V4);
  Ia[1] = V2 * (267 * V2 + 1642 * V3 + 1602 * V4 + 494 * V5) + V3 * (2843 * V3 + 5966 * V4 + 1922 * V5) +
          V4 * (3443 * V4 + 2522 * V5) + V5 * (547 * V5);
  Ia[2] = V3 * (547 * V3 + 2522 * V4 + 1922 * V5 + 494 * V6) + V4 * (3443 * V4 + 5966 * V5 + 1602 * V6) +
          V5 * (2843 * V5 + 1642 * V6) + V6 * (267 * V6);
  Ia[3] = V4 * (2107 * V4 + 9402 * V5 + 7042 * V6 + 1854 * V7) + V5 * (11003 * V5 + 17246 * V6 + 4642 * V7) +
          V6 * (7043 * V6 + 3882 * V7) + V7 * (547 * V7);
}

F3D_HD void weno5_abs_is(const double* V, double* Ia) {
  for (int m = 0; m < 3; ++m) {
    const double a = ::fabs(V[m]) + 2 * ::fabs(V[m + 1]) + ::fabs(V[m + 2]);
    const double c = m == 0 ? ::fabs(V[0]) + 4 * ::fabs(V[1]) + 3 * ::fabs(V[2])
                            : (m == 1 ? ::fabs(V[1]) + ::fabs(V[3]) : 3 * ::fabs(V[2]) + 4 * ::fabs(V[3]) + ::fabs(V[4]));
    Ia[m] = 13 * a * a + 3 * c * c;
  }
}

// d weno7(V)/dV scaled by gW, accumulated into gV[7] (s0, eps are metric/constant inputs).
F3D_HD void weno7_vjp(const double* V, double s0, double eps, double gW, double* gV) {
  const double V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4], V6 = V[5], V7 = V[6];
  const double q[4] = {-3 * V1 + 13 * V2 - 23 * V3 + 25 * V4, V2 - 5 * V3 + 13 * V4 + 3 * V5,
                       -V3 + 7 * V4 + 7 * V5 - V6, 3 * V4 + 13 * V5 - 5 * V6 + V7};
  const double I[4] = {
      V1 * (547 * V1 - 3882 * V2 + 4642 * V3 - 1854 * V4) + V2 * (7043 * V2 - 17246 * V3 + 7042 * V4) +
          V3 * (11003 * V3 - 9402 * V4) + V4 * (2107 * V4),
      V2 * (267 * V2 - 1642 * V3 + 1602 * V4 - 494 * V5) + V3 * (2843 * V3 - 5966 * V4 + 1922 * V5) +
          V4 * (3443 * V4 - 2522 * V5) + V5 * (547 * V5),
      V3 * (547 * V3 - 2522 * V4 + 1922 * V5 - 494 * V6) + V4 * (3443 * V4 - 5966 * V5 + 1602 * V6) +
          V5 * (2843 * V5 - 1642 * V6) + V6 * (267 * V6),
      V4 * (2107 * V4 - 9402 * V5 + 7042 * V6 - 1854 * V7) + V5 * (11003 * V5 - 17246 * V6 + 4642 * V7) +
          V6 * (7043 * V6 - 3882 * V7) + V7 * (547 * V7)};
  const double dl[4] = {1, 12, 18, 4};
  double a[4], b[4], S = 0, acc = 0;
  for (int m = 0; m < 4; ++m) {
    b[m] = eps + I[m] * s0;
    a[m] = dl[m] / (b[m] * b[m]);
    S += a[m];
    acc += a[m] * q[m];
  }
  const double Wv = acc / S;
  double gq[4], gI[4], Ia[4];
  weno7_abs_is(V, Ia);
  for (int m = 0; m < 4; ++m) {
    gq[m] = gW * a[m] / S;
    const double ga = gW * (q[m] - Wv) / S;
    gI[m] = I[m] < kIs7Floor * Ia[m] ? 0.0 : -2 * a[m] / b[m] * ga * s0;
  }
  // candidate polynomials
  gV[0] += -3 * gq[0];
  gV[1] += 13 * gq[0] + gq[1];
  gV[2] += -23 * gq[0] - 5 * gq[1] - gq[2];
  gV[3] += 25 * gq[0] + 13 * gq[1] + 7 * gq[2] + 3 * gq[3];
  gV[4] += 3 * gq[1] + 7 * gq[2] + 13 * gq[3];
  gV[5] += -gq[2] - 5 * gq[3];
  gV[6] += gq[3];
  // smoothness indicators (gradients of the quadratic forms)
  gV[0] += gI[0] * (1094 * V1 - 3882 * V2 + 4642 * V3 - 1854 * V4);
  gV[1] += gI[0] * (-3882 * V1 + 14086 * V2 - 17246 * V3 + 7042 * V4);
  gV[2] += gI[0] * (4642 * V1 - 17246 * V2 + 22006 * V3 - 9402 * V4);
  gV[3] += gI[0] * (-1854 * V1 + 7042 * V2 - 9402 * V3 + 4214 * V4);
  gV[1] += gI[1] * (534 * V2 - 1642 * V3 + 1602 * V4 - 494 * V5);
  gV[2] += gI[1] * (-1642 * V2 + 5686 * V3 - 5966 * V4 + 1922 * V5);
  gV[3] += gI[1] * (1602 * V2 - 5966 * V3 + 6886 * V4 - 2522 * V5);
  gV[4] += gI[1] * (-494 * V2 + 1922 * V3 - 2522 * V4 + 1094 * V5);
  gV[2] += gI[2] * (1094 * V3 - 2522 * V4 + 1922 * V5 - 494 * V6);
  gV[3] += gI[2] * (-2522 * V3 + 6886 * V4 - 5966 * V5 + 1602 * V6);
  gV[4] += gI[2] * (1922 * V3 - 5966 * V4 + 5686 * V5 - 1642 * V6);
  gV[5] += gI[2] * (-494 * V3 + 1602 * V4 - 1642 * V5 + 534 * V6);
  gV[3] += gI[3] * (4214 * V4 - 9402 * V5 + 7042 * V6 - 1854 * V7);
  gV[4] += gI[3] * (-9402 * V4 + 22006 * V5 - 17246 * V6 + 4642 * V7);
  gV[5] += gI[3] * (7042 * V4 - 17246 * V5 + 14086 * V6 - 3882 * V7);
  gV[6] += gI[3] * (-1854 * V4 + 4642 * V5 - 3882 * V6 + 1094 * V7);
}

// d weno5(V)/dV scaled by gW, accumulated into gV[5].
F3D_HD void weno5_vjp(const double* V, double s0, double eps, double gW, double* gV) {
  const double V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4];
  const double a[3] = {V1 - 2 * V2 + V3, V2 - 2 * V3 + V4, V3 - 2 * V4 + V5};
  const double c[3] = {V1 - 4 * V2 + 3 * V3, V2 - V4, 3 * V3 - 4 * V4 + V5};
  const double v[3] = {2 * V1 - 7 * V2 + 11 * V3, -V2 + 5 * V3 + 2 * V4, 2 * V3 + 5 * V4 - V5};
  const double dl[3] = {1, 6, 3};
  double b[3], wt[3], S = 0, acc = 0;
  for (int m = 0; m < 3; ++m) {
    b[m] = eps + (13 * (a[m] * a[m]) + 3 * (c[m] * c[m])) * s0;
    wt[m] = dl[m] / (b[m] * b[m]);
    S += wt[m];
    acc += wt[m] * v[m];
  }
  const double Wv = acc / S;
  double gv[3], ga[3], gc[3], Ia[3];
  weno5_abs_is(V, Ia);
  for (int m = 0; m < 3; ++m) {
    gv[m] = gW * wt[m] / S;
    const double gwt = gW * (v[m] - Wv) / S;
    const double Is = 13 * (a[m] * a[m]) + 3 * (c[m] * c[m]);
    const double gIs = Is < kIs5Floor * Ia[m] ? 0.0 : -2 * wt[m] / b[m] * gwt * s0;
    ga[m] = 26 * a[m] * gIs;
    gc[m] = 6 * c[m] * gIs;
  }
  gV[0] += 2 * gv[0] + ga[0] + gc[0];
  gV[1] += -7 * gv[0] - gv[1] - 2 * ga[0] + ga[1] - 4 * gc[0] + gc[1];
  gV[2] += 11 * gv[0] + 5 * gv[1] + 2 * gv[2] + ga[0] - 2 * ga[1] + ga[2] + 3 * gc[0] + 3 * gc[2];
  gV[3] += 2 * gv[1] + 5 * gv[2] + ga[1] - 2 * ga[2] - gc[1] - 4 * gc[2];
  gV[4] += -gv[2] + ga[2] + gc[2];
}

// ------------------------------------------------------------------ characteristic frame
// Face normal n, tangents l, m and s0 = 2/(SL+SR) from the two cells' metric rows (schemes.jl:756-790).
F3D_HD void face_geometry(const double* AL, const double* AR, double SL, double SR, double& s0, double* n, double* l,
                          double* m) {
  s0 = 2.0 / (SL + SR);
  n[0] = (AR[0] + AL[0]) * 0.5 * s0;
  n[1] = (AR[1] + AL[1]) * 0.5 * s0;
  n[2] = (AR[2] + AL[2]) * 0.5 * s0;
  if (::fabs(n[2]) <= ::fabs(n[1])) {
    const double ss = 1.0 / ::sqrt(n[0] * n[0] + n[1] * n[1]);
    l[0] = -n[1] * ss; l[1] = n[0] * ss; l[2] = 0;
  } else {
    const double ss = 1.0 / ::sqrt(n[0] * n[0] + n[2] * n[2]);
    l[0] = -n[2] * ss; l[1] = 0; l[2] = n[0] * ss;
  }
  m[0] = n[1] * l[2] - n[2] * l[1];
  m[1] = n[2] * l[0] - n[0] * l[2];
  m[2] = n[0] * l[1] - n[1] * l[0];
}

// Left eigenvectors L (rows) and right eigenvectors Rr (face flux = Rr * characteristic flux) at
// the averaged state (u, v, w, T) (schemes.jl:792-811).
template <typename T>
F3D_HD void char_frame(const KP& kp, T u, T v, T w, T Tt, const double* n, const double* l, const double* m,
                       T L[5][5], T Rr[5][5]) {
  using ::sqrt;  // double: the C function; Dv: f3d::sqrt by argument-dependent lookup
  const T qn = u * n[0] + v * n[1] + w * n[2];
  const T ql = u * l[0] + v * l[1] + w * l[2];
  const T qm = u * m[0] + v * m[1] + w * m[2];
  const T q2 = T(0.5) * (u * u + v * v + w * w);
  const T c = sqrt(T(kp.gamma) * T(kp.Rg) * Tt);
  const T ic2 = T(1) / (T(2) * c);
  const T K = (T(kp.gamma) - T(1)) / (c * c);
  const T K2 = K * T(0.5);
  const T H = T(1) / K + q2;
  L[0][0] = K2 * q2 + qn * ic2; L[0][1] = -(K2 * u + n[0] * ic2); L[0][2] = -(K2 * v + n[1] * ic2);
  L[0][3] = -(K2 * w + n[2] * ic2); L[0][4] = K2;
  L[1][0] = T(1) - K * q2; L[1][1] = K * u; L[1][2] = K * v; L[1][3] = K * w; L[1][4] = -K;
  L[2][0] = K2 * q2 - qn * ic2; L[2][1] = -(K2 * u - n[0] * ic2); L[2][2] = -(K2 * v - n[1] * ic2);
  L[2][3] = -(K2 * w - n[2] * ic2); L[2][4] = K2;
  L[3][0] = -ql; L[3][1] = T(l[0]); L[3][2] = T(l[1]); L[3][3] = T(l[2]); L[3][4] = T(0);
  L[4][0] = -qm; L[4][1] = T(m[0]); L[4][2] = T(m[1]); L[4][3] = T(m[2]); L[4][4] = T(0);
  Rr[0][0] = T(1); Rr[0][1] = T(1); Rr[0][2] = T(1); Rr[0][3] = T(0); Rr[0][4] = T(0);
  Rr[1][0] = u - c * n[0]; Rr[1][1] = u; Rr[1][2] = u + c * n[0]; Rr[1][3] = T(l[0]); Rr[1][4] = T(m[0]);
  Rr[2][0] = v - c * n[1]; Rr[2][1] = v; Rr[2][2] = v + c * n[1]; Rr[2][3] = T(l[1]); Rr[2][4] = T(m[1]);
  Rr[3][0] = w - c * n[2]; Rr[3][1] = w; Rr[3][2] = w + c * n[2]; Rr[3][3] = T(l[2]); Rr[3][4] = T(m[2]);
  Rr[4][0] = H - qn * c; Rr[4][1] = q2; Rr[4][2] = H + qn * c; Rr[4][3] = ql; Rr[4][4] = qm;
}

// ------------------------------------------------------------------ one face, reverse
// Reverse of face_flux: gF = adjoint of the face flux (5).  fp[s][n] / fm[s][n] are the split
// fluxes at stencil offset s-4 (s = 0..7) from the right cell; qL, qR = (u,v,w,T) of the left and
// right cells.  Accumulates into gfp, gfm (8x5) and gqL, gqR (4).
F3D_HD void face_flux_vjp(const KP& kp, int sch, const double fp[8][5], const double fm[8][5], const double* qL,
                          const double* qR, const double* AL, const double* AR, double SL, double SR,
                          const double* gF, double gfp[8][5], double gfm[8][5], double* gqL, double* gqR) {
  if (sch == F3D_SCHEME_UP7) {
    for (int q = 0; q < 7; ++q)
      for (int n = 0; n < 5; ++n) {
        gfp[q][n] += kp.lin[q] * gF[n];
        gfm[7 - q][n] += kp.lin[q] * gF[n];
      }
    return;
  }
  if (sch == F3D_SCHEME_NND2) {
    for (int n = 0; n < 5; ++n) {
      // fp = FP(-1) + 1/2 minmod(FP(0) - FP(-1), FP(-1) - FP(-2))
      gfp[3][n] += gF[n];
      {
        const double a = fp[4][n] - fp[3][n], b = fp[3][n] - fp[2][n], g = 0.5 * gF[n];
        if (a * b > 0) {
          if (::fabs(a) > ::fabs(b)) { gfp[3][n] += g; gfp[2][n] -= g; }
          else { gfp[4][n] += g; gfp[3][n] -= g; }
        }
      }
      // fm = FM(0) - 1/2 minmod(FM(1) - FM(0), FM(0) - FM(-1))
      gfm[4][n] += gF[n];
      {
        const double a = fm[5][n] - fm[4][n], b = fm[4][n] - fm[3][n], g = -0.5 * gF[n];
        if (a * b > 0) {
          if (::fabs(a) > ::fabs(b)) { gfm[4][n] += g; gfm[3][n] -= g; }
          else { gfm[5][n] += g; gfm[4][n] -= g; }
        }
      }
    }
    return;
  }
  double s0, nv[3], lv[3], mv[3];
  face_geometry(AL, AR, SL, SR, s0, nv, lv, mv);
  const double th[4] = {0.5 * (qR[0] + qL[0]), 0.5 * (qR[1] + qL[1]), 0.5 * (qR[2] + qL[2]), 0.5 * (qR[3] + qL[3])};
  double L[5][5], Rr[5][5];
  char_frame<double>(kp, th[0], th[1], th[2], th[3], nv, lv, mv, L, Rr);
  const bool w7 = sch == F3D_SCHEME_WENO7;
  const int ns = w7 ? 7 : 5;
  const double eps = w7 ? kp.eps7 : kp.eps5;
  const double scale = w7 ? 1.0 / 12.0 : 1.0 / 6.0;
  double vp[5][7], vm[5][7], flux[5];
  for (int n = 0; n < 5; ++n) {
    for (int s = 0; s < ns; ++s) {
      const int ip = w7 ? s : s + 1, im = w7 ? 7 - s : 6 - s;
      double a = 0, b = 0;
      for (int q = 0; q < 5; ++q) {
        a += fp[ip][q] * L[n][q];
        b += fm[im][q] * L[n][q];
      }
      vp[n][s] = a;
      vm[n][s] = b;
    }
    flux[n] = w7 ? (weno7<double>(vp[n], s0, eps) + weno7<double>(vm[n], s0, eps)) * scale
                 : (weno5<double>(vp[n], s0, eps) + weno5<double>(vm[n], s0, eps)) * scale;
  }
  double gL[5][5] = {}, gRr[5][5];
  double gflux[5];
  for (int n = 0; n < 5; ++n) {
    double a = 0;
    for (int m = 0; m < 5; ++m) {
      a += Rr[m][n] * gF[m];
      gRr[m][n] = gF[m] * flux[n];
    }
    gflux[n] = a;
  }
  for (int n = 0; n < 5; ++n) {
    double gvp[7] = {0, 0, 0, 0, 0, 0, 0}, gvm[7] = {0, 0, 0, 0, 0, 0, 0};
    const double gw = gflux[n] * scale;
    if (w7) {
      weno7_vjp(vp[n], s0, eps, gw, gvp);
      weno7_vjp(vm[n], s0, eps, gw, gvm);
    } else {
      weno5_vjp(vp[n], s0, eps, gw, gvp);
      weno5_vjp(vm[n], s0, eps, gw, gvm);
    }
    for (int s = 0; s < ns; ++s) {
      const int ip = w7 ? s : s + 1, im = w7 ? 7 - s : 6 - s;
      for (int q = 0; q < 5; ++q) {
        gfp[ip][q] += L[n][q] * gvp[s];
        gfm[im][q] += L[n][q] * gvm[s];
        gL[n][q] += fp[ip][q] * gvp[s] + fm[im][q] * gvm[s];
      }
    }
  }
  // frame: one forward-mode pass per averaged input (u, v, w, T)
  for (int k = 0; k < 4; ++k) {
    Dv t[4] = {Dv(th[0]), Dv(th[1]), Dv(th[2]), Dv(th[3])};
    t[k].d = 1.0;
    Dv dL[5][5], dR[5][5];
    char_frame<Dv>(kp, t[0], t[1], t[2], t[3], nv, lv, mv, dL, dR);
    double g = 0;
    for (int a = 0; a < 5; ++a)
      for (int b = 0; b < 5; ++b) g += gL[a][b] * dL[a][b].d + gRr[a][b] * dR[a][b].d;
    gqL[k] += 0.5 * g;
    gqR[k] += 0.5 * g;
  }
}

// ------------------------------------------------------------------ viscous cell, reverse
// Reverse of viscous_cell: go[d][4] = adjoint of the cell's viscous fluxes in direction d.
// Returns gg[4][3] (adjoint of the computational derivatives of u,v,w,T) and gval[4] (adjoint of
// u, v, w, T at the cell, including mu(T) and lambda(T)).
F3D_HD void viscous_cell_vjp(const KP& kp, double u, double v, double w, double Tt, const double g[4][3],
                             const double* m, double Jac, const double go[3][4], double gg[4][3], double* gval) {
  const double sT = ::sqrt(Tt);
  const double mu = kp.Cs * Tt * sT / (Tt + kp.Ts);
  const double cpr = kp.Cp / kp.Pr;
  const double lam = cpr * mu;
  double dx[4], dy[4], dz[4];
  for (int q = 0; q < 4; ++q) {
    dx[q] = (g[q][0] * m[0] + g[q][1] * m[3] + g[q][2] * m[6]) * Jac;
    dy[q] = (g[q][0] * m[1] + g[q][1] * m[4] + g[q][2] * m[7]) * Jac;
    dz[q] = (g[q][0] * m[2] + g[q][1] * m[5] + g[q][2] * m[8]) * Jac;
  }
  const double c23 = 2.0 / 3.0;
  const double dv = dx[0] + dy[1] + dz[2];
  const double S11 = 2 * dx[0] - c23 * dv, S22 = 2 * dy[1] - c23 * dv, S33 = 2 * dz[2] - c23 * dv;
  const double S12 = dy[0] + dx[1], S13 = dz[0] + dx[2], S23 = dy[2] + dz[1];
  const double t11 = mu * S11, t12 = mu * S12, t13 = mu * S13, t22 = mu * S22, t23 = mu * S23, t33 = mu * S33;
  double g11 = 0, g12 = 0, g13 = 0, g22 = 0, g23 = 0, g33 = 0, gE1 = 0, gE2 = 0, gE3 = 0;
  for (int d = 0; d < 3; ++d) {
    const double a = m[3 * d], b = m[3 * d + 1], e = m[3 * d + 2];
    g11 += a * go[d][0];
    g12 += b * go[d][0] + a * go[d][1];
    g13 += e * go[d][0] + a * go[d][2];
    g22 += b * go[d][1];
    g23 += e * go[d][1] + b * go[d][2];
    g33 += e * go[d][2];
    gE1 += a * go[d][3];
    gE2 += b * go[d][3];
    gE3 += e * go[d][3];
  }
  g11 += u * gE1;
  g12 += v * gE1 + u * gE2;
  g13 += w * gE1 + u * gE3;
  g22 += v * gE2;
  g23 += w * gE2 + v * gE3;
  g33 += w * gE3;
  gval[0] = t11 * gE1 + t12 * gE2 + t13 * gE3;
  gval[1] = t12 * gE1 + t22 * gE2 + t23 * gE3;
  gval[2] = t13 * gE1 + t23 * gE2 + t33 * gE3;
  const double glam = dx[3] * gE1 + dy[3] * gE2 + dz[3] * gE3;
  const double gmu = g11 * S11 + g12 * S12 + g13 * S13 + g22 * S22 + g23 * S23 + g33 * S33 + glam * cpr;
  gval[3] = gmu * mu * (1.5 / Tt - 1.0 / (Tt + kp.Ts));
  double gdx[4], gdy[4], gdz[4];
  const double gdv = -c23 * mu * (g11 + g22 + g33);
  gdx[0] = 2 * mu * g11 + gdv;
  gdy[1] = 2 * mu * g22 + gdv;
  gdz[2] = 2 * mu * g33 + gdv;
  gdy[0] = mu * g12; gdx[1] = mu * g12;
  gdz[0] = mu * g13; gdx[2] = mu * g13;
  gdy[2] = mu * g23; gdz[1] = mu * g23;
  gdx[3] = lam * gE1; gdy[3] = lam * gE2; gdz[3] = lam * gE3;
  for (int q = 0; q < 4; ++q) {
    gg[q][0] = Jac * (m[0] * gdx[q] + m[1] * gdy[q] + m[2] * gdz[q]);
    gg[q][1] = Jac * (m[3] * gdx[q] + m[4] * gdy[q] + m[5] * gdz[q]);
    gg[q][2] = Jac * (m[6] * gdx[q] + m[7] * gdy[q] + m[8] * gdz[q]);
  }
}

// ------------------------------------------------------------------ c2Prim, reverse
// Reverse of c2Prim (utils.jl:2-30) at one cell: U[5] conserved, gQ[6] adjoint of
// (rho,u,v,w,p,T) -> gU[5].  The eps floors pass no derivative when active.
F3D_HD void c2prim_vjp(const KP& kp, const double* U, const double* gQ, double* gU) {
  const double eps = 2.220446049250313e-16;
  const double rho = jmax(U[0], eps);
  const double ri = 1.0 / rho;
  const double u = U[1] * ri, v = U[2] * ri, w = U[3] * ri;
  const double k2 = u * u + v * v + w * w;
  const double eraw = U[4] - 0.5 * rho * k2;
  const double ei = jmax(eraw, eps);
  const double p = (kp.gamma - 1) * ei;
  const double Tt = p / (rho * kp.Rg);
  const double gp = gQ[4] + gQ[5] / (rho * kp.Rg);
  double grho = gQ[0] - gQ[5] * Tt / rho;
  const double gei = (eraw >= eps) ? (kp.gamma - 1) * gp : 0.0;
  gU[4] = gei;
  grho -= 0.5 * k2 * gei;
  const double gu = gQ[1] - rho * u * gei, gv = gQ[2] - rho * v * gei, gw = gQ[3] - rho * w * gei;
  gU[1] = gu * ri;
  gU[2] = gv * ri;
  gU[3] = gw * ri;
  grho -= (gu * U[1] + gv * U[2] + gw * U[3]) * ri * ri;
  gU[0] = (U[0] >= eps) ? grho : 0.0;
}

}  // namespace f3d
