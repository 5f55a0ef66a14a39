// CPU oracle for the Flame3D finite-difference RHS + RK3 stage path.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain restatement, on the CPU, of the reference kernel
// bodies of rygan-code/Flame3D-MHD (Julia + CUDA.jl), written from reading those files.  It is the
// checker for the HIP library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it, and never as the thing measured or shipped.
//
// Conventions: 1-based (i,j,k,n) exactly as in the Julia source so that each loop can be read
// against the cited lines; arrays are column-major with i fastest.  T is the "bulk" type (double
// for the FP64 oracle, float for the FP32-bulk form); conserved U, Un and J are always double.
//
// Pinning status of each function (what checks it besides GPU parity) is listed in the header of
// oracle/f3d_oracle.cpp.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#include "../include/f3d.h"
#include "f3d_dual.hpp"

namespace f3d_oracle {

// Extents of one 4-D column-major array (n1,n2,n3,nv), 1-based accessors.
struct Ext {
  long n1, n2, n3;
  size_t at(long i, long j, long k) const {
    return (size_t)(i - 1) + (size_t)n1 * ((size_t)(j - 1) + (size_t)n2 * (size_t)(k - 1));
  }
  size_t at(long i, long j, long k, long n) const {
    return at(i, j, k) + (size_t)n1 * n2 * n3 * (size_t)(n - 1);
  }
};

struct Grid {
  long Nxp, Nyp, Nzp, NG, Nxt, Nyt, Nzt;
  Ext cell, fx, fy, fz, fv;
  explicit Grid(const f3d_params* p)
      : Nxp(p->Nxp), Nyp(p->Nyp), Nzp(p->Nzp), NG(p->NG), Nxt(p->Nxp + 2 * p->NG),
        Nyt(p->Nyp + 2 * p->NG), Nzt(p->Nzp + 2 * p->NG) {
    cell = {Nxt, Nyt, Nzt};
    fx = {Nxp + 1, Nyp, Nzp};
    fy = {Nxp, Nyp + 1, Nzp};
    fz = {Nxp, Nyp, Nzp + 1};
    fv = {Nxp + NG, Nyp + NG, Nzp + NG};
  }
};

// Julia's max for floats propagates NaN (Base.max); C's std::max does not.
template <typename T>
inline T jmax(T a, T b) {
  if (a != a) return a;
  if (b != b) return b;
  return (b > a) ? b : a;
}
// NaN-propagating min (Julia Base.min on floats).
template <typename T>
inline T jmin(T a, T b) {
  return (a != a) ? a : ((b != b) ? b : ((b < a) ? b : a));
}

// Linear = UP7*Linear_phi + CD6*(1-Linear_phi)   run.jl:75-77
template <typename T>
inline void linear_coeffs(const f3d_params* p, T c[7]) {
  const double up7[7] = {-3.0 / 420, 25.0 / 420, -101.0 / 420, 319.0 / 420,
                         214.0 / 420, -38.0 / 420, 4.0 / 420};
  const double cd6[7] = {0.0, 1.0 / 60, -2.0 / 15, 37.0 / 60, 37.0 / 60, -2.0 / 15, 1.0 / 60};
  for (int m = 0; m < 7; ++m)
    c[m] = (T)(up7[m] * p->Linear_phi + cd6[m] * (1.0 - p->Linear_phi));
}

// ---------------------------------------------------------------- utils.jl
// c2Prim  utils.jl:2-30 (interior [NG+1, Np+NG]^3).  The eps floors are the machine epsilon of the
// type the conversion runs in (the reference uses eps(Float32) in its Float32 build).
template <typename T, typename TU = double>
void c2prim(const f3d_params* p, const TU* U, T* Q) {
  Grid g(p);
  const double gam = p->gamma, Rg = p->Rg, eps = std::numeric_limits<double>::epsilon();
#pragma omp parallel for collapse(2)
  for (long k = g.NG + 1; k <= g.Nzp + g.NG; ++k)
    for (long j = g.NG + 1; j <= g.Nyp + g.NG; ++j)
      for (long i = g.NG + 1; i <= g.Nxp + g.NG; ++i) {
        TU rho = jmax(U[g.cell.at(i, j, k, 1)], TU(eps));
        TU rinv = 1 / rho;
        TU u = U[g.cell.at(i, j, k, 2)] * rinv;
        TU v = U[g.cell.at(i, j, k, 3)] * rinv;
        TU w = U[g.cell.at(i, j, k, 4)] * rinv;
        TU ei = jmax(U[g.cell.at(i, j, k, 5)] - 0.5 * rho * (u * u + v * v + w * w), TU(eps));
        TU pr = (gam - 1) * ei;
        TU Tt = pr / (rho * Rg);
        Q[g.cell.at(i, j, k, 1)] = (T)rho;
        Q[g.cell.at(i, j, k, 2)] = (T)u;
        Q[g.cell.at(i, j, k, 3)] = (T)v;
        Q[g.cell.at(i, j, k, 4)] = (T)w;
        Q[g.cell.at(i, j, k, 5)] = (T)pr;
        Q[g.cell.at(i, j, k, 6)] = (T)Tt;
      }
}

// prim2c  utils.jl:33-54 (whole array incl. ghosts)
template <typename T>
void prim2c(const f3d_params* p, double* U, const T* Q) {
  Grid g(p);
  const double gam = p->gamma;
#pragma omp parallel for collapse(2)
  for (long k = 1; k <= g.Nzt; ++k)
    for (long j = 1; j <= g.Nyt; ++j)
      for (long i = 1; i <= g.Nxt; ++i) {
        double rho = Q[g.cell.at(i, j, k, 1)], u = Q[g.cell.at(i, j, k, 2)];
        double v = Q[g.cell.at(i, j, k, 3)], w = Q[g.cell.at(i, j, k, 4)];
        double ei = (double)Q[g.cell.at(i, j, k, 5)] / (gam - 1);
        U[g.cell.at(i, j, k, 1)] = rho;
        U[g.cell.at(i, j, k, 2)] = u * rho;
        U[g.cell.at(i, j, k, 3)] = v * rho;
        U[g.cell.at(i, j, k, 4)] = w * rho;
        U[g.cell.at(i, j, k, 5)] = ei + 0.5 * rho * (u * u + v * v + w * w);
      }
}

// linComb  utils.jl:57-70
inline void lin_comb(const f3d_params* p, double* U, const double* Un, int NV, double a, double b) {
  Grid g(p);
  for (long n = 1; n <= NV; ++n)
#pragma omp parallel for collapse(2)
    for (long k = g.NG + 1; k <= g.Nzp + g.NG; ++k)
      for (long j = g.NG + 1; j <= g.Nyp + g.NG; ++j)
        for (long i = g.NG + 1; i <= g.Nxp + g.NG; ++i) {
          size_t o = g.cell.at(i, j, k, n);
          U[o] = U[o] * a + Un[o] * b;
        }
}

// ---------------------------------------------------------------- schemes.jl
// shockSensor  schemes.jl:1-24, domain [2, N_tot-1]^3
template <typename T>
void shock_sensor(const f3d_params* p, T* phi, const T* Q) {
  Grid g(p);
#pragma omp parallel for collapse(2)
  for (long k = 2; k <= g.Nzt - 1; ++k)
    for (long j = 2; j <= g.Nyt - 1; ++j)
      for (long i = 2; i <= g.Nxt - 1; ++i) {
        T Px1 = Q[g.cell.at(i - 1, j, k, 5)], Px2 = Q[g.cell.at(i, j, k, 5)];
        T Px3 = Q[g.cell.at(i + 1, j, k, 5)];
        T Py1 = Q[g.cell.at(i, j - 1, k, 5)], Py2 = Px2, Py3 = Q[g.cell.at(i, j + 1, k, 5)];
        T Pz1 = Q[g.cell.at(i, j, k - 1, 5)], Pz2 = Px2, Pz3 = Q[g.cell.at(i, j, k + 1, 5)];
        T px = std::fabs(-Px1 + 2 * Px2 - Px3) / (Px1 + 2 * Px2 + Px3);
        T py = std::fabs(-Py1 + 2 * Py2 - Py3) / (Py1 + 2 * Py2 + Py3);
        T pz = std::fabs(-Pz1 + 2 * Pz2 - Pz3) / (Pz1 + 2 * Pz2 + Pz3);
        phi[g.cell.at(i, j, k)] = px + py + pz;
      }
}

// minmod  schemes.jl:26-28
template <typename T>
inline T minmod(T a, T b) {
  return (a * b > 0) ? ((std::fabs(a) > std::fabs(b)) ? b : a) : T(0);
}

// Direction helper: face loop bounds and stencil offsets of advect_{x,y,z}c.
struct Dir {
  int d;  // 0,1,2
  long di, dj, dk;
};
inline Dir make_dir(int d) {
  return Dir{d, d == 0 ? 1L : 0L, d == 1 ? 1L : 0L, d == 2 ? 1L : 0L};
}

// Scheme chosen on one face: max of phi over the 6 cells i-3..i+2 (schemes.jl:724-729), then the
// thresholds of schemes.jl:731, 813, 931.
template <typename T>
inline int face_scheme(const f3d_params* p, const Grid& g, const T* phi, long i, long j, long k,
                       const Dir& D, T* phix_out) {
  T m = phi[g.cell.at(i - 3 * D.di, j - 3 * D.dj, k - 3 * D.dk)];
  for (int o = -2; o <= 2; ++o) m = jmax(m, phi[g.cell.at(i + o * D.di, j + o * D.dj, k + o * D.dk)]);
  if (phix_out) *phix_out = m;
  if (m < (T)p->hybrid_phi1) return F3D_SCHEME_UP7;
  if (m < (T)p->hybrid_phi2) return F3D_SCHEME_WENO7;
  if (m < (T)p->hybrid_phi3) return F3D_SCHEME_WENO5;
  return F3D_SCHEME_NND2;
}

template <typename T>
void scheme_mask(const f3d_params* p, uint8_t* mx, uint8_t* my, uint8_t* mz, const T* phi) {
  Grid g(p);
  uint8_t* out[3] = {mx, my, mz};
  const Ext* fe[3] = {&g.fx, &g.fy, &g.fz};
  for (int d = 0; d < 3; ++d) {
    if (!out[d]) continue;
    Dir D = make_dir(d);
#pragma omp parallel for collapse(2)
    for (long k = g.NG + 1; k <= g.Nzp + g.NG + D.dk; ++k)
      for (long j = g.NG + 1; j <= g.Nyp + g.NG + D.dj; ++j)
        for (long i = g.NG + 1; i <= g.Nxp + g.NG + D.di; ++i)
          out[d][fe[d]->at(i - g.NG, j - g.NG, k - g.NG)] =
              (uint8_t)face_scheme(p, g, phi, i, j, k, D, (T*)nullptr);
  }
}

// Ratio bmin/|b| of the range-safe float weights (bmin the smallest |b| of the stencil, so the
// ratio lies in [0, 1] even where rounding makes a smoothness indicator slightly negative); a
// vanishing |b|, whose weight is unbounded, gets 1 and every other stencil 0.
inline float wratio(float bmin, float b) {
  const float ab = std::fabs(b);
  return ab > 0.f ? bmin / ab : 1.f;
}

// WENO7 nonlinear weights (schemes.jl:871-879).  The reference evaluates them in Float64 (its
// epsilon is a Float64 literal); for T=double that is the literal formula.  For T=float the same
// weights are formed range-safely: every (eps+Is*s0)^2 is divided by the smallest one before
// inverting, which leaves the normalised weights unchanged and avoids 1/(1e-20)^2 overflowing.
template <typename T>
struct Weno7Weights;
template <>
struct Weno7Weights<double> {
  static double combine(double eps, double s0, const double Is[4], const double q[4]) {
    double a1 = 1 / ((eps + Is[0] * s0) * (eps + Is[0] * s0));
    double a2 = 12 / ((eps + Is[1] * s0) * (eps + Is[1] * s0));
    double a3 = 18 / ((eps + Is[2] * s0) * (eps + Is[2] * s0));
    double a4 = 4 / ((eps + Is[3] * s0) * (eps + Is[3] * s0));
    double invsum = 1 / (a1 + a2 + a3 + a4);
    return invsum * (a1 * q[0] + a2 * q[1] + a3 * q[2] + a4 * q[3]);
  }
};
template <>
struct Weno7Weights<long double> {
  static long double combine(long double eps, long double s0, const long double Is[4], const long double q[4]) {
    const long double d[4] = {1, 12, 18, 4};
    long double a[4], sum = 0, acc = 0;
    for (int m = 0; m < 4; ++m) {
      const long double b = eps + Is[m] * s0;
      a[m] = d[m] / (b * b);
      sum += a[m];
      acc += a[m] * q[m];
    }
    return acc / sum;
  }
};
template <>
struct Weno7Weights<Dual> {  // literal form, as for double
  static Dual combine(Dual eps, Dual s0, const Dual Is[4], const Dual q[4]) {
    const double d[4] = {1, 12, 18, 4};
    Dual a[4], sum = 0, acc = 0;
    for (int m = 0; m < 4; ++m) {
      const Dual b = eps + Is[m] * s0;
      a[m] = d[m] / (b * b);
      sum += a[m];
      acc += a[m] * q[m];
    }
    return acc / sum;
  }
};
template <>
struct Weno7Weights<float> {
  static float combine(float eps, float s0, const float Is[4], const float q[4]) {
    float b[4];
    for (int m = 0; m < 4; ++m) b[m] = eps + Is[m] * s0;
    float bmin = std::min(std::min(std::fabs(b[0]), std::fabs(b[1])), std::min(std::fabs(b[2]), std::fabs(b[3])));
    const float d[4] = {1.f, 12.f, 18.f, 4.f};
    float a[4], sum = 0.f, acc = 0.f;
    for (int m = 0; m < 4; ++m) {
      float r = wratio(bmin, b[m]);
      a[m] = d[m] * (r * r);
      sum += a[m];
      acc += a[m] * q[m];
    }
    return acc / sum;
  }
};

// The four WENO7 smoothness indicators as quadratic forms of the first differences
// a_k = V_{k+1} - V_k: the reference's quadratic forms in V (schemes.jl:848-869) regrouped (each is
// invariant to adding a constant to V), the same polynomials exactly in rational arithmetic.  The
// float form evaluates them this way: on smooth data the V form loses the indicator to the
// cancellation of O(V^2) terms (float's 6e-8 x 17246 V^2), which sets the weights by rounding noise.
template <typename T>
inline void weno7_is_diff(const T* V, T Is[4]) {
  const T a1 = V[1] - V[0], a2 = V[2] - V[1], a3 = V[3] - V[2], a4 = V[4] - V[3], a5 = V[5] - V[4],
          a6 = V[6] - V[5];
  Is[0] = a1 * (T(547) * a1 - T(2788) * a2 + T(1854) * a3) + a2 * (T(3708) * a2 - T(5188) * a3) + a3 * (T(2107) * a3);
  Is[1] = a2 * (T(267) * a2 - T(1108) * a3 + T(494) * a4) + a3 * (T(1468) * a3 - T(1428) * a4) + a4 * (T(547) * a4);
  Is[2] = a3 * (T(547) * a3 - T(1428) * a4 + T(494) * a5) + a4 * (T(1468) * a4 - T(1108) * a5) + a5 * (T(267) * a5);
  Is[3] = a4 * (T(2107) * a4 - T(5188) * a5 + T(1854) * a6) + a5 * (T(3708) * a5 - T(2788) * a6) + a6 * (T(547) * a6);
}

// One WENO7 reconstruction of a 7-point characteristic stencil V[0..6]  (schemes.jl:847-879)
template <typename T>
inline T weno7(const T* V, T s0, T eps) {
  const T V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4], V6 = V[5], V7 = V[6];
  T q[4], Is[4];
  q[0] = -3 * V1 + 13 * V2 - 23 * V3 + 25 * V4;
  q[1] = V2 - 5 * V3 + 13 * V4 + 3 * V5;
  q[2] = -V3 + 7 * V4 + 7 * V5 - V6;
  q[3] = 3 * V4 + 13 * V5 - 5 * V6 + V7;
  if (std::is_same<T, float>::value) {
    weno7_is_diff(V, Is);
    return Weno7Weights<T>::combine(eps, s0, Is, q);
  }
  Is[0] = V1 * (547 * V1 - 3882 * V2 + 4642 * V3 - 1854 * V4) +
          V2 * (7043 * V2 - 17246 * V3 + 7042 * V4) + V3 * (11003 * V3 - 9402 * V4) +
          V4 * (2107 * V4);
  Is[1] = V2 * (267 * V2 - 1642 * V3 + 1602 * V4 - 494 * V5) +
          V3 * (2843 * V3 - 5966 * V4 + 1922 * V5) + V4 * (3443 * V4 - 2522 * V5) +
          V5 * (547 * V5);
  Is[2] = V3 * (547 * V3 - 2522 * V4 + 1922 * V5 - 494 * V6) +
          V4 * (3443 * V4 - 5966 * V5 + 1602 * V6) + V5 * (2843 * V5 - 1642 * V6) +
          V6 * (267 * V6);
  Is[3] = V4 * (2107 * V4 - 9402 * V5 + 7042 * V6 - 1854 * V7) +
          V5 * (11003 * V5 - 17246 * V6 + 4642 * V7) + V6 * (7043 * V6 - 3882 * V7) +
          V7 * (547 * V7);
  return Weno7Weights<T>::combine(eps, s0, Is, q);
}

// Forward mode of weno7: the tangent of a smoothness indicator below the rounding floor of its own
// evaluation (Is < 1e-12 x the quadratic form on |V| with |coefficients|) is dropped, so a stencil
// flat to the last bit keeps its weight frozen (with eps = 1e-20 its weight is 1/noise^2 and the
// derivative a ratio of noise terms).  The adjoint uses the same convention.
template <>
inline Dual weno7<Dual>(const Dual* V, Dual s0, Dual eps) {
  const Dual V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4], V6 = V[5], V7 = V[6];
  Dual q[4], Is[4];
  q[0] = -3 * V1 + 13 * V2 - 23 * V3 + 25 * V4;
  q[1] = V2 - 5 * V3 + 13 * V4 + 3 * V5;
  q[2] = -V3 + 7 * V4 + 7 * V5 - V6;
  q[3] = 3 * V4 + 13 * V5 - 5 * V6 + V7;
  Is[0] = V1 * (547 * V1 - 3882 * V2 + 4642 * V3 - 1854 * V4) + V2 * (7043 * V2 - 17246 * V3 + 7042 * V4) +
          V3 * (11003 * V3 - 9402 * V4) + V4 * (2107 * V4);
  Is[1] = V2 * (267 * V2 - 1642 * V3 + 1602 * V4 - 494 * V5) + V3 * (2843 * V3 - 5966 * V4 + 1922 * V5) +
          V4 * (3443 * V4 - 2522 * V5) + V5 * (547 * V5);
  Is[2] = V3 * (547 * V3 - 2522 * V4 + 1922 * V5 - 494 * V6) + V4 * (3443 * V4 - 5966 * V5 + 1602 * V6) +
          V5 * (2843 * V5 - 1642 * V6) + V6 * (267 * V6);
  Is[3] = V4 * (2107 * V4 - 9402 * V5 + 7042 * V6 - 1854 * V7) + V5 * (11003 * V5 - 17246 * V6 + 4642 * V7) +
          V6 * (7043 * V6 - 3882 * V7) + V7 * (547 * V7);
  const double a1 = std::fabs(V1.v), a2 = std::fabs(V2.v), a3 = std::fabs(V3.v), a4 = std::fabs(V4.v),
               a5 = std::fabs(V5.v), a6 = std::fabs(V6.v), a7 = std::fabs(V7.v);
  const double Ia[4] = {
      a1 * (547 * a1 + 3882 * a2 + 4642 * a3 + 1854 * a4) + a2 * (7043 * a2 + 17246 * a3 + 7042 * a4) +
          a3 * (11003 * a3 + 9402 * a4) + a4 * (2107 * a4),
      a2 * (267 * a2 + 1642 * a3 + 1602 * a4 + 494 * a5) + a3 * (2843 * a3 + 5966 * a4 + 1922 * a5) +
          a4 * (3443 * a4 + 2522 * a5) + a5 * (547 * a5),
      a3 * (547 * a3 + 2522 * a4 + 1922 * a5 + 494 * a6) + a4 * (3443 * a4 + 5966 * a5 + 1602 * a6) +
          a5 * (2843 * a5 + 1642 * a6) + a6 * (267 * a6),
      a4 * (2107 * a4 + 9402 * a5 + 7042 * a6 + 1854 * a7) + a5 * (11003 * a5 + 17246 * a6 + 4642 * a7) +
          a6 * (7043 * a6 + 3882 * a7) + a7 * (547 * a7)};
  for (int m = 0; m < 4; ++m)
    if (Is[m].v < 1e-12 * Ia[m]) Is[m].d = 0.0;
  return Weno7Weights<Dual>::combine(eps, s0, Is, q);
}

// One WENO5 reconstruction of a 5-point stencil  (schemes.jl:958-972)
template <typename T>
inline T weno5(const T* V, T s0, T eps) {
  const T V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4];
  const T a1 = V1 - 2 * V2 + V3, b1 = V1 - 4 * V2 + 3 * V3;
  const T a2 = V2 - 2 * V3 + V4, b2 = V2 - V4;
  const T a3 = V3 - 2 * V4 + V5, b3 = 3 * V3 - 4 * V4 + V5;
  T s11 = 13 * (a1 * a1) + 3 * (b1 * b1);
  T s22 = 13 * (a2 * a2) + 3 * (b2 * b2);
  T s33 = 13 * (a3 * a3) + 3 * (b3 * b3);
  s11 = 1 / ((eps + s11 * s0) * (eps + s11 * s0));
  s22 = 6 / ((eps + s22 * s0) * (eps + s22 * s0));
  s33 = 3 / ((eps + s33 * s0) * (eps + s33 * s0));
  T invsum = 1 / (s11 + s22 + s33);
  T v1 = 2 * V1 - 7 * V2 + 11 * V3;
  T v2 = -V2 + 5 * V3 + 2 * V4;
  T v3 = 2 * V3 + 5 * V4 - V5;
  return invsum * (s11 * v1 + s22 * v2 + s33 * v3);
}

// Forward mode of weno5 with the rounding-floor convention of weno7<Dual>: Is = 13 a^2 + 3 c^2 below
// 1e-24 x (13 A^2 + 3 C^2) (A, C the |V|-sums of the differences a, c) passes no tangent.
template <>
inline Dual weno5<Dual>(const Dual* V, Dual s0, Dual eps) {
  const Dual V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4];
  const Dual a1 = V1 - 2 * V2 + V3, b1 = V1 - 4 * V2 + 3 * V3;
  const Dual a2 = V2 - 2 * V3 + V4, b2 = V2 - V4;
  const Dual a3 = V3 - 2 * V4 + V5, b3 = 3 * V3 - 4 * V4 + V5;
  Dual s11 = 13 * (a1 * a1) + 3 * (b1 * b1);
  Dual s22 = 13 * (a2 * a2) + 3 * (b2 * b2);
  Dual s33 = 13 * (a3 * a3) + 3 * (b3 * b3);
  const double x1 = std::fabs(V1.v), x2 = std::fabs(V2.v), x3 = std::fabs(V3.v), x4 = std::fabs(V4.v),
               x5 = std::fabs(V5.v);
  auto fl = [](double A, double C) { return 1e-24 * (13 * A * A + 3 * C * C); };
  if (s11.v < fl(x1 + 2 * x2 + x3, x1 + 4 * x2 + 3 * x3)) s11.d = 0.0;
  if (s22.v < fl(x2 + 2 * x3 + x4, x2 + x4)) s22.d = 0.0;
  if (s33.v < fl(x3 + 2 * x4 + x5, 3 * x3 + 4 * x4 + x5)) s33.d = 0.0;
  s11 = 1 / ((eps + s11 * s0) * (eps + s11 * s0));
  s22 = 6 / ((eps + s22 * s0) * (eps + s22 * s0));
  s33 = 3 / ((eps + s33 * s0) * (eps + s33 * s0));
  Dual invsum = 1 / (s11 + s22 + s33);
  Dual v1 = 2 * V1 - 7 * V2 + 11 * V3;
  Dual v2 = -V2 + 5 * V3 + 2 * V4;
  Dual v3 = 2 * V3 + 5 * V4 - V5;
  return invsum * (s11 * v1 + s22 * v2 + s33 * v3);
}

// advect_xc / advect_yc / advect_zc  schemes.jl:700-1012 / 1014-1328 / 1330-1643 (index
// permutations of one another).  F has the face extents of direction d.
template <typename T>
void advect_c(const f3d_params* p, int d, T* F, const T* phi, const T* S, const T* Fp,
              const T* Fm, const T* Q, const T* Ax, const T* Ay, const T* Az) {
  Grid g(p);
  Dir D = make_dir(d);
  const Ext& fe = d == 0 ? g.fx : (d == 1 ? g.fy : g.fz);
  T c[7];
  linear_coeffs<T>(p, c);
  const T tmp1 = T(1) / T(12), tmp2 = T(1) / T(6);
  const T eps7 = (T)p->weno7_eps, eps5 = (T)p->weno5_eps;
  const T gam = (T)p->gamma, Rg = (T)p->Rg;
#pragma omp parallel for collapse(2)
  for (long k = g.NG + 1; k <= g.Nzp + g.NG + D.dk; ++k)
    for (long j = g.NG + 1; j <= g.Nyp + g.NG + D.dj; ++j)
      for (long i = g.NG + 1; i <= g.Nxp + g.NG + D.di; ++i) {
        auto C = [&](long o) { return g.cell.at(i + o * D.di, j + o * D.dj, k + o * D.dk); };
        auto out = [&](long n) { return fe.at(i - g.NG, j - g.NG, k - g.NG, n); };
        const size_t ncell = (size_t)g.Nxt * g.Nyt * g.Nzt;
        T phix;
        int sch = face_scheme(p, g, phi, i, j, k, D, &phix);
        if (sch == F3D_SCHEME_UP7) {  // schemes.jl:731-754
          for (long n = 1; n <= 5; ++n) {
            const size_t nn = ncell * (n - 1);
            T V1 = Fp[C(-4) + nn], V2 = Fp[C(-3) + nn], V3 = Fp[C(-2) + nn], V4 = Fp[C(-1) + nn];
            T V5 = Fp[C(0) + nn], V6 = Fp[C(1) + nn], V7 = Fp[C(2) + nn];
            T fp = c[0] * V1 + c[1] * V2 + c[2] * V3 + c[3] * V4 + c[4] * V5 + c[5] * V6 + c[6] * V7;
            V1 = Fm[C(3) + nn]; V2 = Fm[C(2) + nn]; V3 = Fm[C(1) + nn]; V4 = Fm[C(0) + nn];
            V5 = Fm[C(-1) + nn]; V6 = Fm[C(-2) + nn]; V7 = Fm[C(-3) + nn];
            T fm = c[0] * V1 + c[1] * V2 + c[2] * V3 + c[3] * V4 + c[4] * V5 + c[5] * V6 + c[6] * V7;
            F[out(n)] = fp + fm;
          }
          continue;
        }
        // characteristic setup  schemes.jl:756-811
        T s0 = 2 / (S[C(-1)] + S[C(0)]);
        T u = T(0.5) * (Q[C(0) + ncell * 1] + Q[C(-1) + ncell * 1]);
        T v = T(0.5) * (Q[C(0) + ncell * 2] + Q[C(-1) + ncell * 2]);
        T w = T(0.5) * (Q[C(0) + ncell * 3] + Q[C(-1) + ncell * 3]);
        T Tt = T(0.5) * (Q[C(0) + ncell * 5] + Q[C(-1) + ncell * 5]);
        T A1 = (Ax[C(0)] + Ax[C(-1)]) * T(0.5);
        T A2 = (Ay[C(0)] + Ay[C(-1)]) * T(0.5);
        T A3 = (Az[C(0)] + Az[C(-1)]) * T(0.5);
        T nx = A1 * s0, ny = A2 * s0, nz = A3 * s0;
        T lx, ly, lz, ss;
        if (std::fabs(nz) <= std::fabs(ny)) {
          ss = 1 / std::sqrt(nx * nx + ny * ny);
          lx = -ny * ss; ly = nx * ss; lz = 0;
        } else {
          ss = 1 / std::sqrt(nx * nx + nz * nz);
          lx = -nz * ss; ly = 0; lz = nx * ss;
        }
        T mx = ny * lz - nz * ly, my = nz * lx - nx * lz, mz = nx * ly - ny * lx;
        T qn = u * nx + v * ny + w * nz;
        T ql = u * lx + v * ly + w * lz;
        T qm = u * mx + v * my + w * mz;
        T q2 = T(0.5) * (u * u + v * v + w * w);
        T cc = std::sqrt(gam * Rg * Tt);
        T c2 = 1 / (2 * cc);
        T K = (gam - 1) / (cc * cc);
        T K2 = K * T(0.5);
        T H = 1 / K + q2;
        T L[5][5] = {{K2 * q2 + qn * c2, -(K2 * u + nx * c2), -(K2 * v + ny * c2), -(K2 * w + nz * c2), K2},
                     {1 - K * q2, K * u, K * v, K * w, -K},
                     {K2 * q2 - qn * c2, -(K2 * u - nx * c2), -(K2 * v - ny * c2), -(K2 * w - nz * c2), K2},
                     {-ql, lx, ly, lz, 0},
                     {-qm, mx, my, mz, 0}};
        T R[5][5] = {{1, 1, 1, 0, 0},
                     {u - cc * nx, u, u + cc * nx, lx, mx},
                     {v - cc * ny, v, v + cc * ny, ly, my},
                     {w - cc * nz, w, w + cc * nz, lz, mz},
                     {H - qn * cc, q2, H + qn * cc, ql, qm}};
        T flux[5];
        if (sch == F3D_SCHEME_WENO7) {  // schemes.jl:813-930
          for (int n = 0; n < 5; ++n) {
            T Fpc[7] = {0, 0, 0, 0, 0, 0, 0}, Fmc[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int m = 0; m < 5; ++m) {
              T ll = L[n][m];
              const size_t mm = ncell * m;
              for (int s = 0; s < 7; ++s) {
                Fpc[s] += Fp[C(s - 4) + mm] * ll;  // Fp[i-4 .. i+2]
                Fmc[s] += Fm[C(3 - s) + mm] * ll;  // Fm[i+3 .. i-3]
              }
            }
            T fp = weno7(Fpc, s0, eps7);
            T fm = weno7(Fmc, s0, eps7);
            flux[n] = (fp + fm) * tmp1;
          }
        } else if (sch == F3D_SCHEME_WENO5) {  // schemes.jl:931-999
          for (int n = 0; n < 5; ++n) {
            T Fpc[5] = {0, 0, 0, 0, 0}, Fmc[5] = {0, 0, 0, 0, 0};
            for (int m = 0; m < 5; ++m) {
              T ll = L[n][m];
              const size_t mm = ncell * m;
              for (int s = 0; s < 5; ++s) {
                Fpc[s] += Fp[C(s - 3) + mm] * ll;  // Fp[i-3 .. i+1]
                Fmc[s] += Fm[C(2 - s) + mm] * ll;  // Fm[i+2 .. i-2]
              }
            }
            T fp = weno5(Fpc, s0, eps5);
            T fm = weno5(Fmc, s0, eps5);
            flux[n] = (fp + fm) * tmp2;
          }
        } else {  // NND2, schemes.jl:1000-1009 (component-wise; L and R are built but unused)
          for (long n = 1; n <= 5; ++n) {
            const size_t nn = ncell * (n - 1);
            T fp = Fp[C(-1) + nn] + T(0.5) * minmod(Fp[C(0) + nn] - Fp[C(-1) + nn], Fp[C(-1) + nn] - Fp[C(-2) + nn]);
            T fm = Fm[C(0) + nn] - T(0.5) * minmod(Fm[C(1) + nn] - Fm[C(0) + nn], Fm[C(0) + nn] - Fm[C(-1) + nn]);
            F[out(n)] = fp + fm;
          }
          continue;
        }
        for (int n = 0; n < 5; ++n)  // F = R * flux  schemes.jl:926-930
          F[out(n + 1)] = flux[0] * R[n][0] + flux[1] * R[n][1] + flux[2] * R[n][2] +
                          flux[3] * R[n][3] + flux[4] * R[n][4];
      }
}

// ---------------------------------------------------------------- split.jl
// fluxSplit_SW  split.jl:2-68, every cell [1, N_tot]^3
template <typename T>
void flux_split_sw(const f3d_params* p, const T* Q, T* Fp, T* Fm, const T* S, const T* Ax,
                   const T* Ay, const T* Az) {
  Grid g(p);
  const T gam = (T)p->gamma, Rg = (T)p->Rg;
#pragma omp parallel for collapse(2)
  for (long k = 1; k <= g.Nzt; ++k)
    for (long j = 1; j <= g.Nyt; ++j)
      for (long i = 1; i <= g.Nxt; ++i) {
        T rho = Q[g.cell.at(i, j, k, 1)], u = Q[g.cell.at(i, j, k, 2)];
        T v = Q[g.cell.at(i, j, k, 3)], w = Q[g.cell.at(i, j, k, 4)];
        T Tt = Q[g.cell.at(i, j, k, 6)];
        T A1 = Ax[g.cell.at(i, j, k)], A2 = Ay[g.cell.at(i, j, k)], A3 = Az[g.cell.at(i, j, k)];
        T ss = S[g.cell.at(i, j, k)];
        T c = std::sqrt(gam * Rg * Tt);
        T E1 = A1 * u + A2 * v + A3 * w;
        T E2 = E1 - c * ss;
        T E3 = E1 + c * ss;
        ss = 1 / ss;
        A1 *= ss; A2 *= ss; A3 *= ss;
        T E1P = (E1 + std::fabs(E1)) * T(0.5);
        T E2P = (E2 + std::fabs(E2)) * T(0.5);
        T E3P = (E3 + std::fabs(E3)) * T(0.5);
        T E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
        T uc1 = u - c * A1, uc2 = u + c * A1;
        T vc1 = v - c * A2, vc2 = v + c * A2;
        T wc1 = w - c * A3, wc2 = w + c * A3;
        T vvc1 = (uc1 * uc1 + vc1 * vc1 + wc1 * wc1) * T(0.5);
        T vvc2 = (uc2 * uc2 + vc2 * vc2 + wc2 * wc2) * T(0.5);
        T vv = (gam - 1) * (u * u + v * v + w * w);
        T W2 = (3 - gam) / (2 * (gam - 1)) * (c * c);
        T tmp1 = rho / (2 * gam);
        T tmp2 = 2 * (gam - 1);
        T* out[2] = {Fp, Fm};
        T e1[2] = {E1P, E1M}, e2[2] = {E2P, E2M}, e3[2] = {E3P, E3M};
        for (int s = 0; s < 2; ++s) {
          out[s][g.cell.at(i, j, k, 1)] = tmp1 * (tmp2 * e1[s] + e2[s] + e3[s]);
          out[s][g.cell.at(i, j, k, 2)] = tmp1 * (tmp2 * e1[s] * u + e2[s] * uc1 + e3[s] * uc2);
          out[s][g.cell.at(i, j, k, 3)] = tmp1 * (tmp2 * e1[s] * v + e2[s] * vc1 + e3[s] * vc2);
          out[s][g.cell.at(i, j, k, 4)] = tmp1 * (tmp2 * e1[s] * w + e2[s] * wc1 + e3[s] * wc2);
          out[s][g.cell.at(i, j, k, 5)] =
              tmp1 * (e1[s] * vv + e2[s] * vvc1 + e3[s] * vvc2 + W2 * (e2[s] + e3[s]));
        }
      }
}

// The three other splittings of split.jl, every cell [1, N_tot]^3.  Each reads rho, u, v, w and p (not T)
// and the metric row A with S = |A|; sound speed c = sqrt(gamma p / rho).
// fluxSplit_LF  split.jl:70-108: F± = (F(A) ± lambda0 U)/2, lambda0 = |A.u| + c S.
// fluxSplit_VL  split.jl:110-183: van Leer mass-flux splitting in M = A.u/(c S).
// fluxSplit_AUSM  split.jl:185-255: AUSM mass-flux and pressure splitting in M.
template <typename T>
void flux_split_other(const f3d_params* p, int method, const T* Q, T* Fp, T* Fm, const T* S, const T* Ax,
                      const T* Ay, const T* Az) {
  Grid g(p);
  const T gam = (T)p->gamma;
#pragma omp parallel for collapse(2)
  for (long k = 1; k <= g.Nzt; ++k)
    for (long j = 1; j <= g.Nyt; ++j)
      for (long i = 1; i <= g.Nxt; ++i) {
        const T rho = Q[g.cell.at(i, j, k, 1)], u = Q[g.cell.at(i, j, k, 2)];
        const T v = Q[g.cell.at(i, j, k, 3)], w = Q[g.cell.at(i, j, k, 4)];
        const T pr = Q[g.cell.at(i, j, k, 5)];
        const T A1 = Ax[g.cell.at(i, j, k)], A2 = Ay[g.cell.at(i, j, k)], A3 = Az[g.cell.at(i, j, k)];
        const T ss = S[g.cell.at(i, j, k)];
        const T c = std::sqrt(gam * pr / rho);
        T un = A1 * u + A2 * v + A3 * w;
        T fp[5], fm[5];
        if (method == F3D_SPLIT_LF) {  // split.jl:92-107
          const T l0 = std::fabs(un) + c * ss;
          const T E0 = pr / (gam - 1) + T(0.5) * rho * (u * u + v * v + w * w);
          fp[0] = T(0.5) * (rho * un + l0 * rho);
          fp[1] = T(0.5) * (rho * u * un + A1 * pr + l0 * rho * u);
          fp[2] = T(0.5) * (rho * v * un + A2 * pr + l0 * rho * v);
          fp[3] = T(0.5) * (rho * w * un + A3 * pr + l0 * rho * w);
          fp[4] = T(0.5) * ((E0 + pr) * un + l0 * E0);
          fm[0] = T(0.5) * (rho * un - l0 * rho);
          fm[1] = T(0.5) * (rho * u * un + A1 * pr - l0 * rho * u);
          fm[2] = T(0.5) * (rho * v * un + A2 * pr - l0 * rho * v);
          fm[3] = T(0.5) * (rho * w * un + A3 * pr - l0 * rho * w);
          fm[4] = T(0.5) * ((E0 + pr) * un - l0 * E0);
        } else {
          const T M = un / (c * ss);
          // full flux on the upwind side when |M| >= 1 (split.jl:136-159, 211-234)
          const T E0 = pr / (gam - 1) + T(0.5) * rho * (u * u + v * v + w * w);
          const T h = gam * pr / (rho * (gam - 1)) + T(0.5) * (u * u + v * v + w * w);
          const T full[5] = {rho * un, rho * u * un + A1 * pr, rho * v * un + A2 * pr, rho * w * un + A3 * pr,
                             method == F3D_SPLIT_VL ? (E0 + pr) * un : rho * h * un};
          if (M <= -1) {
            for (int q = 0; q < 5; ++q) {
              fp[q] = 0;
              fm[q] = full[q];
            }
          } else if (M >= 1) {
            for (int q = 0; q < 5; ++q) {
              fm[q] = 0;
              fp[q] = full[q];
            }
          } else if (method == F3D_SPLIT_VL) {  // split.jl:161-181
            un = un / ss;
            const T f1p = rho * c * T(0.25) * ((M + 1) * (M + 1)) * ss;
            const T f1m = -rho * c * T(0.25) * ((M - 1) * (M - 1)) * ss;
            const T t1 = (-un + 2 * c) / (gam * ss);
            const T t2 = (-un - 2 * c) / (gam * ss);
            const T t3 = 1 / (2 * (gam * gam - 1));
            const T t4 = T(0.5) * (u * u + v * v + w * w - un * un);
            fp[0] = f1p;
            fp[1] = f1p * (A1 * t1 + u);
            fp[2] = f1p * (A2 * t1 + v);
            fp[3] = f1p * (A3 * t1 + w);
            fp[4] = f1p * (((gam - 1) * un + 2 * c) * ((gam - 1) * un + 2 * c) * t3 + t4);
            fm[0] = f1m;
            fm[1] = f1m * (A1 * t2 + u);
            fm[2] = f1m * (A2 * t2 + v);
            fm[3] = f1m * (A3 * t2 + w);
            fm[4] = f1m * (((gam - 1) * un - 2 * c) * ((gam - 1) * un - 2 * c) * t3 + t4);
          } else {  // AUSM, split.jl:236-253
            const T f1p = T(0.25) * ((M + 1) * (M + 1)) * rho * c * ss;
            const T f1m = -T(0.25) * ((M - 1) * (M - 1)) * rho * c * ss;
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
        }
        for (int q = 0; q < 5; ++q) {
          Fp[g.cell.at(i, j, k, q + 1)] = fp[q];
          Fm[g.cell.at(i, j, k, q + 1)] = fm[q];
        }
      }
}

// advect_x / advect_y / advect_z  schemes.jl:86-290 / 291-495 / 496-698: the component-wise form of
// advect_*c (the same sensor, UP7, WENO7 with the Float64 epsilon, WENO5 and NND2 stencils on each of
// the first NV split-flux components, no characteristic projection).
template <typename T>
void advect_comp(const f3d_params* p, int d, T* F, const T* phi, const T* S, const T* Fp, const T* Fm, int NV) {
  Grid g(p);
  Dir D = make_dir(d);
  const Ext& fe = d == 0 ? g.fx : (d == 1 ? g.fy : g.fz);
  T c[7];
  linear_coeffs<T>(p, c);
  const T tmp1 = T(1) / T(12), tmp2 = T(1) / T(6);
  const T eps7 = (T)p->weno7_eps, eps5 = (T)p->weno5_eps;
  const size_t ncell = (size_t)g.Nxt * g.Nyt * g.Nzt;
#pragma omp parallel for collapse(2)
  for (long k = g.NG + 1; k <= g.Nzp + g.NG + D.dk; ++k)
    for (long j = g.NG + 1; j <= g.Nyp + g.NG + D.dj; ++j)
      for (long i = g.NG + 1; i <= g.Nxp + g.NG + D.di; ++i) {
        auto C = [&](long o) { return g.cell.at(i + o * D.di, j + o * D.dj, k + o * D.dk); };
        const int sch = face_scheme(p, g, phi, i, j, k, D, (T*)nullptr);
        const T ss = 2 / (S[C(-1)] + S[C(0)]);  // schemes.jl:109
        for (long n = 1; n <= NV; ++n) {
          const size_t nn = ncell * (n - 1);
          T f;
          if (sch == F3D_SCHEME_UP7) {  // schemes.jl:119-143
            T fp = 0, fm = 0;
            for (int q = 0; q < 7; ++q) fp += c[q] * Fp[C(q - 4) + nn];
            for (int q = 0; q < 7; ++q) fm += c[q] * Fm[C(3 - q) + nn];
            f = fp + fm;
          } else if (sch == F3D_SCHEME_WENO7) {  // schemes.jl:144-224
            T Vp[7], Vm[7];
            for (int q = 0; q < 7; ++q) {
              Vp[q] = Fp[C(q - 4) + nn];
              Vm[q] = Fm[C(3 - q) + nn];
            }
            f = (weno7(Vp, ss, eps7) + weno7(Vm, ss, eps7)) * tmp1;
          } else if (sch == F3D_SCHEME_WENO5) {  // schemes.jl:225-276
            T Vp[5], Vm[5];
            for (int q = 0; q < 5; ++q) {
              Vp[q] = Fp[C(q - 3) + nn];
              Vm[q] = Fm[C(2 - q) + nn];
            }
            f = (weno5(Vp, ss, eps5) + weno5(Vm, ss, eps5)) * tmp2;
          } else {  // schemes.jl:277-286
            const T fp = Fp[C(-1) + nn] + T(0.5) * minmod(Fp[C(0) + nn] - Fp[C(-1) + nn], Fp[C(-1) + nn] - Fp[C(-2) + nn]);
            const T fm = Fm[C(0) + nn] - T(0.5) * minmod(Fm[C(1) + nn] - Fm[C(0) + nn], Fm[C(0) + nn] - Fm[C(-1) + nn]);
            f = fp + fm;
          }
          F[fe.at(i - g.NG, j - g.NG, k - g.NG, n)] = f;
        }
      }
}

// ---------------------------------------------------------------- viscous.jl
// Eddy viscosity of the LES branches  viscous.jl:66-106, added to mu and lam (lam_t = Cp mu_t / Prt,
// Prt = 0.9), mu_t capped at twice the laminar mu; filter width Delta = J^(-1/3).
//   Smagorinsky (66-79): mu_t = rho (0.1 Delta)^2 sqrt(2 S_ij S_ij).
//   WALE (80-106): mu_t = rho (0.325 Delta)^2 Sd^3 / (S^5 + Sd^(5/2)), S = sqrt(S_ij S_ij),
//   Sd = sqrt(Sd_ij Sd_ij) of the traceless symmetric part of g^2 (g_ij = du_i/dx_j), in the
//   reference's expanded products.  Where S = Sd = 0 the reference's 0/0 is replaced by D = 0 (the
//   model's value without a velocity gradient).  The float form evaluates D from the gradient scaled by
//   its largest component (S^5 overflows float beyond gradients of ~5e7 1/s); FP64 keeps the literal form.
// g[3][3]: rows u, v, w; columns x, y, z.
template <typename T>
void les_eddy(const f3d_params* p, T rho, T Jac, const T g[3][3], T& mu, T& lam) {
  if (p->les == F3D_LES_NONE) return;
  const T Prt = T(0.9);
  const T half = T(0.5);
  const T dudx = g[0][0], dudy = g[0][1], dudz = g[0][2];
  const T dvdx = g[1][0], dvdy = g[1][1], dvdz = g[1][2];
  const T dwdx = g[2][0], dwdy = g[2][1], dwdz = g[2][2];
  const T s12 = half * (dudy + dvdx), s13 = half * (dudz + dwdx), s23 = half * (dvdz + dwdy);
  const T SS = dudx * dudx + dvdy * dvdy + dwdz * dwdz + 2 * (s12 * s12 + s13 * s13 + s23 * s23);
  const T cb = std::cbrt(Jac);
  T mut;
  if (p->les == F3D_LES_SMAG) {
    const T Sijmag = std::sqrt(2 * SS);
    const T dl = T(0.1) / cb;
    mut = jmin(rho * (dl * dl) * Sijmag, 2 * mu);
  } else {
    T sc = T(1), G[3][3];
    if (sizeof(T) == 4) {
      T gm = T(0);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) gm = std::max(gm, (T)std::fabs(g[a][b]));
      sc = gm > T(0) ? gm : T(1);
    }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) G[a][b] = g[a][b] / sc;
    const T ux = G[0][0], uy = G[0][1], uz = G[0][2], vx = G[1][0], vy = G[1][1], vz = G[1][2];
    const T wx = G[2][0], wy = G[2][1], wz = G[2][2];
    const T t12 = half * (uy + vx), t13 = half * (uz + wx), t23 = half * (vz + wy);
    const T S = std::sqrt(ux * ux + vy * vy + wz * wz + 2 * (t12 * t12 + t13 * t13 + t23 * t23));
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
    const T Sd = std::sqrt(Sd11 * Sd11 + Sd22 * Sd22 + Sd33 * Sd33 + 2 * (Sd12 * Sd12 + Sd13 * Sd13 + Sd23 * Sd23));
    const T den = S * S * S * S * S + Sd * Sd * std::sqrt(Sd);
    const T D = den > T(0) ? sc * (Sd * Sd * Sd / den) : T(0);
    const T dl = T(0.325) / cb;
    mut = jmin(rho * (dl * dl) * D, 2 * mu);
  }
  mu += mut;
  lam += (T)p->Cp * mut / Prt;
}

// viscousFlux  viscous.jl:1-134 (laminar + the LES eddy viscosity selected by p->les), domain
// [3, N_tot-2]^3, stored at (i-2, j-2, k-2).  J is double in both forms (cast to T here).
template <typename T>
void viscous_flux(const f3d_params* p, T* Fvx, T* Fvy, T* Fvz, const T* Q, const T* const m[9],
                  const double* J) {
  Grid g(p);
  const T Cs = (T)p->C_s, Ts = (T)p->T_s, Cp = (T)p->Cp, Pr = (T)p->Pr;
  const T c1 = T(1) / T(12), c2 = T(-2) / T(3), c23 = T(2) / T(3);
  const size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
#pragma omp parallel for collapse(2)
  for (long k = 3; k <= g.Nzt - 2; ++k)
    for (long j = 3; j <= g.Nyt - 2; ++j)
      for (long i = 3; i <= g.Nxt - 2; ++i) {
        const size_t o = g.cell.at(i, j, k);
        T xx = m[0][o], xy = m[1][o], xz = m[2][o];
        T ex = m[3][o], ey = m[4][o], ez = m[5][o];
        T zx = m[6][o], zy = m[7][o], zz = m[8][o];
        T Jac = (T)J[o];
        T Tt = Q[o + 5 * nc];
        T mu = Cs * Tt * std::sqrt(Tt) / (Tt + Ts);
        T lam = Cp * mu / Pr;
        auto D = [&](long di, long dj, long dk, int n) {
          const size_t nn = nc * (n - 1);
          return c1 * (Q[g.cell.at(i - 2 * di, j - 2 * dj, k - 2 * dk) + nn] -
                       Q[g.cell.at(i + 2 * di, j + 2 * dj, k + 2 * dk) + nn]) +
                 c2 * (Q[g.cell.at(i - di, j - dj, k - dk) + nn] -
                       Q[g.cell.at(i + di, j + dj, k + dk) + nn]);
        };
        T uxi = D(1, 0, 0, 2), vxi = D(1, 0, 0, 3), wxi = D(1, 0, 0, 4), Txi = D(1, 0, 0, 6);
        T ueta = D(0, 1, 0, 2), veta = D(0, 1, 0, 3), weta = D(0, 1, 0, 4), Teta = D(0, 1, 0, 6);
        T uzeta = D(0, 0, 1, 2), vzeta = D(0, 0, 1, 3), wzeta = D(0, 0, 1, 4), Tzeta = D(0, 0, 1, 6);
        T u = Q[o + nc], v = Q[o + 2 * nc], w = Q[o + 3 * nc];
        T dudx = (uxi * xx + ueta * ex + uzeta * zx) * Jac;
        T dudy = (uxi * xy + ueta * ey + uzeta * zy) * Jac;
        T dudz = (uxi * xz + ueta * ez + uzeta * zz) * Jac;
        T dvdx = (vxi * xx + veta * ex + vzeta * zx) * Jac;
        T dvdy = (vxi * xy + veta * ey + vzeta * zy) * Jac;
        T dvdz = (vxi * xz + veta * ez + vzeta * zz) * Jac;
        T dwdx = (wxi * xx + weta * ex + wzeta * zx) * Jac;
        T dwdy = (wxi * xy + weta * ey + wzeta * zy) * Jac;
        T dwdz = (wxi * xz + weta * ez + wzeta * zz) * Jac;
        T dTdx = (Txi * xx + Teta * ex + Tzeta * zx) * Jac;
        T dTdy = (Txi * xy + Teta * ey + Tzeta * zy) * Jac;
        T dTdz = (Txi * xz + Teta * ez + Tzeta * zz) * Jac;
        T dv = dudx + dvdy + dwdz;
        if constexpr (std::is_floating_point<T>::value) {  // (the dual-number forms are laminar only)
          if (p->les != F3D_LES_NONE) {
            const T gr[3][3] = {{dudx, dudy, dudz}, {dvdx, dvdy, dvdz}, {dwdx, dwdy, dwdz}};
            les_eddy<T>(p, Q[o], Jac, gr, mu, lam);
          }
        }
        T t11 = mu * (2 * dudx - c23 * dv);
        T t12 = mu * (dudy + dvdx);
        T t13 = mu * (dudz + dwdx);
        T t22 = mu * (2 * dvdy - c23 * dv);
        T t23 = mu * (dwdy + dvdz);
        T t33 = mu * (2 * dwdz - c23 * dv);
        T E1 = u * t11 + v * t12 + w * t13 + lam * dTdx;
        T E2 = u * t12 + v * t22 + w * t23 + lam * dTdy;
        T E3 = u * t13 + v * t23 + w * t33 + lam * dTdz;
        T* Fo[3] = {Fvx, Fvy, Fvz};
        T mr[3][3] = {{xx, xy, xz}, {ex, ey, ez}, {zx, zy, zz}};
        for (int d = 0; d < 3; ++d) {
          Fo[d][g.fv.at(i - 2, j - 2, k - 2, 1)] = mr[d][0] * t11 + mr[d][1] * t12 + mr[d][2] * t13;
          Fo[d][g.fv.at(i - 2, j - 2, k - 2, 2)] = mr[d][0] * t12 + mr[d][1] * t22 + mr[d][2] * t23;
          Fo[d][g.fv.at(i - 2, j - 2, k - 2, 3)] = mr[d][0] * t13 + mr[d][1] * t23 + mr[d][2] * t33;
          Fo[d][g.fv.at(i - 2, j - 2, k - 2, 4)] = mr[d][0] * E1 + mr[d][1] * E2 + mr[d][2] * E3;
        }
      }
}

// ---------------------------------------------------------------- div.jl
// div  div.jl:1-50 : interior (i,j,k) in [1, Np]^3, U index +NG.  The flux differences accumulate
// in T; the conservative update is in double.
// ldt != nullptr: div_LTS  div.jl:52-101, the same update with the per-cell dt ldt[cell] (interior cells
// of a cell-shaped array) in place of the scalar dt.
template <typename T>
void div(const f3d_params* p, double* U, const T* Fx, const T* Fy, const T* Fz, const T* Fvx,
         const T* Fvy, const T* Fvz, double dt, const double* J, const double* ldt = nullptr) {
  Grid g(p);
  const T c1 = T(1) / T(12), c2 = T(-2) / T(3);
  const long NG = g.NG;
#pragma omp parallel for collapse(2)
  for (long k = 1; k <= g.Nzp; ++k)
    for (long j = 1; j <= g.Nyp; ++j)
      for (long i = 1; i <= g.Nxp; ++i) {
        const size_t oc = g.cell.at(i + NG, j + NG, k + NG);
        double Jact = J[oc] * (ldt ? ldt[oc] : dt);
        T dV[3][4];
        for (int n = 1; n <= 4; ++n) {
          dV[0][n - 1] = c1 * (Fvx[g.fv.at(i, j + 2, k + 2, n)] - Fvx[g.fv.at(i + 4, j + 2, k + 2, n)]) +
                         c2 * (Fvx[g.fv.at(i + 1, j + 2, k + 2, n)] - Fvx[g.fv.at(i + 3, j + 2, k + 2, n)]);
          dV[1][n - 1] = c1 * (Fvy[g.fv.at(i + 2, j, k + 2, n)] - Fvy[g.fv.at(i + 2, j + 4, k + 2, n)]) +
                         c2 * (Fvy[g.fv.at(i + 2, j + 1, k + 2, n)] - Fvy[g.fv.at(i + 2, j + 3, k + 2, n)]);
          dV[2][n - 1] = c1 * (Fvz[g.fv.at(i + 2, j + 2, k, n)] - Fvz[g.fv.at(i + 2, j + 2, k + 4, n)]) +
                         c2 * (Fvz[g.fv.at(i + 2, j + 2, k + 1, n)] - Fvz[g.fv.at(i + 2, j + 2, k + 3, n)]);
        }
        for (long n = 1; n <= 5; ++n) {
          T s = Fx[g.fx.at(i, j, k, n)] - Fx[g.fx.at(i + 1, j, k, n)] + Fy[g.fy.at(i, j, k, n)] -
                Fy[g.fy.at(i, j + 1, k, n)] + Fz[g.fz.at(i, j, k, n)] - Fz[g.fz.at(i, j, k + 1, n)];
          U[g.cell.at(i + NG, j + NG, k + NG, n)] += (double)s * Jact;
        }
        for (int n = 2; n <= 5; ++n) {
          T s = dV[0][n - 2] + dV[1][n - 2] + dV[2][n - 2];
          U[g.cell.at(i + NG, j + NG, k + NG, n)] += (double)s * Jact;
        }
      }
}

// ---------------------------------------------------------------- filters  utils.jl:102-316
// pre_x/_y/_z  utils.jl:102-169: shock sensor of the nonlinear filter along d from the pressure
// (Q index 5, 1-based) at d-offsets -2..2:
//   dp0 = (-p4 + 2p3 - p2)/4, dp1 = (-p5 + 2p4 - p3)/4, dp2 = (-p3 + 2p2 - p1)/4,
//   r = ((dp0 - dp1)^2 + (dp0 - dp2)^2)/2 / p3^2 + 1e-16,  sc = (1 - rth/r + |1 - rth/r|)/2.
// sc is cell-shaped here and written on the interior and on the first ghost layer of each side along
// d, which the filter's boundary faces read (the reference allocates sc as (Nxp, Nyp, Nzp) but indexes
// it with the interior's ghost-offset range, so its boundary neighbours alias other cells or lie
// outside the allocation).  Double arithmetic from the bulk-type pressure.
template <typename T>
void pre_sensor(const f3d_params* p, int d, double* sc, const T* Q, double rth) {
  Grid g(p);
  const size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  const Dir D = make_dir(d);
  const long lo[3] = {g.NG + 1 - D.di, g.NG + 1 - D.dj, g.NG + 1 - D.dk};
  const long hi[3] = {g.Nxp + g.NG + D.di, g.Nyp + g.NG + D.dj, g.Nzp + g.NG + D.dk};
  for (long k = lo[2]; k <= hi[2]; ++k)
    for (long j = lo[1]; j <= hi[1]; ++j)
      for (long i = lo[0]; i <= hi[0]; ++i) {
        auto P = [&](long s) { return (double)Q[g.cell.at(i + s * D.di, j + s * D.dj, k + s * D.dk) + 4 * nc]; };
        const double p1 = P(-2), p2 = P(-1), p3 = P(0), p4 = P(1), p5 = P(2);
        const double dp0 = 0.25 * (-p4 + 2 * p3 - p2);
        const double dp1 = 0.25 * (-p5 + 2 * p4 - p3);
        const double dp2 = 0.25 * (-p3 + 2 * p2 - p1);
        const double ri = 0.5 * ((dp0 - dp1) * (dp0 - dp1) + (dp0 - dp2) * (dp0 - dp2)) / (p3 * p3) + 1e-16;
        sc[g.cell.at(i, j, k)] = 0.5 * (1.0 - rth / ri + std::fabs(1.0 - rth / ri));
      }
}

// filter_x/_y/_z  utils.jl:171-241: conservative nonlinear (shock-capturing) filter of the interior,
//   U = Un - s0 (sc+ D+ - sc- D-),  sc+- = (sc_i + sc_(i+-1))/2,
//   D+ = c1 (Un_(i+1) - Un_i) + c2 (Un_(i+2) - Un_(i-1)),  D- = c1 (Un_i - Un_(i-1)) + c2 (Un_(i+1) - Un_(i-2)),
// c1 = -0.210383, c2 = 0.039617, for the 5 conserved variables.
inline void filter_nonlinear(const f3d_params* p, int d, double* U, const double* Un, const double* sc, double s0) {
  Grid g(p);
  const Dir D = make_dir(d);
  const double c1 = -0.210383, c2 = 0.039617;
  for (long k = g.NG + 1; k <= g.Nzp + g.NG; ++k)
    for (long j = g.NG + 1; j <= g.Nyp + g.NG; ++j)
      for (long i = g.NG + 1; i <= g.Nxp + g.NG; ++i) {
        auto C = [&](long s) { return g.cell.at(i + s * D.di, j + s * D.dj, k + s * D.dk); };
        const double sc1 = 0.5 * (sc[C(0)] + sc[C(1)]);
        const double sc2 = 0.5 * (sc[C(0)] + sc[C(-1)]);
        for (long n = 1; n <= 5; ++n) {
          const size_t nn = (size_t)(n - 1) * g.cell.n1 * g.cell.n2 * g.cell.n3;
          auto V = [&](long s) { return Un[C(s) + nn]; };
          U[C(0) + nn] = V(0) - s0 * (sc1 * (c1 * (V(1) - V(0)) + c2 * (V(2) - V(-1))) -
                                      sc2 * (c1 * (V(0) - V(-1)) + c2 * (V(1) - V(-2))));
        }
      }
}

// linearFilter_x/_y/_z  utils.jl:243-316: 9-point linear filter of the interior,
//   U = Un - s0 (d0 Un_i + sum_(m=1..4) d_m (Un_(i-m) + Un_(i+m))),
// d = (0.24352749312, -0.20478888064, 0.12000759168, -0.04521111936, 0.00822866176).
inline void filter_linear(const f3d_params* p, int d, double* U, const double* Un, double s0) {
  Grid g(p);
  const Dir D = make_dir(d);
  const double dc[5] = {0.243527493120, -0.204788880640, 0.120007591680, -0.045211119360, 0.008228661760};
  for (long k = g.NG + 1; k <= g.Nzp + g.NG; ++k)
    for (long j = g.NG + 1; j <= g.Nyp + g.NG; ++j)
      for (long i = g.NG + 1; i <= g.Nxp + g.NG; ++i) {
        auto C = [&](long s) { return g.cell.at(i + s * D.di, j + s * D.dj, k + s * D.dk); };
        for (long n = 1; n <= 5; ++n) {
          const size_t nn = (size_t)(n - 1) * g.cell.n1 * g.cell.n2 * g.cell.n3;
          auto V = [&](long s) { return Un[C(s) + nn]; };
          const double f = dc[0] * V(0) + dc[1] * (V(-1) + V(1)) + dc[2] * (V(-2) + V(2)) +
                           dc[3] * (V(-3) + V(3)) + dc[4] * (V(-4) + V(4));
          U[C(0) + nn] = V(0) - s0 * f;
        }
      }
}

// The filtering block of the time loop  solver.jl:446-470: per direction x, y, z, Un <- U, then the
// nonlinear filter with its sensor from Q (filtering_nonlinear) or the linear filter.  Q is not
// updated afterwards (as in the reference: the next step starts from the filtered U and the
// unfiltered Q).
template <typename T>
void filter_block(const f3d_params* p, double* U, double* Un, const T* Q, double* sc) {
  Grid g(p);
  const size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  for (int d = 0; d < 3; ++d) {
    std::memcpy(Un, U, nc * 5 * sizeof(double));
    if (p->filtering_nonlinear) {
      pre_sensor<T>(p, d, sc, Q, p->filtering_rth);
      filter_nonlinear(p, d, U, Un, sc, p->filtering_s0);
    } else {
      filter_linear(p, d, U, Un, p->filtering_s0);
    }
  }
}

// ---------------------------------------------------------------- mpi.jl / boundary.jl
// exchange_ghost with a single rank  mpi.jl:2-103 (+ pack/unpack 105-283): for each periodic
// dimension in the order x, y, z: low ghosts <- Q[Np+g], high ghosts <- Q[NG+g], over the full
// extent (ghosts included) of the other two dimensions.
template <typename T>
void exchange_ghost(const f3d_params* p, T* Q, int NV) {
  Grid g(p);
  const long Np[3] = {g.Nxp, g.Nyp, g.Nzp};
  for (int d = 0; d < 3; ++d) {
    if (p->bc[2 * d] != F3D_BC_PERIODIC || p->bc[2 * d + 1] != F3D_BC_PERIODIC) continue;
    Dir D = make_dir(d);
    for (long n = 1; n <= NV; ++n)
#pragma omp parallel for collapse(2)
      for (long k = 1; k <= g.Nzt; ++k)
        for (long j = 1; j <= g.Nyt; ++j)
          for (long i = 1; i <= g.Nxt; ++i) {
            long idx = d == 0 ? i : (d == 1 ? j : k);
            long src;
            if (idx <= g.NG) src = idx + Np[d];
            else if (idx > Np[d] + g.NG) src = idx - Np[d];
            else continue;
            long si = i + (src - idx) * D.di, sj = j + (src - idx) * D.dj, sk = k + (src - idx) * D.dk;
            Q[g.cell.at(i, j, k, n)] = Q[g.cell.at(si, sj, sk, n)];
          }
  }
}

// Ghost fill  boundary.jl:1-156 generalised to per-face kinds; applied in the order fill_x,
// fill_y, fill_z (boundary.jl:153-155), each over the full extent of the other two dimensions.
// P is the type the ghost values are formed in (double, or Dual for the forward-mode reference);
// qinf / Tw override p->q_inf / p->T_wall when given.
template <typename T, typename P = double>
void fill_ghost(const f3d_params* p, T* Q, double* U, const T* inflow, const P* qinf = nullptr,
                const P* Tw = nullptr) {
  Grid g(p);
  const long Np[3] = {g.Nxp, g.Nyp, g.Nzp};
  const double gam = p->gamma, Rg = p->Rg;
  const size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  for (int d = 0; d < 3; ++d) {
    for (int side = 0; side < 2; ++side) {
      const int bc = p->bc[2 * d + side];
#pragma omp parallel for collapse(2)
      for (long k = 1; k <= g.Nzt; ++k)
        for (long j = 1; j <= g.Nyt; ++j)
          for (long i = 1; i <= g.Nxt; ++i) {
            const long idx = d == 0 ? i : (d == 1 ? j : k);
            const bool ghost = side == 0 ? (idx <= g.NG) : (idx > Np[d] + g.NG);
            if (!ghost) continue;
            auto shifted = [&](long to) {
              return g.cell.at(d == 0 ? to : i, d == 1 ? to : j, d == 2 ? to : k);
            };
            const size_t o = g.cell.at(i, j, k);
            P q[6];
            if (bc == F3D_BC_PERIODIC) {  // fill_z  boundary.jl:132-146
              const size_t s = shifted(side == 0 ? idx + Np[d] : idx - Np[d]);
              for (int n = 0; n < 6; ++n) Q[o + n * nc] = Q[s + n * nc];
              if (U)
                for (int n = 0; n < 5; ++n) U[o + n * nc] = U[s + n * nc];
              continue;
            } else if (bc == F3D_BC_ZEROGRAD) {  // boundary.jl:42-51
              const size_t s = shifted(side == 0 ? g.NG + 1 : Np[d] + g.NG);
              for (int n = 0; n < 6; ++n) Q[o + n * nc] = Q[s + n * nc];
              if (U)
                for (int n = 0; n < 5; ++n) U[o + n * nc] = U[s + n * nc];
              continue;
            } else if (bc == F3D_BC_DIRICHLET) {  // boundary.jl:12-38, 98-118
              for (int n = 0; n < 6; ++n) q[n] = qinf ? qinf[n] : P(p->q_inf[n]);
            } else if (bc == F3D_BC_INFLOW) {
              for (int n = 0; n < 6; ++n)
                q[n] = (P)inflow[(size_t)(j - 1) + (size_t)g.Nyt * ((size_t)(k - 1) + (size_t)g.Nzt * n)];
            } else {  // walls: mirror index 2NG+1-idx (boundary.jl:69), or its high-side image
              const long mi = side == 0 ? 2 * g.NG + 1 - idx : 2 * (Np[d] + g.NG) + 1 - idx;
              const size_t s = shifted(mi);
              q[0] = (P)Q[s];
              q[1] = -(P)Q[s + nc];
              q[2] = -(P)Q[s + 2 * nc];
              q[3] = -(P)Q[s + 3 * nc];
              q[4] = (P)Q[s + 4 * nc];
              q[5] = (P)Q[s + 5 * nc];
              if (bc == F3D_BC_WALL_ISOTHERMAL) {
                q[5] = 2 * (Tw ? *Tw : P(p->T_wall)) - q[5];
                q[0] = q[4] / (Rg * q[5]);
              }
            }
            for (int n = 0; n < 6; ++n) Q[o + n * nc] = (T)q[n];
            if (U) {
              // U from the stored (rounded) Q, as boundary.jl:82-92 reads Q back
              double rho = (double)Q[o], u = (double)Q[o + nc], v = (double)Q[o + 2 * nc], w = (double)Q[o + 3 * nc];
              double pr = (double)Q[o + 4 * nc];
              U[o] = rho;
              U[o + nc] = rho * u;
              U[o + 2 * nc] = rho * v;
              U[o + 3 * nc] = rho * w;
              U[o + 4 * nc] = pr / (gam - 1) + 0.5 * rho * (u * u + v * v + w * w);
            }
          }
    }
  }
}

// ---------------------------------------------------------------- solver.jl
// flowAdvance, finite-difference branch  solver.jl:133-188 (SW splitting, characteristic advect)
// (solver.jl:134-184: per direction the split selected by splitMethod, then advect_*c when
// `character`, else the component-wise advect_*)
template <typename T>
void flow_advance(const f3d_params* p, T* Q, T* Fp, T* Fm, T* Fx, T* Fy, T* Fz, T* Fvx, T* Fvy,
                  T* Fvz, const T* s1, const T* s2, const T* s3, const T* const m[9],
                  const double* J, T* phi) {
  const T* S[3] = {s1, s2, s3};
  T* F[3] = {Fx, Fy, Fz};
  for (int d = 0; d < 3; ++d) {
    if (p->split == F3D_SPLIT_SW)
      flux_split_sw<T>(p, Q, Fp, Fm, S[d], m[3 * d], m[3 * d + 1], m[3 * d + 2]);
    else
      flux_split_other<T>(p, p->split, Q, Fp, Fm, S[d], m[3 * d], m[3 * d + 1], m[3 * d + 2]);
    if (p->character)
      advect_c<T>(p, d, F[d], phi, S[d], Fp, Fm, Q, m[3 * d], m[3 * d + 1], m[3 * d + 2]);
    else
      advect_comp<T>(p, d, F[d], phi, S[d], Fp, Fm, 5);
  }
  if (p->viscous) viscous_flux<T>(p, Fvx, Fvy, Fvz, Q, m, J);
}

// Scratch arrays of the reference's time_step (solver.jl:294-308), sized for one params; ldt is the
// per-cell dt of local time stepping (LTS_dt, solver.jl:304-306), double here.
template <typename T>
struct Scratch {
  std::vector<T> phi, Fp, Fm, Fx, Fy, Fz, Fvx, Fvy, Fvz;
  std::vector<double> ldt, fsc;  // LTS dt; the nonlinear filter's sensor (solver.jl:314-316)
  explicit Scratch(const f3d_params* p) {
    Grid g(p);
    size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
    phi.assign(nc, 0);
    if (p->lts) ldt.assign(nc, 0.0);
    if (p->filtering) fsc.assign(nc, 0.0);
    Fp.assign(nc * 5, 0);
    Fm.assign(nc * 5, 0);
    Fx.assign((size_t)(g.Nxp + 1) * g.Nyp * g.Nzp * 5, 0);
    Fy.assign((size_t)g.Nxp * (g.Nyp + 1) * g.Nzp * 5, 0);
    Fz.assign((size_t)g.Nxp * g.Nyp * (g.Nzp + 1) * 5, 0);
    size_t nv = (size_t)(g.Nxp + g.NG) * (g.Nyp + g.NG) * (g.Nzp + g.NG) * 4;
    Fvx.assign(nv, 0);
    Fvy.assign(nv, 0);
    Fvz.assign(nv, 0);
  }
};

// |grad xi_d| per cell  solver.jl:286-288
template <typename T>
void metric_norms(const f3d_params* p, const T* const m[9], T* s1, T* s2, T* s3) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  T* s[3] = {s1, s2, s3};
  for (int d = 0; d < 3; ++d)
    for (size_t o = 0; o < nc; ++o) {
      T a = m[3 * d][o], b = m[3 * d + 1][o], c = m[3 * d + 2][o];
      s[d][o] = std::sqrt(a * a + b * b + c * c);
    }
}

// compute_dt  utils.jl:73-100: the local time step of every interior cell,
//   dt = LTS_CFL * min(dx/(u+c), dy/(v+c), dz/(w+c)),  dx = (1/J)/S1, dy = (1/J)/S2, dz = (1/J)/S3,
// with the Cartesian velocity components u, v, w against the xi/eta/zeta sizes and u+c (not |u|+c)
// as written (a cell whose u < -c gets a negative dt there).  In double from the bulk-type inputs;
// Julia's NaN-propagating min.
template <typename T>
void compute_dt(const f3d_params* p, double* dt, const T* Q, const double* J, const T* S1, const T* S2,
                const T* S3) {
  Grid g(p);
  const size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  for (long k = g.NG + 1; k <= g.Nzp + g.NG; ++k)
    for (long j = g.NG + 1; j <= g.Nyp + g.NG; ++j)
      for (long i = g.NG + 1; i <= g.Nxp + g.NG; ++i) {
        const size_t o = g.cell.at(i, j, k);
        const double u = Q[o + nc], v = Q[o + 2 * nc], w = Q[o + 3 * nc], Tt = Q[o + 5 * nc];
        const double V = 1.0 / J[o];
        const double c = std::sqrt(p->gamma * p->Rg * Tt);
        const double dx = V / (double)S1[o], dy = V / (double)S2[o], dz = V / (double)S3[o];
        const double dtx = dx / (u + c), dty = dy / (v + c), dtz = dz / (w + c);
        dt[o] = jmin(jmin(dtx, dty), dtz) * p->lts_cfl;
      }
}

// One RK3 stage of the time loop body  solver.jl:416-444 (filtering and I/O off).  Local time
// stepping (p->lts): compute_dt from the stage-1 state at KRK == 1 (solver.jl:419-421), div_LTS
// with it in every stage (426-428); the scalar dt is then unused.
template <typename T>
void rk3_stage(const f3d_params* p, Scratch<T>& sc, int KRK, double* U, double* Un, T* Q,
               const T* s1, const T* s2, const T* s3, const T* const m[9], const double* J,
               double dt, const T* inflow) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  if (KRK == 1) std::memcpy(Un, U, nc * 5 * sizeof(double));  // solver.jl:418
  if (KRK == 1 && p->lts) compute_dt<T>(p, sc.ldt.data(), Q, J, s1, s2, s3);  // solver.jl:419-421
  shock_sensor<T>(p, sc.phi.data(), Q);                        // solver.jl:424
  flow_advance<T>(p, Q, sc.Fp.data(), sc.Fm.data(), sc.Fx.data(), sc.Fy.data(), sc.Fz.data(),
                  sc.Fvx.data(), sc.Fvy.data(), sc.Fvz.data(), s1, s2, s3, m, J, sc.phi.data());
  div<T>(p, U, sc.Fx.data(), sc.Fy.data(), sc.Fz.data(), sc.Fvx.data(), sc.Fvy.data(),
         sc.Fvz.data(), dt, J, p->lts ? sc.ldt.data() : nullptr);  // solver.jl:426-430
  if (KRK == 2) lin_comb(p, U, Un, 5, 0.25, 0.75);             // solver.jl:433
  if (KRK == 3) lin_comb(p, U, Un, 5, 2.0 / 3.0, 1.0 / 3.0);   // solver.jl:435
  c2prim<T>(p, U, Q);                                          // solver.jl:438
  exchange_ghost<T>(p, Q, 6);                                  // solver.jl:439
  fill_ghost<T>(p, Q, U, inflow);                              // solver.jl:443
}

// Global CFL dt (see f3d_cfl_dt_*): cfl * min over interior of 1/(J * max_d(|A_d.u| + c|A_d|)).
template <typename T>
double cfl_dt(const f3d_params* p, const T* Q, const T* const m[9], const double* J, double cfl) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  double best = std::numeric_limits<double>::infinity();
#pragma omp parallel for collapse(2) reduction(min : best)
  for (long k = g.NG + 1; k <= g.Nzp + g.NG; ++k)
    for (long j = g.NG + 1; j <= g.Nyp + g.NG; ++j)
      for (long i = g.NG + 1; i <= g.Nxp + g.NG; ++i) {
        size_t o = g.cell.at(i, j, k);
        double u = Q[o + nc], v = Q[o + 2 * nc], w = Q[o + 3 * nc], Tt = Q[o + 5 * nc];
        double c = std::sqrt(p->gamma * p->Rg * Tt);
        double lam = 0;
        for (int d = 0; d < 3; ++d) {
          double a = m[3 * d][o], b = m[3 * d + 1][o], e = m[3 * d + 2][o];
          double s = std::sqrt(a * a + b * b + e * e);
          lam = std::max(lam, std::fabs(a * u + b * v + e * w) + c * s);
        }
        best = std::min(best, 1.0 / (J[o] * lam));
      }
  return cfl * best;
}

// Metrics from coordinates  parse_mesh.jl:139-248 (FP64 throughout).
inline void metrics(const f3d_params* p, const double* x, const double* y, const double* z,
                    double* const m[9], double* J) {
  Grid g(p);
  const long N[3] = {g.Nxt, g.Nyt, g.Nzt};
  const size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  std::vector<double> d(9 * nc);  // dx/dxi, dx/deta, dx/dzeta, dy/..., dz/...
  const double* X[3] = {x, y, z};
  for (int c = 0; c < 3; ++c)
    for (int dir = 0; dir < 3; ++dir) {
      Dir D = make_dir(dir);
      double* out = &d[(size_t)(3 * c + dir) * nc];
#pragma omp parallel for collapse(2)
      for (long k = 1; k <= g.Nzt; ++k)
        for (long j = 1; j <= g.Nyt; ++j)
          for (long i = 1; i <= g.Nxt; ++i) {
            long idx = dir == 0 ? i : (dir == 1 ? j : k);
            auto f = [&](long o) {  // value at idx+o along dir
              return X[c][g.cell.at(i + o * D.di, j + o * D.dj, k + o * D.dk)];
            };
            double r;
            if (idx >= 4 && idx <= N[dir] - 3)  // CD6  parse_mesh.jl:139-142
              r = 1.0 / 60 * (f(3) - f(-3)) - 3.0 / 20 * (f(2) - f(-2)) + 3.0 / 4 * (f(1) - f(-1));
            else if (idx <= 3)  // CD2_L on f[i..i+2]  parse_mesh.jl:144-147
              r = 2 * f(1) - 0.5 * f(2) - 1.5 * f(0);
            else  // CD2_R on f[i-2..i]  parse_mesh.jl:149-152
              r = -2 * f(-1) + 0.5 * f(-2) + 1.5 * f(0);
            out[g.cell.at(i, j, k)] = r;
          }
    }
  const double *xa = &d[0], *xb = &d[nc], *xc = &d[2 * nc];
  const double *ya = &d[3 * nc], *yb = &d[4 * nc], *yc = &d[5 * nc];
  const double *za = &d[6 * nc], *zb = &d[7 * nc], *zc = &d[8 * nc];
#pragma omp parallel for
  for (size_t o = 0; o < nc; ++o) {
    J[o] = 1 / (xa[o] * (yb[o] * zc[o] - yc[o] * zb[o]) - xb[o] * (ya[o] * zc[o] - yc[o] * za[o]) +
                xc[o] * (ya[o] * zb[o] - yb[o] * za[o]));  // parse_mesh.jl:237
    m[0][o] = yb[o] * zc[o] - yc[o] * zb[o];  // parse_mesh.jl:240-248
    m[1][o] = xc[o] * zb[o] - xb[o] * zc[o];
    m[2][o] = xb[o] * yc[o] - xc[o] * yb[o];
    m[3][o] = yc[o] * za[o] - ya[o] * zc[o];
    m[4][o] = xa[o] * zc[o] - xc[o] * za[o];
    m[5][o] = xc[o] * ya[o] - xa[o] * yc[o];
    m[6][o] = ya[o] * zb[o] - yb[o] * za[o];
    m[7][o] = xb[o] * za[o] - xa[o] * zb[o];
    m[8][o] = xa[o] * yb[o] - xb[o] * ya[o];
  }
}

}  // namespace f3d_oracle
