// Device building blocks of the finite-volume branch (FVM.jl) for gfx950.
//
// Face between cells c and c+e_d (0-based cell offsets; stride `st` along d).  Geometry comes from
// the mesh points x, y, z (double in both numeric forms) and is formed in FP64; the reconstruction
// runs in the bulk type T on the conserved state U (double in memory, rounded to T on load).
#pragma once
#include "f3d_device.hpp"

namespace f3d {

// WENO epsilons of the finite-volume kernels (FVM.jl:91), distinct from the finite-difference ones.
constexpr double kFvmEps7 = 1e-10;
constexpr double kFvmEps5 = 1e-8;

// Transverse strides of direction d in the order FVM.jl uses them: i faces (j, k), j faces (i, k),
// k faces (i, j); `flip` = 1 on j faces, whose normal FVM.jl:258-260 negates (that makes every
// normal point along +d).
struct FvmDir {
  long long st, sa, sb;
  int flip;
};
F3D_HD FvmDir fvm_dir(const KP& kp, int d) {
  const long long sy = kp.Nxt, sz = (long long)kp.Nxt * kp.Nyt;
  if (d == 0) return FvmDir{1, sy, sz, 0};
  if (d == 1) return FvmDir{sy, 1, sz, 1};
  return FvmDir{sz, 1, sy, 0};
}

// Per-cell halves of the face diagonals of FVM.jl:16-21, differences first:
//   D1(c) = (X[c+a]-X[c-a]) + (X[c+b]-X[c-b]) + (X[c+a+b]-X[c-a-b])
//   D2(c) = (X[c+b]-X[c-b]) - (X[c+a]-X[c-a]) + (X[c-a+b]-X[c+a-b])
// so that the face between c and c+e has d1 = (D1(c) + D1(c+e))/8, d2 = (D2(c) + D2(c+e))/8 (the
// reference's 12-term sums regrouped; pairing each point with its mirror keeps the cancellation of
// nearly equal coordinates exact on stretched meshes).  P = (D1x, D1y, D1z, D2x, D2y, D2z).
F3D_HD void fvm_partials(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                         long long o, long long sa, long long sb, double P[6]) {
  const double* X[3] = {x, y, z};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double* A = X[c];
    const double da = A[o + sa] - A[o - sa], db = A[o + sb] - A[o - sb];
    const double dpp = A[o + sa + sb] - A[o - sa - sb], dmp = A[o - sa + sb] - A[o + sa - sb];
    P[c] = (da + db) + dpp;
    P[3 + c] = (db - da) + dmp;
  }
}

// Unit normal and area from the partials of the two cells (FVM.jl:22-31), in FP64.
template <typename T>
F3D_HD void fvm_geom(const double* Pc, const double* Pn, int flip, T n[3], T* area) {
  double d1[3], d2[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    d1[c] = 0.125 * (Pc[c] + Pn[c]);
    d2[c] = 0.125 * (Pc[3 + c] + Pn[3 + c]);
  }
  double nx = 0.5 * (d1[1] * d2[2] - d1[2] * d2[1]);
  double ny = 0.5 * (d1[2] * d2[0] - d1[0] * d2[2]);
  double nz = 0.5 * (d1[0] * d2[1] - d1[1] * d2[0]);
  if (flip) {
    nx = -nx;
    ny = -ny;
    nz = -nz;
  }
  const double A = sqrt(nx * nx + ny * ny + nz * nz + 1e-20);
  n[0] = (T)(nx / A);
  n[1] = (T)(ny / A);
  n[2] = (T)(nz / A);
  *area = (T)A;
}

// HLLC flux (FVM.jl:1224-1358) through unit normal n.
template <typename T>
F3D_HD void fvm_hllc(const KP& kp, const T* UL, const T* UR, T nx, T ny, T nz, T* __restrict__ F) {
  const T gam = (T)kp.gamma, half = T(0.5), K = gam - T(1);
  const T rL = UL[0], irL = T(1) / rL, uL = UL[1] * irL, vL = UL[2] * irL, wL = UL[3] * irL, EL = UL[4];
  const T pL = K * (EL - half * rL * (uL * uL + vL * vL + wL * wL));
  const T cL = sqrt(fabs(gam * pL * irL));
  const T qL = uL * nx + vL * ny + wL * nz;
  const T rR = UR[0], irR = T(1) / rR, uR = UR[1] * irR, vR = UR[2] * irR, wR = UR[3] * irR, ER = UR[4];
  const T pR = K * (ER - half * rR * (uR * uR + vR * vR + wR * wR));
  const T cR = sqrt(fabs(gam * pR * irR));
  const T qR = uR * nx + vR * ny + wR * nz;
  const T sqL = sqrt(rL), sqR = sqrt(rR);
  const T isq = T(1) / (sqL + sqR);
  const T ur = (sqL * uL + sqR * uR) * isq, vr = (sqL * vL + sqR * vR) * isq, wr = (sqL * wL + sqR * wR) * isq;
  const T qr = ur * nx + vr * ny + wr * nz;
  const T Hr = (sqL * ((EL + pL) * irL) + sqR * ((ER + pR) * irR)) * isq;
  const T cr = sqrt(fabs(K * (Hr - half * (ur * ur + vr * vr + wr * wr))));
  const T SL = jmin(qL - cL, qr - cr);
  const T SR = jmax(qR + cR, qr + cr);
  const T Ss = (pR - pL + rL * qL * (SL - qL) - rR * qR * (SR - qR)) / (rL * (SL - qL) - rR * (SR - qR) + T(1e-12));
  // one of the four fluxes: side K (L unless the fan is entirely left-going or S* < 0), star or not
  const bool left = SL >= T(0) || (!(SR <= T(0)) && Ss >= T(0));
  const bool star = !(SL >= T(0)) && !(SR <= T(0));
  const T r = left ? rL : rR, u = left ? uL : uR, v = left ? vL : vR, w = left ? wL : wR;
  const T E = left ? EL : ER, p = left ? pL : pR, q = left ? qL : qR, ir = left ? irL : irR;
  const T S = left ? SL : SR;
  F[0] = r * q;
  F[1] = r * u * q + p * nx;
  F[2] = r * v * q + p * ny;
  F[3] = r * w * q + p * nz;
  F[4] = (E + p) * q;
  if (star) {
    const T* Uk = left ? UL : UR;
    const T f = r * (S - q) / (S - Ss);
    const T dq = Ss - q;
    const T Us[5] = {f, f * (u + dq * nx), f * (v + dq * ny), f * (w + dq * nz),
                     f * (E * ir + dq * (Ss + p / (r * (S - q))))};
#pragma unroll
    for (int m = 0; m < 5; ++m) F[m] = F[m] + S * (Us[m] - Uk[m]);
  }
}

// Normalised nonlinear weights d_k/(eps + Is_k*ss)^2.  Double: the literal form.  Float: the same
// normalised weights from ratios to the smallest |eps + Is*ss| (wratio), so that the squares cannot
// overflow float on large conserved values (E ~ 1e4-1e5 in SI units).
template <typename T>
struct FvmW {
  static F3D_HD T mix4(T e, T ss, T I1, T I2, T I3, T I4, T q1, T q2, T q3, T q4) {
    const T t1 = e + I1 * ss, t2 = e + I2 * ss, t3 = e + I3 * ss, t4 = e + I4 * ss;
    const T a1 = T(1) / (t1 * t1), a2 = T(12) / (t2 * t2), a3 = T(18) / (t3 * t3), a4 = T(4) / (t4 * t4);
    const T inv = T(1) / (a1 + a2 + a3 + a4);
    return inv * (a1 * q1 + a2 * q2 + a3 * q3 + a4 * q4);
  }
  static F3D_HD T mix3(T e, T ss, T s1, T s2, T s3, T v1, T v2, T v3) {
    const T t1 = e + s1 * ss, t2 = e + s2 * ss, t3 = e + s3 * ss;
    const T a1 = T(1) / (t1 * t1), a2 = T(6) / (t2 * t2), a3 = T(3) / (t3 * t3);
    const T inv = T(1) / (a1 + a2 + a3);
    return inv * (a1 * v1 + a2 * v2 + a3 * v3);
  }
};
template <>
struct FvmW<float> {
  static F3D_HD float mix4(float e, float ss, float I1, float I2, float I3, float I4, float q1, float q2, float q3,
                           float q4) {
    const float t1 = e + I1 * ss, t2 = e + I2 * ss, t3 = e + I3 * ss, t4 = e + I4 * ss;
    const float tm = fminf(fminf(fabsf(t1), fabsf(t2)), fminf(fabsf(t3), fabsf(t4)));
    const float r1 = wratio(tm, t1), r2 = wratio(tm, t2), r3 = wratio(tm, t3), r4 = wratio(tm, t4);
    const float a1 = r1 * r1, a2 = 12.f * (r2 * r2), a3 = 18.f * (r3 * r3), a4 = 4.f * (r4 * r4);
    return (a1 * q1 + a2 * q2 + a3 * q3 + a4 * q4) / (a1 + a2 + a3 + a4);
  }
  static F3D_HD float mix3(float e, float ss, float s1, float s2, float s3, float v1, float v2, float v3) {
    const float t1 = e + s1 * ss, t2 = e + s2 * ss, t3 = e + s3 * ss;
    const float tm = fminf(fminf(fabsf(t1), fabsf(t2)), fabsf(t3));
    const float r1 = wratio(tm, t1), r2 = wratio(tm, t2), r3 = wratio(tm, t3);
    const float a1 = r1 * r1, a2 = 6.f * (r2 * r2), a3 = 3.f * (r3 * r3);
    return (a1 * v1 + a2 * v2 + a3 * v3) / (a1 + a2 + a3);
  }
};

// WENO7 of FVM.jl:116-147 on V[0..6] (eps 1e-10), divided by 12; float: first-difference indicators.
template <typename T>
F3D_HD T fvm_weno7(const T* V, T ss) {
  const T V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4], V6 = V[5], V7 = V[6];
  const T q1 = -3 * V1 + 13 * V2 - 23 * V3 + 25 * V4;
  const T q2 = V2 - 5 * V3 + 13 * V4 + 3 * V5;
  const T q3 = -V3 + 7 * V4 + 7 * V5 - V6;
  const T q4 = 3 * V4 + 13 * V5 - 5 * V6 + V7;
  if (sizeof(T) == 4) {  // float form: first-difference indicators (weno7_is_diff)
    T J1, J2, J3, J4;
    weno7_is_diff(V, J1, J2, J3, J4);
    return FvmW<T>::mix4((T)kFvmEps7, ss, J1, J2, J3, J4, q1, q2, q3, q4) * (T(1) / T(12));
  }
  const T I1 = V1 * (547 * V1 - 3882 * V2 + 4642 * V3 - 1854 * V4) + V2 * (7043 * V2 - 17246 * V3 + 7042 * V4) +
               V3 * (11003 * V3 - 9402 * V4) + V4 * (2107 * V4);
  const T I2 = V2 * (267 * V2 - 1642 * V3 + 1602 * V4 - 494 * V5) + V3 * (2843 * V3 - 5966 * V4 + 1922 * V5) +
               V4 * (3443 * V4 - 2522 * V5) + V5 * (547 * V5);
  const T I3 = V3 * (547 * V3 - 2522 * V4 + 1922 * V5 - 494 * V6) + V4 * (3443 * V4 - 5966 * V5 + 1602 * V6) +
               V5 * (2843 * V5 - 1642 * V6) + V6 * (267 * V6);
  const T I4 = V4 * (2107 * V4 - 9402 * V5 + 7042 * V6 - 1854 * V7) + V5 * (11003 * V5 - 17246 * V6 + 4642 * V7) +
               V6 * (7043 * V6 - 3882 * V7) + V7 * (547 * V7);
  return FvmW<T>::mix4((T)kFvmEps7, ss, I1, I2, I3, I4, q1, q2, q3, q4) * (T(1) / T(12));
}

// WENO5 of FVM.jl:151-201 on V[1..5] of the 7-point stencil (eps 1e-8), divided by 6.
template <typename T>
F3D_HD T fvm_weno5(const T* V, T ss) {
  const T V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4], V6 = V[5];
  T t1 = V2 - 2 * V3 + V4, t2 = V2 - 4 * V3 + 3 * V4;
  const T s1 = 13 * t1 * t1 + 3 * t2 * t2;
  t1 = V3 - 2 * V4 + V5;
  t2 = V3 - V5;
  const T s2 = 13 * t1 * t1 + 3 * t2 * t2;
  t1 = V4 - 2 * V5 + V6;
  t2 = 3 * V4 - 4 * V5 + V6;
  const T s3 = 13 * t1 * t1 + 3 * t2 * t2;
  const T v1 = 2 * V2 - 7 * V3 + 11 * V4;
  const T v2 = -V3 + 5 * V4 + 2 * V5;
  const T v3 = 2 * V4 + 5 * V5 - V6;
  return FvmW<T>::mix3((T)kFvmEps5, ss, s1, s2, s3, v1, v2, v3) * (T(1) / T(6));
}

// Left / right values of one stencil pair under a non-UP7 scheme (FVM.jl:116-212).  The right
// minmod state limits with minmod(V4R - V3R, V4R - V5R) as the reference writes it.
template <typename T>
F3D_HD void fvm_pair(int sch, const T* VL, const T* VR, T ss, T& vl, T& vr) {
  if (sch == F3D_SCHEME_WENO7) {
    vl = fvm_weno7(VL, ss);
    vr = fvm_weno7(VR, ss);
  } else if (sch == F3D_SCHEME_WENO5) {
    vl = fvm_weno5(VL, ss);
    vr = fvm_weno5(VR, ss);
  } else {
    vl = VL[3] + T(0.5) * minmod(VL[3] - VL[2], VL[4] - VL[3]);
    vr = VR[3] - T(0.5) * minmod(VR[3] - VR[2], VR[3] - VR[4]);
  }
}

// UP7 (Linear) states of FVM.jl:43-61 from the 8-cell stencil W[s] = U at c-3+s.
template <typename T>
F3D_HD void fvm_up7_states(const KP& kp, const T W[8][5], T* UL, T* UR) {
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    T l = 0, r = 0;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      l += Coef<T>::lin(kp, q) * W[q][m];
      r += Coef<T>::lin(kp, q) * W[7 - q][m];
    }
    UL[m] = l;
    UR[m] = r;
  }
}

// Non-UP7 states: component-wise (Conser_reconstruct, FVM.jl:742-842) or characteristic in the
// Roe frame of cells c, c+1 (Eigen_reconstruct, FVM.jl:70-226).  q0/q1: (rho,u,v,w,p) of c, c+1.
template <typename T, bool EIGEN>
F3D_HD void fvm_char_states(const KP& kp, int sch, const T W[8][5], const T* q0, const T* q1, const T n[3], T ss,
                            T* __restrict__ UL, T* __restrict__ UR) {
  if (!EIGEN) {
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      T VL[7], VR[7];
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        VL[s] = W[s][m];
        VR[s] = W[7 - s][m];
      }
      fvm_pair<T>(sch, VL, VR, ss, UL[m], UR[m]);
    }
    return;
  }
  const T gam = (T)kp.gamma;
  const T sr0 = sqrt(q0[0]), sr1 = sqrt(q1[0]), ssum = sr0 + sr1;
  const T u = (sr0 * q0[1] + sr1 * q1[1]) / ssum, v = (sr0 * q0[2] + sr1 * q1[2]) / ssum,
          w = (sr0 * q0[3] + sr1 * q1[3]) / ssum;
  const T g1 = gam / (gam - T(1));
  const T HL = g1 * q0[4] / q0[0] + T(0.5) * (q0[1] * q0[1] + q0[2] * q0[2] + q0[3] * q0[3]);
  const T HR = g1 * q1[4] / q1[0] + T(0.5) * (q1[1] * q1[1] + q1[2] * q1[2] + q1[3] * q1[3]);
  const T H = (sr0 * HL + sr1 * HR) / ssum;
  const T v2 = T(0.5) * (u * u + v * v + w * w);
  const T c = sqrt((gam - T(1)) * (H - v2));
  const T nx = n[0], ny = n[1], nz = n[2];
  T lx, ly, lz;
  if (fabs(nz) <= fabs(ny)) {
    const T den = sqrt(nx * nx + ny * ny + T(1e-12));
    lx = -ny / den; ly = nx / den; lz = 0;
  } else {
    const T den = sqrt(nx * nx + nz * nz + T(1e-12));
    lx = -nz / den; ly = 0; lz = nx / den;
  }
  const T mx = ny * lz - nz * ly, my = nz * lx - nx * lz, mz = nx * ly - ny * lx;
  const T invc = T(1) / c, invc2 = invc * invc, K = gam - T(1);
  const T Ku = K * u * invc2, Kv = K * v * invc2, Kw = K * w * invc2, Kv2 = K * v2 * invc2, Kc2 = K * invc2;
  const T un = u * nx + v * ny + w * nz, ul = u * lx + v * ly + w * lz, um = u * mx + v * my + w * mz;
  const T unc = un * invc, nxc = nx * invc, nyc = ny * invc, nzc = nz * invc;
  const T L[5][5] = {{T(0.5) * (Kv2 + unc), T(-0.5) * (Ku + nxc), T(-0.5) * (Kv + nyc), T(-0.5) * (Kw + nzc), T(0.5) * Kc2},
                     {T(1) - Kv2, Ku, Kv, Kw, -Kc2},
                     {T(0.5) * (Kv2 - unc), T(-0.5) * (Ku - nxc), T(-0.5) * (Kv - nyc), T(-0.5) * (Kw - nzc), T(0.5) * Kc2},
                     {-ul, lx, ly, lz, T(0)},
                     {-um, mx, my, mz, T(0)}};
  T val[2][5];
#pragma unroll
  for (int nn = 0; nn < 5; ++nn) {
    T VL[7], VR[7];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      VL[s] = L[nn][0] * W[s][0] + L[nn][1] * W[s][1] + L[nn][2] * W[s][2] + L[nn][3] * W[s][3] + L[nn][4] * W[s][4];
      VR[s] = L[nn][0] * W[7 - s][0] + L[nn][1] * W[7 - s][1] + L[nn][2] * W[7 - s][2] + L[nn][3] * W[7 - s][3] +
              L[nn][4] * W[7 - s][4];
    }
    fvm_pair<T>(sch, VL, VR, ss, val[0][nn], val[1][nn]);
  }
  // back-projection with the right eigenvectors (FVM.jl:215-226)
#pragma unroll
  for (int sd = 0; sd < 2; ++sd) {
    const T* f = val[sd];
    T* out = sd == 0 ? UL : UR;
    out[0] = f[0] + f[1] + f[2];
    out[1] = f[0] * (u - nx * c) + f[1] * u + f[2] * (u + nx * c) + f[3] * lx + f[4] * mx;
    out[2] = f[0] * (v - ny * c) + f[1] * v + f[2] * (v + ny * c) + f[3] * ly + f[4] * my;
    out[3] = f[0] * (w - nz * c) + f[1] * w + f[2] * (w + nz * c) + f[3] * lz + f[4] * mz;
    out[4] = f[0] * (H - un * c) + f[1] * v2 + f[2] * (H + un * c) + f[3] * ul + f[4] * um;
  }
}

// Scheme from the sensor window of the face's left cell c (FVM.jl:34): max phi over c-2..c+3.
template <typename T>
__device__ __forceinline__ int fvm_scheme(const KP& kp, const T* __restrict__ phi, long long oc, long long st) {
  T m = phi[oc - 2 * st];
#pragma unroll
  for (int s = -1; s <= 3; ++s) m = jmax(m, phi[oc + s * st]);
  return scheme_of<T>(kp, m);
}

// Face-array index of the face (I, J, K) of direction d (extents Nxp+[d==0], Nyp+[d==1], Nzp+[d==2],
// the layout of Fx / Fy / Fz and of the face geometry arrays), and the number of faces.
F3D_HD long long fidx(const KP& kp, int d, int I, int J, int K) {
  const long long ni = kp.Nxp + (d == 0), nj = kp.Nyp + (d == 1);
  return (long long)I + ni * ((long long)J + nj * (long long)K);
}
F3D_HD long long fcount(const KP& kp, int d) {
  return (long long)(kp.Nxp + (d == 0)) * (kp.Nyp + (d == 1)) * (kp.Nzp + (d == 2));
}

// Flux (times area) through the face between cells oc and oc+st of unit normal n and area A, for
// scheme sch: stencil loads, states, HLLC.  Sc, Sc1: |grad xi_d| of the two cells.
template <typename T, bool EIGEN>
__device__ __forceinline__ void fvm_face_flux_n(const KP& kp, int sch, const T* __restrict__ Q,
                                                const double* __restrict__ U, long long oc, long long st,
                                                const T n[3], T A, T Sc, T Sc1, T* __restrict__ F) {
  const long long nc = kp.nc;
  T W[8][5];
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int m = 0; m < 5; ++m) W[s][m] = (T)U[oc + (s - 3) * st + m * nc];
  T UL[5], UR[5];
  if (sch == F3D_SCHEME_UP7) {
    fvm_up7_states<T>(kp, W, UL, UR);
  } else {
    T q0[5], q1[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      q0[m] = Q[oc + m * nc];
      q1[m] = Q[oc + st + m * nc];
    }
    fvm_char_states<T, EIGEN>(kp, sch, W, q0, q1, n, T(2) / (Sc1 + Sc), UL, UR);
  }
  fvm_hllc<T>(kp, UL, UR, n[0], n[1], n[2], F);
#pragma unroll
  for (int m = 0; m < 5; ++m) F[m] *= A;
}

// As fvm_face_flux_n with the geometry formed from the mesh points.
template <typename T, bool EIGEN>
__device__ __forceinline__ void fvm_face_flux(const KP& kp, int sch, const T* __restrict__ Q,
                                              const double* __restrict__ U, const double* __restrict__ x,
                                              const double* __restrict__ y, const double* __restrict__ z,
                                              long long oc, const FvmDir& dd, T Sc, T Sc1, T* __restrict__ F) {
  double Pc[6], Pn[6];
  fvm_partials(x, y, z, oc, dd.sa, dd.sb, Pc);
  fvm_partials(x, y, z, oc + dd.st, dd.sa, dd.sb, Pn);
  T n[3], A;
  fvm_geom<T>(Pc, Pn, dd.flip, n, &A);
  fvm_face_flux_n<T, EIGEN>(kp, sch, Q, U, oc, dd.st, n, A, Sc, Sc1, F);
}

}  // namespace f3d
