// CPU oracle, finite-volume branch of flowAdvance (solver.jl:21-31; FVM.jl:6-1358).
//
// TEST INFRASTRUCTURE ONLY, like f3d_oracle.hpp: a plain CPU restatement, written from reading
// FVM.jl, in the reference's 1-based indexing and evaluation order.  T is the bulk type (double,
// float for the FP32-bulk form, long double for the rounding-floor reference); U and the mesh
// coordinates are double arrays in every form.  The face geometry is formed in FP64 (long double
// for T = long double) and rounded to T: that is this build's precision choice for the FP32-bulk
// form (the reference forms it in Float32 from Float32 coordinates), and so is the float form's
// evaluation of the WENO7 indicators from first differences (weno7_is_diff).  The FP64 form uses the decimal
// values of the reference's Float32 literals (1e-10, 1e-8, 1e-12, 1e-20).
//
// One behaviour of the reference is kept as written: the minmod branch of the right state
// (FVM.jl:210-211, 416-417, 634-635, 826-827, ...) limits with minmod(V4R - V3R, V4R - V5R), two
// differences of opposite sign on monotone data, so the right state is first order there (the
// left state uses the usual minmod(V4L - V3L, V5L - V4L)).
#pragma once
#include "f3d_oracle.hpp"

namespace f3d_oracle {

// Type the face geometry is formed in.
template <typename T>
struct GeoT {
  using type = double;
};
template <>
struct GeoT<long double> {
  using type = long double;
};


// The two transverse dimensions of direction d in the order FVM.jl uses them: i faces (j, k)
// (FVM.jl:16-21), j faces (i, k) (FVM.jl:250-255), k faces (i, j) (FVM.jl:465-470).
inline void fvm_transverse(int d, int& a, int& b) {
  a = d == 0 ? 1 : 0;
  b = d == 2 ? 1 : 2;
}

// Unit normal and area of the face between cells c and c + e_d, c = (i, j, k) 1-based:
// two diagonal vectors d1, d2 from 12-point differences of the mesh points around the face
// (FVM.jl:16-21), n = (d1 x d2)/2 (FVM.jl:22-24; negated on j faces, FVM.jl:258-260),
// Area = sqrt(|n|^2 + 1e-20), n /= Area (FVM.jl:25-28).
template <typename G>
void fvm_face_geometry(const Grid& g, int d, const double* const X[3], long i, long j, long k, G n[3], G* area) {
  int ta, tb;
  fvm_transverse(d, ta, tb);
  auto at = [&](int c, int oe, int oa, int ob) -> G {
    long o[3] = {i, j, k};
    o[d] += oe;
    o[ta] += oa;
    o[tb] += ob;
    return (G)X[c][g.cell.at(o[0], o[1], o[2])];
  };
  G d1[3], d2[3];
  for (int c = 0; c < 3; ++c) {
    d1[c] = G(0.125) * (at(c, 0, 1, 0) + at(c, 1, 1, 0) + at(c, 0, 0, 1) + at(c, 1, 0, 1) + at(c, 0, 1, 1) +
                        at(c, 1, 1, 1) - at(c, 0, -1, 0) - at(c, 1, -1, 0) - at(c, 0, 0, -1) - at(c, 1, 0, -1) -
                        at(c, 0, -1, -1) - at(c, 1, -1, -1));
    d2[c] = G(0.125) * (at(c, 0, -1, 0) + at(c, 1, -1, 0) + at(c, 0, 0, 1) + at(c, 1, 0, 1) + at(c, 0, -1, 1) +
                        at(c, 1, -1, 1) - at(c, 0, 1, 0) - at(c, 1, 1, 0) - at(c, 0, 0, -1) - at(c, 1, 0, -1) -
                        at(c, 0, 1, -1) - at(c, 1, 1, -1));
  }
  G nx, ny, nz;
  if (d == 1) {  // FVM.jl:258-260
    nx = G(0.5) * (-d1[1] * d2[2] + d1[2] * d2[1]);
    ny = G(0.5) * (-d1[2] * d2[0] + d1[0] * d2[2]);
    nz = G(0.5) * (-d1[0] * d2[1] + d1[1] * d2[0]);
  } else {  // FVM.jl:22-24, 473-475
    nx = G(0.5) * (d1[1] * d2[2] - d1[2] * d2[1]);
    ny = G(0.5) * (d1[2] * d2[0] - d1[0] * d2[2]);
    nz = G(0.5) * (d1[0] * d2[1] - d1[1] * d2[0]);
  }
  const G A = std::sqrt(nx * nx + ny * ny + nz * nz + G(1e-20));
  n[0] = nx / A;
  n[1] = ny / A;
  n[2] = nz / A;
  *area = A;
}

// HLLC flux through a face of unit normal n (FVM.jl:1224-1358): Roe-averaged wave-speed estimates
// SL = min(qL - cL, q_roe - c_roe), SR = max(qR + cR, q_roe + c_roe); contact speed S* with a 1e-12
// guard in its denominator; F_L, F_R, or the star-state fluxes F*_K = F_K + S_K (U*_K - U_K).
template <typename T>
void hllc_flux(const f3d_params* p, const T UL[5], const T UR[5], T nx, T ny, T nz, T F[5]) {
  const T gam = (T)p->gamma, half = T(0.5);
  const T rL = UL[0], irL = T(1) / rL, uL = UL[1] * irL, vL = UL[2] * irL, wL = UL[3] * irL, EL = UL[4];
  const T pL = (gam - 1) * (EL - half * rL * (uL * uL + vL * vL + wL * wL));
  const T cL = std::sqrt(std::fabs(gam * pL * irL));
  const T qL = uL * nx + vL * ny + wL * nz;
  const T rR = UR[0], irR = T(1) / rR, uR = UR[1] * irR, vR = UR[2] * irR, wR = UR[3] * irR, ER = UR[4];
  const T pR = (gam - 1) * (ER - half * rR * (uR * uR + vR * vR + wR * wR));
  const T cR = std::sqrt(std::fabs(gam * pR * irR));
  const T qR = uR * nx + vR * ny + wR * nz;
  const T sqL = std::sqrt(rL), sqR = std::sqrt(rR);
  const T isq = T(1) / (sqL + sqR);
  const T ur = (sqL * uL + sqR * uR) * isq, vr = (sqL * vL + sqR * vR) * isq, wr = (sqL * wL + sqR * wR) * isq;
  const T qr = ur * nx + vr * ny + wr * nz;
  const T HL = (EL + pL) * irL, HR = (ER + pR) * irR;
  const T Hr = (sqL * HL + sqR * HR) * isq;
  const T v2r = ur * ur + vr * vr + wr * wr;
  const T cr = std::sqrt(std::fabs((gam - 1) * (Hr - half * v2r)));
  const T SL = jmin(qL - cL, qr - cr);
  const T SR = jmax(qR + cR, qr + cr);
  const T Ss = (pR - pL + rL * qL * (SL - qL) - rR * qR * (SR - qR)) / (rL * (SL - qL) - rR * (SR - qR) + T(1e-12));
  if (SL >= 0) {  // FVM.jl:1284-1292
    F[0] = rL * qL;
    F[1] = rL * uL * qL + pL * nx;
    F[2] = rL * vL * qL + pL * ny;
    F[3] = rL * wL * qL + pL * nz;
    F[4] = (EL + pL) * qL;
  } else if (SR <= 0) {  // FVM.jl:1294-1302
    F[0] = rR * qR;
    F[1] = rR * uR * qR + pR * nx;
    F[2] = rR * vR * qR + pR * ny;
    F[3] = rR * wR * qR + pR * nz;
    F[4] = (ER + pR) * qR;
  } else if (SL < 0 && Ss >= 0) {  // FVM.jl:1304-1327
    const T FL[5] = {rL * qL, rL * uL * qL + pL * nx, rL * vL * qL + pL * ny, rL * wL * qL + pL * nz, (EL + pL) * qL};
    const T f = rL * (SL - qL) / (SL - Ss);
    const T Us[5] = {f, f * (uL + (Ss - qL) * nx), f * (vL + (Ss - qL) * ny), f * (wL + (Ss - qL) * nz),
                     f * (EL * irL + (Ss - qL) * (Ss + pL / (rL * (SL - qL))))};
    for (int q = 0; q < 5; ++q) F[q] = FL[q] + SL * (Us[q] - UL[q]);
  } else {  // FVM.jl:1329-1356
    const T FR[5] = {rR * qR, rR * uR * qR + pR * nx, rR * vR * qR + pR * ny, rR * wR * qR + pR * nz, (ER + pR) * qR};
    const T f = rR * (SR - qR) / (SR - Ss);
    const T Us[5] = {f, f * (uR + (Ss - qR) * nx), f * (vR + (Ss - qR) * ny), f * (wR + (Ss - qR) * nz),
                     f * (ER * irR + (Ss - qR) * (Ss + pR / (rR * (SR - qR))))};
    for (int q = 0; q < 5; ++q) F[q] = FR[q] + SR * (Us[q] - UR[q]);
  }
}

// Nonlinear weights of the finite-volume WENO kernels: the literal form of FVM.jl:140-145 and
// 176-181 for double and long double; for float (this build's FP32-bulk form) the same normalised
// weights from ratios to the smallest |eps + Is*ss| (wratio), so that the squares cannot overflow
// float on large conserved values.
template <typename T>
struct FvmMix {
  static T w7(T e, T ss, T Is1, T Is2, T Is3, T Is4, T q1, T q2, T q3, T q4) {
    const T t1 = e + Is1 * ss, t2 = e + Is2 * ss, t3 = e + Is3 * ss, t4 = e + Is4 * ss;
    const T a1 = T(1) / (t1 * t1), a2 = T(12) / (t2 * t2), a3 = T(18) / (t3 * t3), a4 = T(4) / (t4 * t4);
    const T invsum = T(1) / (a1 + a2 + a3 + a4);
    return invsum * (a1 * q1 + a2 * q2 + a3 * q3 + a4 * q4);
  }
  static T w5(T e, T ss, T s1, T s2, T s3, T v1, T v2, T v3) {
    const T d1 = e + s1 * ss, d2 = e + s2 * ss, d3 = e + s3 * ss;
    const T a1 = T(1) / (d1 * d1), a2 = T(6) / (d2 * d2), a3 = T(3) / (d3 * d3);
    const T invsum = T(1) / (a1 + a2 + a3);
    return invsum * (a1 * v1 + a2 * v2 + a3 * v3);
  }
};
template <>
struct FvmMix<float> {
  static float w7(float e, float ss, float Is1, float Is2, float Is3, float Is4, float q1, float q2, float q3,
                  float q4) {
    const float t1 = e + Is1 * ss, t2 = e + Is2 * ss, t3 = e + Is3 * ss, t4 = e + Is4 * ss;
    const float tm = std::min(std::min(std::fabs(t1), std::fabs(t2)), std::min(std::fabs(t3), std::fabs(t4)));
    const float r1 = wratio(tm, t1), r2 = wratio(tm, t2), r3 = wratio(tm, t3), r4 = wratio(tm, t4);
    const float a1 = r1 * r1, a2 = 12.f * (r2 * r2), a3 = 18.f * (r3 * r3), a4 = 4.f * (r4 * r4);
    return (a1 * q1 + a2 * q2 + a3 * q3 + a4 * q4) / (a1 + a2 + a3 + a4);
  }
  static float w5(float e, float ss, float s1, float s2, float s3, float v1, float v2, float v3) {
    const float t1 = e + s1 * ss, t2 = e + s2 * ss, t3 = e + s3 * ss;
    const float tm = std::min(std::min(std::fabs(t1), std::fabs(t2)), std::fabs(t3));
    const float r1 = wratio(tm, t1), r2 = wratio(tm, t2), r3 = wratio(tm, t3);
    const float a1 = r1 * r1, a2 = 6.f * (r2 * r2), a3 = 3.f * (r3 * r3);
    return (a1 * v1 + a2 * v2 + a3 * v3) / (a1 + a2 + a3);
  }
};

// WENO7 of the finite-volume kernels (FVM.jl:116-147): candidates q1..q4, Balsara-Shu indicators,
// alpha_k = d_k / (1e-10 + Is_k*ss)^2 with d = (1, 12, 18, 4), result / 12.
template <typename T>
inline T fvm_weno7(const T V[7], T ss) {
  const T V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4], V6 = V[5], V7 = V[6];
  const T e = T(1e-10), tmp1 = T(1) / T(12);
  const T q1 = T(-3) * V1 + T(13) * V2 - T(23) * V3 + T(25) * V4;
  const T q2 = T(1) * V2 - T(5) * V3 + T(13) * V4 + T(3) * V5;
  const T q3 = T(-1) * V3 + T(7) * V4 + T(7) * V5 - T(1) * V6;
  const T q4 = T(3) * V4 + T(13) * V5 - T(5) * V6 + T(1) * V7;
  const T Is1 = V1 * (T(547) * V1 - T(3882) * V2 + T(4642) * V3 - T(1854) * V4) +
                V2 * (T(7043) * V2 - T(17246) * V3 + T(7042) * V4) + V3 * (T(11003) * V3 - T(9402) * V4) +
                V4 * (T(2107) * V4);
  const T Is2 = V2 * (T(267) * V2 - T(1642) * V3 + T(1602) * V4 - T(494) * V5) +
                V3 * (T(2843) * V3 - T(5966) * V4 + T(1922) * V5) + V4 * (T(3443) * V4 - T(2522) * V5) +
                V5 * (T(547) * V5);
  const T Is3 = V3 * (T(547) * V3 - T(2522) * V4 + T(1922) * V5 - T(494) * V6) +
                V4 * (T(3443) * V4 - T(5966) * V5 + T(1602) * V6) + V5 * (T(2843) * V5 - T(1642) * V6) +
                V6 * (T(267) * V6);
  const T Is4 = V4 * (T(2107) * V4 - T(9402) * V5 + T(7042) * V6 - T(1854) * V7) +
                V5 * (T(11003) * V5 - T(17246) * V6 + T(4642) * V7) + V6 * (T(7043) * V6 - T(3882) * V7) +
                V7 * (T(547) * V7);
  if (std::is_same<T, float>::value) {  // the float form's indicators: first-difference quadratic forms
    T J[4];
    weno7_is_diff(V, J);
    return FvmMix<T>::w7(e, ss, J[0], J[1], J[2], J[3], q1, q2, q3, q4) * tmp1;
  }
  return FvmMix<T>::w7(e, ss, Is1, Is2, Is3, Is4, q1, q2, q3, q4) * tmp1;
}

// WENO5 of the finite-volume kernels (FVM.jl:151-201) on V2..V6 of the 7-point stencil:
// beta_k = 13 t1^2 + 3 t2^2, alpha = (1, 6, 3) / (1e-8 + beta*ss)^2, result / 6.
template <typename T>
inline T fvm_weno5(const T V[7], T ss) {
  const T V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4], V6 = V[5];
  const T e = T(1e-8), tmp2 = T(1) / T(6);
  T t1 = V2 - T(2) * V3 + V4, t2 = V2 - T(4) * V3 + T(3) * V4;
  const T s1 = T(13) * t1 * t1 + T(3) * t2 * t2;
  t1 = V3 - T(2) * V4 + V5;
  t2 = V3 - V5;
  const T s2 = T(13) * t1 * t1 + T(3) * t2 * t2;
  t1 = V4 - T(2) * V5 + V6;
  t2 = T(3) * V4 - T(4) * V5 + V6;
  const T s3 = T(13) * t1 * t1 + T(3) * t2 * t2;
  const T v1 = T(2) * V2 - T(7) * V3 + T(11) * V4;
  const T v2 = T(-1) * V3 + T(5) * V4 + T(2) * V5;
  const T v3 = T(2) * V4 + T(5) * V5 - T(1) * V6;
  return FvmMix<T>::w5(e, ss, s1, s2, s3, v1, v2, v3) * tmp2;
}

// Reconstructed value of one 7-point stencil VL (left state: cells c-3..c+3) and VR (right state:
// cells c+4..c-2) under scheme sch != UP7 (FVM.jl:116-212).
template <typename T>
inline void fvm_reconstruct_pair(int sch, const T VL[7], const T VR[7], T ss, T& valL, T& valR) {
  if (sch == F3D_SCHEME_WENO7) {
    valL = fvm_weno7(VL, ss);
    valR = fvm_weno7(VR, ss);
  } else if (sch == F3D_SCHEME_WENO5) {
    valL = fvm_weno5(VL, ss);
    valR = fvm_weno5(VR, ss);
  } else {  // FVM.jl:208-211 (right state as written: minmod(V4R - V3R, V4R - V5R))
    valL = VL[3] + T(0.5) * minmod(VL[3] - VL[2], VL[4] - VL[3]);
    valR = VR[3] - T(0.5) * minmod(VR[3] - VR[2], VR[3] - VR[4]);
  }
}

// States on both sides of the face between cells c and c + e_d (c = (i, j, k)), FVM.jl:33-214
// (eigen) / 707-846 (conser), and the scheme the 6-cell sensor maximum selects.
template <typename T>
int fvm_states(const f3d_params* p, const Grid& g, int d, bool eigen, const T* Q, const double* U, const T* phi,
               const T* S, long i, long j, long k, const T n[3], T UL[5], T UR[5]) {
  const Dir D = make_dir(d);
  auto cell = [&](long s) { return g.cell.at(i + s * D.di, j + s * D.dj, k + s * D.dk); };
  const size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  auto Uv = [&](long s, int m) { return (T)U[cell(s) + (size_t)m * nc]; };  // m 0-based
  // FVM.jl:34
  T phx = phi[cell(-2)];
  for (long s = -1; s <= 3; ++s) phx = jmax(phx, phi[cell(s)]);
  int sch;
  if (phx < (T)p->hybrid_phi1) sch = F3D_SCHEME_UP7;
  else if (phx < (T)p->hybrid_phi2) sch = F3D_SCHEME_WENO7;
  else if (phx < (T)p->hybrid_phi3) sch = F3D_SCHEME_WENO5;
  else sch = F3D_SCHEME_NND2;

  if (sch == F3D_SCHEME_UP7) {  // FVM.jl:43-61
    T c[7];
    linear_coeffs<T>(p, c);
    for (int m = 0; m < 5; ++m) {
      UL[m] = c[0] * Uv(-3, m) + c[1] * Uv(-2, m) + c[2] * Uv(-1, m) + c[3] * Uv(0, m) + c[4] * Uv(1, m) +
              c[5] * Uv(2, m) + c[6] * Uv(3, m);
      UR[m] = c[0] * Uv(4, m) + c[1] * Uv(3, m) + c[2] * Uv(2, m) + c[3] * Uv(1, m) + c[4] * Uv(0, m) +
              c[5] * Uv(-1, m) + c[6] * Uv(-2, m);
    }
    return sch;
  }
  const T ss = T(2) / (S[cell(1)] + S[cell(0)]);  // FVM.jl:107
  if (!eigen) {  // FVM.jl:742-842: component-wise
    for (int m = 0; m < 5; ++m) {
      T VL[7], VR[7];
      for (int s = 0; s < 7; ++s) {
        VL[s] = Uv(s - 3, m);
        VR[s] = Uv(4 - s, m);
      }
      fvm_reconstruct_pair(sch, VL, VR, ss, UL[m], UR[m]);
    }
    return sch;
  }
  // Roe-averaged frame of cells c, c+1  FVM.jl:70-102
  const T gam = (T)p->gamma;
  const size_t o0 = cell(0), o1 = cell(1);
  const T r0 = Q[o0], r1 = Q[o1];
  const T sr0 = std::sqrt(r0), sr1 = std::sqrt(r1);
  const T u = (sr0 * Q[o0 + nc] + sr1 * Q[o1 + nc]) / (sr0 + sr1);
  const T v = (sr0 * Q[o0 + 2 * nc] + sr1 * Q[o1 + 2 * nc]) / (sr0 + sr1);
  const T w = (sr0 * Q[o0 + 3 * nc] + sr1 * Q[o1 + 3 * nc]) / (sr0 + sr1);
  const T HL = gam / (gam - T(1)) * Q[o0 + 4 * nc] / r0 +
               T(0.5) * (Q[o0 + nc] * Q[o0 + nc] + Q[o0 + 2 * nc] * Q[o0 + 2 * nc] + Q[o0 + 3 * nc] * Q[o0 + 3 * nc]);
  const T HR = gam / (gam - T(1)) * Q[o1 + 4 * nc] / r1 +
               T(0.5) * (Q[o1 + nc] * Q[o1 + nc] + Q[o1 + 2 * nc] * Q[o1 + 2 * nc] + Q[o1 + 3 * nc] * Q[o1 + 3 * nc]);
  const T H = (sr0 * HL + sr1 * HR) / (sr0 + sr1);
  const T v2 = T(0.5) * (u * u + v * v + w * w);
  const T c = std::sqrt((gam - T(1)) * (H - v2));
  const T nx = n[0], ny = n[1], nz = n[2];
  T lx, ly, lz;
  if (std::fabs(nz) <= std::fabs(ny)) {
    const T den = std::sqrt(nx * nx + ny * ny + T(1e-12));
    lx = -ny / den;
    ly = nx / den;
    lz = 0;
  } else {
    const T den = std::sqrt(nx * nx + nz * nz + T(1e-12));
    lx = -nz / den;
    ly = 0;
    lz = nx / den;
  }
  const T mx = ny * lz - nz * ly, my = nz * lx - nx * lz, mz = nx * ly - ny * lx;
  const T invc = T(1) / c, invc2 = invc * invc;
  const T K = gam - T(1);
  const T Ku = K * u * invc2, Kv = K * v * invc2, Kw = K * w * invc2;
  const T Kv2 = K * v2 * invc2, Kc2 = K * invc2;
  const T un = u * nx + v * ny + w * nz, ul = u * lx + v * ly + w * lz, um = u * mx + v * my + w * mz;
  const T un_invc = un * invc, nx_invc = nx * invc, ny_invc = ny * invc, nz_invc = nz * invc;
  const T half = T(0.5), mhalf = T(-0.5);
  const T L[5][5] = {  // FVM.jl:111-116
      {half * (Kv2 + un_invc), mhalf * (Ku + nx_invc), mhalf * (Kv + ny_invc), mhalf * (Kw + nz_invc), half * Kc2},
      {T(1) - Kv2, Ku, Kv, Kw, -Kc2},
      {half * (Kv2 - un_invc), mhalf * (Ku - nx_invc), mhalf * (Kv - ny_invc), mhalf * (Kw - nz_invc), half * Kc2},
      {-ul, lx, ly, lz, T(0)},
      {-um, mx, my, mz, T(0)}};
  const T R[5][5] = {  // rows n of FVM.jl:215-220: (rn1..rn5)
      {T(1), u - nx * c, v - ny * c, w - nz * c, H - un * c},
      {T(1), u, v, w, v2},
      {T(1), u + nx * c, v + ny * c, w + nz * c, H + un * c},
      {T(0), lx, ly, lz, ul},
      {T(0), mx, my, mz, um}};
  for (int m = 0; m < 5; ++m) UL[m] = UR[m] = T(0);
  for (int nn = 0; nn < 5; ++nn) {
    const T* l = L[nn];
    T VL[7], VR[7];
    for (int s = 0; s < 7; ++s) {  // FVM.jl:119-134
      VL[s] = l[0] * Uv(s - 3, 0) + l[1] * Uv(s - 3, 1) + l[2] * Uv(s - 3, 2) + l[3] * Uv(s - 3, 3) +
              l[4] * Uv(s - 3, 4);
      VR[s] = l[0] * Uv(4 - s, 0) + l[1] * Uv(4 - s, 1) + l[2] * Uv(4 - s, 2) + l[3] * Uv(4 - s, 3) +
              l[4] * Uv(4 - s, 4);
    }
    T valL, valR;
    fvm_reconstruct_pair(sch, VL, VR, ss, valL, valR);
    for (int m = 0; m < 5; ++m) {  // FVM.jl:222-226
      UL[m] += R[nn][m] * valL;
      UR[m] += R[nn][m] * valR;
    }
  }
  return sch;
}

// Eigen_reconstruct_{i,j,k} (eigen = true; FVM.jl:6-670) / Conser_reconstruct_{i,j,k}
// (FVM.jl:672-1222): face flux F = HLLC(UL, UR, n) * Area into the face array of direction d.
// Optional outputs per face: UL/UR (5 each, face layout) and geo (nx, ny, nz, Area).
template <typename T>
void fvm_reconstruct(const f3d_params* p, int d, bool eigen, const T* Q, const double* U, const T* phi, const T* S,
                     T* F, const double* x, const double* y, const double* z, T* ULo = nullptr, T* URo = nullptr,
                     T* geo = nullptr) {
  typedef typename GeoT<T>::type G;
  Grid g(p);
  const Dir D = make_dir(d);
  const Ext& fe = d == 0 ? g.fx : (d == 1 ? g.fy : g.fz);
  const double* const X[3] = {x, y, z};
#pragma omp parallel for collapse(2)
  for (long k = g.NG + 1 - D.dk; k <= g.Nzp + g.NG; ++k)
    for (long j = g.NG + 1 - D.dj; j <= g.Nyp + g.NG; ++j)
      for (long i = g.NG + 1 - D.di; i <= g.Nxp + g.NG; ++i) {
        G nG[3], AG;
        fvm_face_geometry<G>(g, d, X, i, j, k, nG, &AG);
        const T n[3] = {(T)nG[0], (T)nG[1], (T)nG[2]}, A = (T)AG;
        T UL[5], UR[5], Fl[5];
        fvm_states<T>(p, g, d, eigen, Q, U, phi, S, i, j, k, n, UL, UR);
        hllc_flux<T>(p, UL, UR, n[0], n[1], n[2], Fl);
        const long fi = i - g.NG + D.di, fj = j - g.NG + D.dj, fk = k - g.NG + D.dk;
        for (int m = 0; m < 5; ++m) {
          F[fe.at(fi, fj, fk, m + 1)] = Fl[m] * A;
          if (ULo) ULo[fe.at(fi, fj, fk, m + 1)] = UL[m];
          if (URo) URo[fe.at(fi, fj, fk, m + 1)] = UR[m];
        }
        if (geo)
          for (int m = 0; m < 3; ++m) geo[fe.at(fi, fj, fk, m + 1)] = n[m];
        if (geo) geo[fe.at(fi, fj, fk, 4)] = A;
      }
}

// flowAdvance, finite-volume branch  solver.jl:21-31, then viscousFlux (solver.jl:185-187).
template <typename T>
void flow_advance_fvm(const f3d_params* p, const double* U, const T* Q, T* Fx, T* Fy, T* Fz, T* Fvx, T* Fvy, T* Fvz,
                      const T* s1, const T* s2, const T* s3, const T* const m[9], const double* J, const double* x,
                      const double* y, const double* z, const T* phi) {
  const bool eigen = p->fvm == F3D_FVM_EIGEN;
  fvm_reconstruct<T>(p, 0, eigen, Q, U, phi, s1, Fx, x, y, z);
  fvm_reconstruct<T>(p, 1, eigen, Q, U, phi, s2, Fy, x, y, z);
  fvm_reconstruct<T>(p, 2, eigen, Q, U, phi, s3, Fz, x, y, z);
  if (p->viscous) viscous_flux<T>(p, Fvx, Fvy, Fvz, Q, m, J);
}

// One RK3 stage with the finite-volume branch (solver.jl:416-444), local time stepping as in
// rk3_stage; fillGhost also refills U's ghost cells, which the next stage's reconstruction reads.
template <typename T>
void rk3_stage_fvm(const f3d_params* p, Scratch<T>& sc, int KRK, double* U, double* Un, T* Q, const T* s1,
                   const T* s2, const T* s3, const T* const m[9], const double* J, const double* x, const double* y,
                   const double* z, double dt, const T* inflow) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  if (KRK == 1) std::memcpy(Un, U, nc * 5 * sizeof(double));
  if (KRK == 1 && p->lts) compute_dt<T>(p, sc.ldt.data(), Q, J, s1, s2, s3);
  shock_sensor<T>(p, sc.phi.data(), Q);
  flow_advance_fvm<T>(p, U, Q, sc.Fx.data(), sc.Fy.data(), sc.Fz.data(), sc.Fvx.data(), sc.Fvy.data(), sc.Fvz.data(),
                      s1, s2, s3, m, J, x, y, z, sc.phi.data());
  div<T>(p, U, sc.Fx.data(), sc.Fy.data(), sc.Fz.data(), sc.Fvx.data(), sc.Fvy.data(), sc.Fvz.data(), dt, J,
         p->lts ? sc.ldt.data() : nullptr);
  if (KRK == 2) lin_comb(p, U, Un, 5, 0.25, 0.75);
  if (KRK == 3) lin_comb(p, U, Un, 5, 2.0 / 3.0, 1.0 / 3.0);
  c2prim<T>(p, U, Q);
  exchange_ghost<T>(p, Q, 6);
  fill_ghost<T>(p, Q, U, inflow);
}

}  // namespace f3d_oracle
