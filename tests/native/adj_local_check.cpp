// CPU check of the local reverse-mode pieces of the HIP library (f3d_adjoint.hpp) against
// forward-mode dual-number instantiations of the forward pieces (f3d_device.hpp) on random
// inputs.  Test infrastructure: built host-only by tests/test_adjoint_local.py.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../flame3d-mhd_amd/csrc/f3d_adjoint.hpp"
#include "../../oracle/f3d_dual.hpp"

using f3d_oracle::Dual;

namespace f3d {
template <>
struct Coef<Dual> {
  static Dual lin(const KP& kp, int q) { return Dual(kp.lin[q]); }
  static Dual gRg(const KP& kp) { return Dual(kp.gRg); }
  static Dual gm1(const KP& kp) { return Dual(kp.gm1); }
  static Dual w2c(const KP& kp) { return Dual(kp.w2c); }
  static Dual inv2g(const KP& kp) { return Dual(kp.inv2g); }
};
template <>
struct W7<Dual> {
  static Dual mix(Dual eps, Dual s0, Dual I1, Dual I2, Dual I3, Dual I4, Dual q1, Dual q2, Dual q3, Dual q4) {
    const Dual b1 = eps + I1 * s0, b2 = eps + I2 * s0, b3 = eps + I3 * s0, b4 = eps + I4 * s0;
    const Dual a1 = 1.0 / (b1 * b1), a2 = 12.0 / (b2 * b2), a3 = 18.0 / (b3 * b3), a4 = 4.0 / (b4 * b4);
    const Dual inv = 1.0 / (a1 + a2 + a3 + a4);
    return inv * (a1 * q1 + a2 * q2 + a3 * q3 + a4 * q4);
  }
};
}  // namespace f3d

static f3d::KP make_kp_local() {
  f3d::KP kp{};
  kp.gamma = 1.4; kp.Rg = 287.0; kp.Cp = 287.0 * 1.4 / 0.4; kp.Cs = 1.458e-6; kp.Ts = 110.4; kp.Pr = 0.72;
  kp.eps7 = 1e-20; kp.eps5 = 1e-16;
  const double up7[7] = {-3.0 / 420, 25.0 / 420, -101.0 / 420, 319.0 / 420, 214.0 / 420, -38.0 / 420, 4.0 / 420};
  for (int q = 0; q < 7; ++q) kp.lin[q] = up7[q];
  kp.gRg = kp.gamma * kp.Rg; kp.gm1 = kp.gamma - 1; kp.w2c = (3 - kp.gamma) / (2 * (kp.gamma - 1));
  kp.inv2g = 1 / (2 * kp.gamma);
  return kp;
}

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

static int failures = 0;
static void report(const char* what, double err, double scale, double tol) {
  const double rel = err / (scale > 0 ? scale : 1);
  std::printf("%-28s rel %.3e (scale %.3e)\n", what, rel, scale);
  if (!(rel <= tol)) ++failures;
}

static void check_split(const f3d::KP& kp, double mach) {
  for (int trial = 0; trial < 50; ++trial) {
    double x[5] = {U(0.2, 2.0), U(-1, 1), U(-1, 1), U(-1, 1), U(100, 600)};
    const double c = std::sqrt(kp.gRg * x[4]);
    for (int q = 1; q < 4; ++q) x[q] *= mach * c;
    const double A[3] = {U(-1, 1), U(-1, 1), U(-1, 1)};
    const double S = std::sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
    double gp[5], gm[5];
    for (int n = 0; n < 5; ++n) { gp[n] = U(-1, 1); gm[n] = U(-1, 1); }
    double g[5];
    f3d::sw_split_vjp(kp, x[0], x[1], x[2], x[3], x[4], A[0], A[1], A[2], S, gp, gm, g);
    double err = 0, scale = 0;
    for (int i = 0; i < 5; ++i) {
      Dual xd[5];
      for (int q = 0; q < 5; ++q) xd[q] = Dual(x[q], q == i ? 1.0 : 0.0);
      Dual fp[5], fm[5];
      f3d::sw_split<Dual>(kp, xd[0], xd[1], xd[2], xd[3], xd[4], Dual(A[0]), Dual(A[1]), Dual(A[2]), Dual(S), fp, fm);
      double ref = 0;
      for (int n = 0; n < 5; ++n) ref += gp[n] * fp[n].d + gm[n] * fm[n].d;
      err = std::fmax(err, std::fabs(ref - g[i]) * (i == 4 ? x[4] : (i == 0 ? 1.0 : c)));
      scale = std::fmax(scale, std::fabs(ref) * (i == 4 ? x[4] : (i == 0 ? 1.0 : c)));
    }
    if (trial == 49) report(mach > 1 ? "sw_split (supersonic)" : "sw_split (subsonic)", err, scale, 1e-12);
  }
}

// dir 0: x-like normal, 1: y-like, 2: pure z (tangent branch 2); flat2d: v = w = 0 exactly
static void check_face(const f3d::KP& kp, int sch, const char* name, double roughness, int dir = 0,
                       bool flat2d = false, double uinf = 200) {
  double worst = 0, wscale = 0;
  for (int trial = 0; trial < 30; ++trial) {
    // a smooth-ish cell state along the stencil with a jump of size `roughness`
    double fp[8][5], fm[8][5], qs[8][4];
    double A8[8][3], S8[8];
    for (int s = 0; s < 8; ++s) {
      const double pert = 1 + 0.05 * U(-1, 1);
      if (dir == 0) { A8[s][0] = 0.8 * pert; A8[s][1] = 0.2 * U(-1, 1); A8[s][2] = flat2d ? 0.0 : 0.2 * U(-1, 1); }
      else if (dir == 1) { A8[s][0] = 0.3 * U(-1, 1); A8[s][1] = 0.9 * pert; A8[s][2] = flat2d ? 0.0 : 0.1 * U(-1, 1); }
      else { A8[s][0] = 0; A8[s][1] = 0; A8[s][2] = 0.7 * pert; }
      S8[s] = std::sqrt(A8[s][0] * A8[s][0] + A8[s][1] * A8[s][1] + A8[s][2] * A8[s][2]);
    }
    double Q[8][5];
    for (int s = 0; s < 8; ++s) {
      const double jump = s >= 4 ? roughness : 0.0;
      Q[s][0] = 1.0 + 0.05 * U(-1, 1) + jump;
      Q[s][1] = uinf + 0.1 * uinf * U(-1, 1) - 0.5 * uinf * jump;
      Q[s][2] = flat2d ? 0.0 : 10 * U(-1, 1);
      Q[s][3] = flat2d ? 0.0 : 10 * U(-1, 1);
      Q[s][4] = 300 + 10 * U(-1, 1) + 200 * jump;
      f3d::sw_split<double>(kp, Q[s][0], Q[s][1], Q[s][2], Q[s][3], Q[s][4], A8[s][0], A8[s][1], A8[s][2], S8[s],
                            fp[s], fm[s]);
      for (int q = 0; q < 4; ++q) qs[s][q] = Q[s][q + 1];
    }
    const double* A = A8[3];
    const double* AR_ = A8[4];
    const double S = S8[3], SR_ = S8[4];
    const double* qL = qs[3];
    const double* qR = qs[4];
    double gF[5];
    for (int n = 0; n < 5; ++n) gF[n] = U(-1, 1);
    double gfp[8][5] = {}, gfm[8][5] = {}, gqL[4] = {}, gqR[4] = {};
    f3d::face_flux_vjp(kp, sch, fp, fm, qL, qR, A, AR_, S, SR_, gF, gfp, gfm, gqL, gqR);
    // forward mode over all 88 inputs
    for (int i = 0; i < 88; ++i) {
      Dual FPd[8][5], FMd[8][5], qLd[4], qRd[4], Ad[3] = {Dual(A[0]), Dual(A[1]), Dual(A[2])};
      Dual ARd[3] = {Dual(AR_[0]), Dual(AR_[1]), Dual(AR_[2])};
      for (int s = 0; s < 8; ++s)
        for (int n = 0; n < 5; ++n) {
          FPd[s][n] = Dual(fp[s][n], i == s * 5 + n ? 1.0 : 0.0);
          FMd[s][n] = Dual(fm[s][n], i == 40 + s * 5 + n ? 1.0 : 0.0);
        }
      for (int q = 0; q < 4; ++q) {
        qLd[q] = Dual(qL[q], i == 80 + q ? 1.0 : 0.0);
        qRd[q] = Dual(qR[q], i == 84 + q ? 1.0 : 0.0);
      }
      auto FP = [&](int s, int n) { return FPd[s + 4][n]; };
      auto FM = [&](int s, int n) { return FMd[s + 4][n]; };
      Dual out[5];
      f3d::face_flux<Dual>(kp, sch, FP, FM, qLd, qRd, Ad, ARd, Dual(S), Dual(SR_), out);
      double ref = 0;
      for (int n = 0; n < 5; ++n) ref += gF[n] * out[n].d;
      double got;
      double xs;  // input scale, so that the error is in output units
      if (i < 40) { got = gfp[i / 5][i % 5]; xs = std::fabs(fp[i / 5][i % 5]) + 1e-3; }
      else if (i < 80) { got = gfm[(i - 40) / 5][(i - 40) % 5]; xs = std::fabs(fm[(i - 40) / 5][(i - 40) % 5]) + 1e-3; }
      else if (i < 84) { got = gqL[i - 80]; xs = std::fabs(qL[i - 80]) + 1; }
      else { got = gqR[i - 84]; xs = std::fabs(qR[i - 84]) + 1; }
      worst = std::fmax(worst, std::fabs(ref - got) * xs);
      wscale = std::fmax(wscale, std::fabs(ref) * xs);
    }
  }
  report(name, worst, wscale, 1e-10);
}

static void check_viscous(const f3d::KP& kp) {
  double worst = 0, wscale = 0;
  for (int trial = 0; trial < 30; ++trial) {
    const double val[4] = {U(-100, 100), U(-100, 100), U(-100, 100), U(100, 600)};
    double g[4][3], m[9];
    for (int q = 0; q < 4; ++q)
      for (int d = 0; d < 3; ++d) g[q][d] = U(-10, 10);
    for (int q = 0; q < 9; ++q) m[q] = U(-1e-2, 1e-2);
    const double Jac = U(1e4, 1e6);
    double go[3][4];
    for (int d = 0; d < 3; ++d)
      for (int n = 0; n < 4; ++n) go[d][n] = U(-1, 1);
    double gg[4][3], gval[4];
    f3d::viscous_cell_vjp(kp, val[0], val[1], val[2], val[3], g, m, Jac, go, gg, gval);
    for (int i = 0; i < 16; ++i) {
      Dual vd[4], gd[4][3], md[9];
      for (int q = 0; q < 4; ++q) vd[q] = Dual(val[q], i == 12 + q ? 1.0 : 0.0);
      for (int q = 0; q < 4; ++q)
        for (int d = 0; d < 3; ++d) gd[q][d] = Dual(g[q][d], i == q * 3 + d ? 1.0 : 0.0);
      for (int q = 0; q < 9; ++q) md[q] = Dual(m[q]);
      Dual out[3][4];
      f3d::viscous_cell<Dual>(kp, vd[0], vd[1], vd[2], vd[3], gd, md, Dual(Jac), out);
      double ref = 0;
      for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 4; ++n) ref += go[d][n] * out[d][n].d;
      const double got = i < 12 ? gg[i / 3][i % 3] : gval[i - 12];
      const double xs = i < 12 ? std::fabs(g[i / 3][i % 3]) : std::fabs(val[i - 12]);
      worst = std::fmax(worst, std::fabs(ref - got) * xs);
      wscale = std::fmax(wscale, std::fabs(ref) * xs);
    }
  }
  report("viscous_cell", worst, wscale, 1e-12);
}

// forward c2Prim of one cell (utils.jl:2-30) in Dual, for the check only
static void c2prim_dual(const f3d::KP& kp, const Dual* Uc, Dual* Q) {
  const double eps = 2.220446049250313e-16;
  Dual rho = (Dual(eps) > Uc[0]) ? Dual(eps) : Uc[0];
  Dual ri = 1.0 / rho;
  Dual u = Uc[1] * ri, v = Uc[2] * ri, w = Uc[3] * ri;
  Dual er = Uc[4] - 0.5 * rho * (u * u + v * v + w * w);
  Dual ei = (Dual(eps) > er) ? Dual(eps) : er;
  Dual p = (kp.gamma - 1) * ei;
  Q[0] = rho; Q[1] = u; Q[2] = v; Q[3] = w; Q[4] = p; Q[5] = p / (rho * kp.Rg);
}

static void check_c2prim(const f3d::KP& kp) {
  double worst = 0, wscale = 0;
  for (int trial = 0; trial < 50; ++trial) {
    const double rho = U(0.1, 3), u = U(-500, 500), v = U(-500, 500), w = U(-500, 500), p = U(1e4, 2e5);
    const double Uc[5] = {rho, rho * u, rho * v, rho * w, p / (kp.gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)};
    double gQ[6];
    for (int n = 0; n < 6; ++n) gQ[n] = U(-1, 1);
    double gU[5];
    f3d::c2prim_vjp(kp, Uc, gQ, gU);
    for (int i = 0; i < 5; ++i) {
      Dual Ud[5], Q[6];
      for (int q = 0; q < 5; ++q) Ud[q] = Dual(Uc[q], q == i ? 1.0 : 0.0);
      c2prim_dual(kp, Ud, Q);
      double ref = 0, sc = 0;
      for (int n = 0; n < 6; ++n) {
        ref += gQ[n] * Q[n].d;
        sc += std::fabs(gQ[n] * Q[n].d);
      }
      worst = std::fmax(worst, std::fabs(ref - gU[i]) * std::fabs(Uc[i]));
      wscale = std::fmax(wscale, sc * std::fabs(Uc[i]));
    }
  }
  report("c2prim", worst, wscale, 1e-12);
}

int main() {
  const f3d::KP kp = make_kp_local();
  check_split(kp, 0.5);
  check_split(kp, 3.0);
  check_face(kp, F3D_SCHEME_UP7, "face UP7", 0.0);
  check_face(kp, F3D_SCHEME_WENO7, "face WENO7 (smooth)", 0.0);
  check_face(kp, F3D_SCHEME_WENO7, "face WENO7 (jump)", 0.5);
  check_face(kp, F3D_SCHEME_WENO5, "face WENO5 (smooth)", 0.0);
  check_face(kp, F3D_SCHEME_WENO5, "face WENO5 (jump)", 0.5);
  check_face(kp, F3D_SCHEME_NND2, "face NND", 0.5);
  const char* dn[3] = {"x", "y", "z"};
  char buf[96];
  for (int d = 0; d < 3; ++d)
    for (int flat = 0; flat < 2; ++flat)
      for (int sch = 1; sch <= 3; ++sch) {
        std::snprintf(buf, sizeof buf, "face sch%d %s-normal%s M3", sch, dn[d], flat ? " 2d" : "");
        check_face(kp, sch, buf, 0.5, d, flat != 0, 1000.0);
      }
  check_viscous(kp);
  check_c2prim(kp);
  std::printf("%s\n", failures ? "FAIL" : "OK");
  return failures ? 1 : 0;
}
