// extern "C" entry points of libf3d_hip.so: argument checking, error reporting and dispatch to
// the launchers.  See include/f3d.h for the contract of every function.
#include <cstdarg>
#include <cstring>

#include "f3d_internal.hpp"

namespace f3d {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* last_error() { return g_err; }

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return (int)e;
  }
  return F3D_OK;
}

KP make_kp(const f3d_params* p) {
  KP k;
  std::memset(&k, 0, sizeof(k));
  k.Nxp = p->Nxp;
  k.Nyp = p->Nyp;
  k.Nzp = p->Nzp;
  k.NG = p->NG;
  k.Nxt = p->Nxp + 2 * p->NG;
  k.Nyt = p->Nyp + 2 * p->NG;
  k.Nzt = p->Nzp + 2 * p->NG;
  k.viscous = p->viscous;
  k.fvm = p->fvm;
  k.split = p->split;
  k.character = p->character;
  k.les = p->les;
  k.lts = p->lts;
  for (int i = 0; i < 6; ++i) k.bc[i] = p->bc[i];
  k.nc = (long long)k.Nxt * k.Nyt * k.Nzt;
  k.gamma = p->gamma;
  k.Rg = p->Rg;
  k.Cp = p->Cp;
  k.Cs = p->C_s;
  k.Ts = p->T_s;
  k.Pr = p->Pr;
  k.phi1 = p->hybrid_phi1;
  k.phi2 = p->hybrid_phi2;
  k.phi3 = p->hybrid_phi3;
  k.eps7 = p->weno7_eps;
  k.eps5 = p->weno5_eps;
  k.Twall = p->T_wall;
  k.lts_cfl = p->lts_cfl;
  // Linear = UP7*Linear_phi + CD6*(1-Linear_phi)  run.jl:75-77
  const double up7[7] = {-3.0 / 420, 25.0 / 420, -101.0 / 420, 319.0 / 420, 214.0 / 420, -38.0 / 420, 4.0 / 420};
  const double cd6[7] = {0.0, 1.0 / 60, -2.0 / 15, 37.0 / 60, 37.0 / 60, -2.0 / 15, 1.0 / 60};
  for (int m = 0; m < 7; ++m) k.lin[m] = up7[m] * p->Linear_phi + cd6[m] * (1.0 - p->Linear_phi);
  auto c = [&](int q) { return (q >= 0 && q <= 6) ? k.lin[q] : 0.0; };
  for (int r = -4; r <= 3; ++r) k.da[r + 4] = c(r + 4) - c(r + 3);
  for (int r = -3; r <= 4; ++r) k.db[r + 3] = c(3 - r) - c(4 - r);
  for (int m = 0; m < 7; ++m) k.linf[m] = (float)k.lin[m];
  for (int m = 0; m < 8; ++m) {
    k.daf[m] = (float)k.da[m];
    k.dbf[m] = (float)k.db[m];
  }
  for (int i = 0; i < 6; ++i) k.qinf[i] = p->q_inf[i];
  k.gRg = p->gamma * p->Rg;
  k.gm1 = p->gamma - 1;
  k.w2c = (3 - p->gamma) / (2 * (p->gamma - 1));
  k.inv2g = 1 / (2 * p->gamma);
  const float gf = (float)p->gamma, rf = (float)p->Rg;
  k.gamf = gf;
  k.gRgf = gf * rf;
  k.gm1f = gf - 1;
  k.w2cf = (3 - gf) / (2 * (gf - 1));
  k.inv2gf = 1 / (2 * gf);
  return k;
}

}  // namespace f3d

using namespace f3d;

#define F3D_API extern "C" __attribute__((visibility("default")))

F3D_API int f3d_version(void) { return F3D_VERSION; }
F3D_API const char* f3d_last_error(void) { return last_error(); }

F3D_API void f3d_params_init(f3d_params* p, int32_t Nxp, int32_t Nyp, int32_t Nzp) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->Nxp = Nxp;
  p->Nyp = Nyp;
  p->Nzp = Nzp;
  p->NG = 4;
  p->viscous = 1;
  // boundary set of boundary.jl: inlet Dirichlet, outlet zero-gradient, adiabatic wall below,
  // Dirichlet farfield above, periodic span (Iperiodic z, run.jl:19)
  p->bc[0] = F3D_BC_DIRICHLET;
  p->bc[1] = F3D_BC_ZEROGRAD;
  p->bc[2] = F3D_BC_WALL_ADIABATIC;
  p->bc[3] = F3D_BC_DIRICHLET;
  p->bc[4] = F3D_BC_PERIODIC;
  p->bc[5] = F3D_BC_PERIODIC;
  p->gamma = 1.4;
  p->Rg = 287.0;
  p->Cp = p->Rg * p->gamma / (p->gamma - 1);
  p->C_s = 1.458e-6;
  p->T_s = 110.4;
  p->Pr = 0.72;
  p->hybrid_phi1 = 0.05;
  p->hybrid_phi2 = 1.0;
  p->hybrid_phi3 = 10.0;
  p->Linear_phi = 1.0;
  p->weno7_eps = 1e-20;
  p->weno5_eps = 1e-16;
  // flag-selected alternatives, run.jl:4-5, 21-22, 50-54, 68-69
  p->split = F3D_SPLIT_SW;
  p->character = 1;
  p->les = F3D_LES_NONE;
  p->lts = 0;
  p->lts_cfl = 0.5;
  p->filtering = 0;
  p->filtering_nonlinear = 0;
  p->filtering_interval = 100;
  p->filtering_rth = 1e-5;
  p->filtering_s0 = 1.0;
  // freestream of boundary.jl:14-19
  p->q_inf[0] = 1.0;
  p->q_inf[1] = 0.5;
  p->q_inf[2] = 0.0;
  p->q_inf[3] = 0.0;
  p->q_inf[4] = 1.0 / 1.4;
  p->q_inf[5] = p->q_inf[4] / (p->q_inf[0] * p->Rg);
  p->T_wall = 0.0;
}

F3D_API int f3d_params_check(const f3d_params* p) {
  if (!p) {
    set_error("params: NULL");
    return F3D_E_NULL;
  }
  if (p->NG != 4) {
    set_error("params: NG=%d, only NG=4 is supported (viscous.jl:119, div.jl:14-39 assume it)", p->NG);
    return F3D_E_NG;
  }
  if (p->Nxp < 4 || p->Nyp < 4 || p->Nzp < 4) {
    set_error("params: interior dims (%d,%d,%d) must each be >= NG", p->Nxp, p->Nyp, p->Nzp);
    return F3D_E_DIMS;
  }
  const long long nc = (long long)(p->Nxp + 8) * (p->Nyp + 8) * (p->Nzp + 8);
  if (nc * 6 >= (1LL << 40)) {
    set_error("params: grid too large");
    return F3D_E_DIMS;
  }
  for (int i = 0; i < 6; ++i) {
    const int b = p->bc[i];
    if (b < F3D_BC_PERIODIC || b > F3D_BC_INFLOW) {
      set_error("params: bc[%d]=%d unknown", i, b);
      return F3D_E_ARG;
    }
    if (b == F3D_BC_INFLOW && i != 0) {
      set_error("params: F3D_BC_INFLOW is only defined on the x-low face");
      return F3D_E_ARG;
    }
  }
  if (p->split < F3D_SPLIT_SW || p->split > F3D_SPLIT_AUSM) {
    set_error("params: split=%d unknown (F3D_SPLIT_SW, _LF, _VL, _AUSM)", p->split);
    return F3D_E_ARG;
  }
  if ((p->character != 0 && p->character != 1) || (p->lts != 0 && p->lts != 1) ||
      (p->filtering != 0 && p->filtering != 1) || (p->filtering_nonlinear != 0 && p->filtering_nonlinear != 1)) {
    set_error("params: character, lts, filtering and filtering_nonlinear are 0 or 1");
    return F3D_E_ARG;
  }
  if (p->les < F3D_LES_NONE || p->les > F3D_LES_WALE) {
    set_error("params: les=%d unknown (F3D_LES_NONE, _SMAG, _WALE)", p->les);
    return F3D_E_ARG;
  }
  if (p->lts && !(p->lts_cfl > 0)) {
    set_error("params: local time stepping needs lts_cfl > 0");
    return F3D_E_ARG;
  }
  if (p->filtering && (p->filtering_interval < 1 || !(p->filtering_s0 >= 0))) {
    set_error("params: filtering needs filtering_interval >= 1 and filtering_s0 >= 0");
    return F3D_E_ARG;
  }
  if (p->fvm < F3D_FDM || p->fvm > F3D_FVM_CONSER) {
    set_error("params: fvm=%d unknown (F3D_FDM, F3D_FVM_EIGEN or F3D_FVM_CONSER)", p->fvm);
    return F3D_E_ARG;
  }
  for (int d = 0; d < 3; ++d)
    if ((p->bc[2 * d] == F3D_BC_PERIODIC) != (p->bc[2 * d + 1] == F3D_BC_PERIODIC)) {
      set_error("params: periodic faces must come in pairs (dimension %d)", d);
      return F3D_E_ARG;
    }
  if (!(p->gamma > 1) || !(p->Rg > 0)) {
    set_error("params: gamma must be > 1 and Rg > 0");
    return F3D_E_ARG;
  }
  if ((p->bc[2] == F3D_BC_WALL_ISOTHERMAL || p->bc[3] == F3D_BC_WALL_ISOTHERMAL || p->bc[0] == F3D_BC_WALL_ISOTHERMAL ||
       p->bc[1] == F3D_BC_WALL_ISOTHERMAL || p->bc[4] == F3D_BC_WALL_ISOTHERMAL ||
       p->bc[5] == F3D_BC_WALL_ISOTHERMAL) &&
      !(p->T_wall > 0)) {
    set_error("params: isothermal wall needs T_wall > 0");
    return F3D_E_ARG;
  }
  return F3D_OK;
}

F3D_API size_t f3d_cell_count(const f3d_params* p) {
  if (!p) return 0;
  return (size_t)(p->Nxp + 2 * p->NG) * (p->Nyp + 2 * p->NG) * (p->Nzp + 2 * p->NG);
}

#define CHECK_P()                          \
  do {                                     \
    int rc_ = f3d_params_check(p);         \
    if (rc_) return rc_;                   \
  } while (0)
#define NEED(x)                                               \
  do {                                                        \
    if (!(x)) {                                               \
      set_error("%s: argument %s is NULL", __func__, #x);     \
      return F3D_E_NULL;                                      \
    }                                                         \
  } while (0)

// ---------------------------------------------------------------- sensor / mask / split
#define SENSOR_IMPL(SUF, T)                                                                      \
  F3D_API int f3d_shock_sensor_##SUF(const f3d_params* p, T* phi, const T* Q, f3d_stream_t s) { \
    CHECK_P();                                                                                   \
    NEED(phi);                                                                                   \
    NEED(Q);                                                                                     \
    return launch_shock_sensor<T>(make_kp(p), phi, Q, (hipStream_t)s);                           \
  }
SENSOR_IMPL(f64, double)
SENSOR_IMPL(f32b, float)

#define MASK_IMPL(SUF, T)                                                                                \
  F3D_API int f3d_scheme_mask_##SUF(const f3d_params* p, uint8_t* mx, uint8_t* my, uint8_t* mz,         \
                                    const T* phi, f3d_stream_t s) {                                      \
    CHECK_P();                                                                                           \
    NEED(phi);                                                                                           \
    return launch_scheme_mask<T>(make_kp(p), mx, my, mz, phi, (hipStream_t)s);                           \
  }
MASK_IMPL(f64, double)
MASK_IMPL(f32b, float)

// fluxSplit_SW / _LF / _VL / _AUSM  split.jl:2 / 70 / 110 / 185 (the method is the entry point's)
#define SPLIT_IMPL(NAME, METHOD, SUF, T)                                                                       \
  F3D_API int f3d_flux_split_##NAME##_##SUF(const f3d_params* p, const T* Q, T* Fp, T* Fm, const T* S,         \
                                            const T* Ax, const T* Ay, const T* Az, f3d_stream_t s) {           \
    CHECK_P();                                                                                                 \
    NEED(Q);                                                                                                   \
    NEED(Fp);                                                                                                  \
    NEED(Fm);                                                                                                  \
    NEED(S);                                                                                                   \
    NEED(Ax);                                                                                                  \
    NEED(Ay);                                                                                                  \
    NEED(Az);                                                                                                  \
    KP kp = make_kp(p);                                                                                        \
    kp.split = METHOD;                                                                                         \
    return launch_flux_split<T>(kp, Q, Fp, Fm, S, Ax, Ay, Az, (hipStream_t)s);                                 \
  }
SPLIT_IMPL(sw, F3D_SPLIT_SW, f64, double)
SPLIT_IMPL(sw, F3D_SPLIT_SW, f32b, float)
SPLIT_IMPL(lf, F3D_SPLIT_LF, f64, double)
SPLIT_IMPL(lf, F3D_SPLIT_LF, f32b, float)
SPLIT_IMPL(vl, F3D_SPLIT_VL, f64, double)
SPLIT_IMPL(vl, F3D_SPLIT_VL, f32b, float)
SPLIT_IMPL(ausm, F3D_SPLIT_AUSM, f64, double)
SPLIT_IMPL(ausm, F3D_SPLIT_AUSM, f32b, float)

#define ADVECT_IMPL(NAME, D, SUF, T)                                                                         \
  F3D_API int f3d_advect_##NAME##_##SUF(const f3d_params* p, T* F, const T* phi, const T* S, const T* Fp,    \
                                        const T* Fm, const T* Q, const T* Ax, const T* Ay, const T* Az,      \
                                        f3d_stream_t s) {                                                    \
    CHECK_P();                                                                                               \
    NEED(F);                                                                                                 \
    NEED(phi);                                                                                               \
    NEED(S);                                                                                                 \
    NEED(Fp);                                                                                                \
    NEED(Fm);                                                                                                \
    NEED(Q);                                                                                                 \
    NEED(Ax);                                                                                                \
    NEED(Ay);                                                                                                \
    NEED(Az);                                                                                                \
    KP kp = make_kp(p);                                                                                      \
    kp.character = 1;                                                                                        \
    return launch_advect<T>(kp, D, F, phi, S, Fp, Fm, Q, Ax, Ay, Az, (hipStream_t)s);                        \
  }
ADVECT_IMPL(xc, 0, f64, double)
ADVECT_IMPL(yc, 1, f64, double)
ADVECT_IMPL(zc, 2, f64, double)
ADVECT_IMPL(xc, 0, f32b, float)
ADVECT_IMPL(yc, 1, f32b, float)
ADVECT_IMPL(zc, 2, f32b, float)

// advect_x / _y / _z  schemes.jl:86 / 291 / 496: component-wise, the first NV split-flux components
#define CADVECT_IMPL(NAME, D, SUF, T)                                                                         \
  F3D_API int f3d_advect_##NAME##_##SUF(const f3d_params* p, T* F, const T* phi, const T* S, const T* Fp,     \
                                        const T* Fm, int32_t NV, f3d_stream_t s) {                            \
    CHECK_P();                                                                                                \
    NEED(F);                                                                                                  \
    NEED(phi);                                                                                                \
    NEED(S);                                                                                                  \
    NEED(Fp);                                                                                                 \
    NEED(Fm);                                                                                                 \
    if (NV < 1 || NV > 5) {                                                                                   \
      set_error("advect_" #NAME ": NV=%d out of 1..5", NV);                                                   \
      return F3D_E_ARG;                                                                                       \
    }                                                                                                         \
    KP kp = make_kp(p);                                                                                       \
    kp.character = 0;                                                                                         \
    return launch_advect<T>(kp, D, F, phi, S, Fp, Fm, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, \
                            (const T*)nullptr, (hipStream_t)s, NV);                                           \
  }
CADVECT_IMPL(x, 0, f64, double)
CADVECT_IMPL(y, 1, f64, double)
CADVECT_IMPL(z, 2, f64, double)
CADVECT_IMPL(x, 0, f32b, float)
CADVECT_IMPL(y, 1, f32b, float)
CADVECT_IMPL(z, 2, f32b, float)

#define VISC_IMPL(SUF, T)                                                                                      \
  F3D_API int f3d_viscous_flux_##SUF(const f3d_params* p, T* Fvx, T* Fvy, T* Fvz, const T* Q, const T* m0,     \
                                     const T* m1, const T* m2, const T* m3, const T* m4, const T* m5,          \
                                     const T* m6, const T* m7, const T* m8, const double* J, f3d_stream_t s) { \
    CHECK_P();                                                                                                 \
    NEED(Fvx);                                                                                                 \
    NEED(Fvy);                                                                                                 \
    NEED(Fvz);                                                                                                 \
    NEED(Q);                                                                                                   \
    NEED(J);                                                                                                   \
    MetricPtrs<T> m{{m0, m1, m2, m3, m4, m5, m6, m7, m8}};                                                     \
    for (int q = 0; q < 9; ++q)                                                                                \
      if (!m.m[q]) {                                                                                           \
        set_error("viscous_flux: metric %d is NULL", q);                                                       \
        return F3D_E_NULL;                                                                                     \
      }                                                                                                        \
    return launch_viscous<T>(make_kp(p), Fvx, Fvy, Fvz, Q, m, J, (hipStream_t)s);                              \
  }
VISC_IMPL(f64, double)
VISC_IMPL(f32b, float)

#define DIV_IMPL(SUF, T)                                                                                     \
  F3D_API int f3d_div_##SUF(const f3d_params* p, double* U, const T* Fx, const T* Fy, const T* Fz,           \
                            const T* Fvx, const T* Fvy, const T* Fvz, double dt, const double* J,            \
                            f3d_stream_t s) {                                                                \
    CHECK_P();                                                                                               \
    NEED(U);                                                                                                 \
    NEED(Fx);                                                                                                \
    NEED(Fy);                                                                                                \
    NEED(Fz);                                                                                                \
    NEED(Fvx);                                                                                               \
    NEED(Fvy);                                                                                               \
    NEED(Fvz);                                                                                               \
    NEED(J);                                                                                                 \
    return launch_div<T>(make_kp(p), U, Fx, Fy, Fz, Fvx, Fvy, Fvz, dt, J, (hipStream_t)s);                   \
  }
DIV_IMPL(f64, double)
DIV_IMPL(f32b, float)

// div_LTS  div.jl:52-101 (per-cell dt) and compute_dt  utils.jl:73-100
#define LTS_IMPL(SUF, T)                                                                                     \
  F3D_API int f3d_div_lts_##SUF(const f3d_params* p, double* U, const T* Fx, const T* Fy, const T* Fz,       \
                                const T* Fvx, const T* Fvy, const T* Fvz, const double* dt, const double* J, \
                                f3d_stream_t s) {                                                            \
    CHECK_P();                                                                                               \
    NEED(U);                                                                                                 \
    NEED(Fx);                                                                                                \
    NEED(Fy);                                                                                                \
    NEED(Fz);                                                                                                \
    NEED(Fvx);                                                                                               \
    NEED(Fvy);                                                                                               \
    NEED(Fvz);                                                                                               \
    NEED(dt);                                                                                                \
    NEED(J);                                                                                                 \
    return launch_div<T>(make_kp(p), U, Fx, Fy, Fz, Fvx, Fvy, Fvz, 0.0, J, (hipStream_t)s, dt);              \
  }                                                                                                          \
  F3D_API int f3d_compute_dt_##SUF(const f3d_params* p, double* dt, const T* Q, const double* J, const T* S1, \
                                   const T* S2, const T* S3, f3d_stream_t s) {                               \
    CHECK_P();                                                                                               \
    NEED(dt);                                                                                                \
    NEED(Q);                                                                                                 \
    NEED(J);                                                                                                 \
    NEED(S1);                                                                                                \
    NEED(S2);                                                                                                \
    NEED(S3);                                                                                                \
    return launch_compute_dt<T>(make_kp(p), dt, Q, J, S1, S2, S3, nullptr, (hipStream_t)s);                  \
  }
LTS_IMPL(f64, double)
LTS_IMPL(f32b, float)

// pre_x/_y/_z  utils.jl:102-169 (sc cell-shaped double), filter_x/_y/_z 171-241, linearFilter_x/_y/_z 243-316
#define PRE_IMPL(X, D, SUF, T)                                                                                \
  F3D_API int f3d_pre_##X##_##SUF(const f3d_params* p, double* sc, const T* Q, double rth, f3d_stream_t s) {  \
    CHECK_P();                                                                                                \
    NEED(sc);                                                                                                 \
    NEED(Q);                                                                                                  \
    return launch_pre_sensor<T>(make_kp(p), D, sc, Q, rth, (hipStream_t)s);                                   \
  }
PRE_IMPL(x, 0, f64, double)
PRE_IMPL(y, 1, f64, double)
PRE_IMPL(z, 2, f64, double)
PRE_IMPL(x, 0, f32b, float)
PRE_IMPL(y, 1, f32b, float)
PRE_IMPL(z, 2, f32b, float)
#define FILTER_IMPL(X, D)                                                                                     \
  F3D_API int f3d_filter_##X(const f3d_params* p, double* U, const double* Un, const double* sc, double s0,   \
                             f3d_stream_t s) {                                                                \
    CHECK_P();                                                                                                \
    NEED(U);                                                                                                  \
    NEED(Un);                                                                                                 \
    NEED(sc);                                                                                                 \
    return launch_filter(make_kp(p), D, U, Un, sc, s0, (hipStream_t)s);                                       \
  }                                                                                                           \
  F3D_API int f3d_linear_filter_##X(const f3d_params* p, double* U, const double* Un, double s0,              \
                                    f3d_stream_t s) {                                                         \
    CHECK_P();                                                                                                \
    NEED(U);                                                                                                  \
    NEED(Un);                                                                                                 \
    return launch_filter(make_kp(p), D, U, Un, nullptr, s0, (hipStream_t)s);                                  \
  }
FILTER_IMPL(x, 0)
FILTER_IMPL(y, 1)
FILTER_IMPL(z, 2)

F3D_API int f3d_lin_comb(const f3d_params* p, double* U, const double* Un, int32_t NV, double a, double b,
                         f3d_stream_t s) {
  CHECK_P();
  NEED(U);
  NEED(Un);
  if (NV < 1 || NV > 6) {
    set_error("lin_comb: NV=%d out of range", NV);
    return F3D_E_ARG;
  }
  return launch_lin_comb(make_kp(p), U, Un, NV, a, b, (hipStream_t)s);
}

#define CONV_IMPL(SUF, T)                                                                           \
  F3D_API int f3d_c2prim_##SUF(const f3d_params* p, const double* U, T* Q, f3d_stream_t s) {        \
    CHECK_P();                                                                                      \
    NEED(U);                                                                                        \
    NEED(Q);                                                                                        \
    return launch_c2prim<T>(make_kp(p), U, Q, (hipStream_t)s);                                      \
  }                                                                                                 \
  F3D_API int f3d_prim2c_##SUF(const f3d_params* p, double* U, const T* Q, f3d_stream_t s) {        \
    CHECK_P();                                                                                      \
    NEED(U);                                                                                        \
    NEED(Q);                                                                                        \
    return launch_prim2c<T>(make_kp(p), U, Q, (hipStream_t)s);                                      \
  }                                                                                                 \
  F3D_API int f3d_exchange_ghost_##SUF(const f3d_params* p, T* Q, int32_t NV, f3d_stream_t s) {     \
    CHECK_P();                                                                                      \
    NEED(Q);                                                                                        \
    if (NV < 1 || NV > 6) {                                                                         \
      set_error("exchange_ghost: NV=%d out of range", NV);                                          \
      return F3D_E_ARG;                                                                             \
    }                                                                                               \
    return launch_exchange<T>(make_kp(p), Q, NV, (hipStream_t)s);                                   \
  }                                                                                                 \
  F3D_API int f3d_fill_ghost_##SUF(const f3d_params* p, T* Q, double* U, const T* inflow,           \
                                   f3d_stream_t s) {                                                \
    CHECK_P();                                                                                      \
    NEED(Q);                                                                                        \
    if (p->bc[0] == F3D_BC_INFLOW && !inflow) {                                                     \
      set_error("fill_ghost: F3D_BC_INFLOW needs an inflow profile");                               \
      return F3D_E_NULL;                                                                            \
    }                                                                                               \
    return launch_fill<T>(make_kp(p), Q, U, inflow, (hipStream_t)s);                                \
  }
CONV_IMPL(f64, double)
CONV_IMPL(f32b, float)

// flowAdvance  solver.jl:19-188: finite-volume branch (solver.jl:21-31) or finite-difference branch
// (solver.jl:133-184), then viscousFlux (solver.jl:185-187).
template <typename T>
static int flow_advance_impl(const f3d_params* p, const double* U, T* Q, T* Fp, T* Fm, T* Fx, T* Fy, T* Fz, T* Fvx,
                             T* Fvy, T* Fvz, const T* s1, const T* s2, const T* s3, const MetricPtrs<T>& m,
                             const double* J, const double* x, const double* y, const double* z, T* phi,
                             hipStream_t s) {
  KP kp = make_kp(p);
  const T* S[3] = {s1, s2, s3};
  T* F[3] = {Fx, Fy, Fz};
  for (int d = 0; d < 3; ++d) {
    int rc;
    if (kp.fvm) {
      rc = launch_fvm_face<T>(kp, d, Q, U, phi, S[d], nullptr, nullptr, nullptr, F[d], x, y, z, nullptr, s);
    } else {
      rc = launch_flux_split<T>(kp, Q, Fp, Fm, S[d], m.m[3 * d], m.m[3 * d + 1], m.m[3 * d + 2], s);
      if (!rc)
        rc = launch_advect<T>(kp, d, F[d], phi, S[d], Fp, Fm, Q, m.m[3 * d], m.m[3 * d + 1], m.m[3 * d + 2], s);
    }
    if (rc) return rc;
  }
  if (p->viscous) return launch_viscous<T>(kp, Fvx, Fvy, Fvz, Q, m, J, s);
  return F3D_OK;
}

#define FA_IMPL(SUF, T)                                                                                        \
  F3D_API int f3d_flow_advance_##SUF(                                                                          \
      const f3d_params* p, double* U, T* Q, T* Fp, T* Fm, T* Fx, T* Fy, T* Fz, T* Fvx, T* Fvy, T* Fvz,          \
      const T* s1, const T* s2, const T* s3, const T* m0, const T* m1, const T* m2, const T* m3, const T* m4,   \
      const T* m5, const T* m6, const T* m7, const T* m8, const double* J, const double* x, const double* y,    \
      const double* z, T* phi, f3d_stream_t s) {                                                               \
    CHECK_P();                                                                                                 \
    NEED(Q);                                                                                                   \
    NEED(Fx);                                                                                                  \
    NEED(Fy);                                                                                                  \
    NEED(Fz);                                                                                                  \
    NEED(s1);                                                                                                  \
    NEED(s2);                                                                                                  \
    NEED(s3);                                                                                                  \
    NEED(phi);                                                                                                 \
    if (p->fvm) {                                                                                              \
      NEED(U);                                                                                                 \
      NEED(x);                                                                                                 \
      NEED(y);                                                                                                 \
      NEED(z);                                                                                                 \
    } else {                                                                                                   \
      NEED(Fp);                                                                                                \
      NEED(Fm);                                                                                                \
    }                                                                                                          \
    MetricPtrs<T> m{{m0, m1, m2, m3, m4, m5, m6, m7, m8}};                                                     \
    for (int q = 0; q < 9; ++q)                                                                                \
      if (!m.m[q] && (!p->fvm || p->viscous)) {                                                                \
        set_error("flow_advance: metric %d is NULL", q);                                                       \
        return F3D_E_NULL;                                                                                     \
      }                                                                                                        \
    if (p->viscous && (!Fvx || !Fvy || !Fvz || !J)) {                                                          \
      set_error("flow_advance: viscous needs Fv_x, Fv_y, Fv_z and J");                                         \
      return F3D_E_NULL;                                                                                       \
    }                                                                                                          \
    return flow_advance_impl<T>(p, U, Q, Fp, Fm, Fx, Fy, Fz, Fvx, Fvy, Fvz, s1, s2, s3, m, J, x, y, z, phi,    \
                                (hipStream_t)s);                                                               \
  }
FA_IMPL(f64, double)
FA_IMPL(f32b, float)

// Eigen_reconstruct_{i,j,k} / Conser_reconstruct_{i,j,k}  FVM.jl:6-670 / 672-1222.  The variant is
// the entry point's, whatever p->fvm says.
#define FVM_IMPL(NAME, SUF, T, D, EIG)                                                                         \
  F3D_API int f3d_##NAME##_##SUF(const f3d_params* p, const T* Q, const double* U, const T* phi, const T* S,   \
                                 T* F, const double* x, const double* y, const double* z, f3d_stream_t s) {   \
    CHECK_P();                                                                                                 \
    NEED(Q);                                                                                                   \
    NEED(U);                                                                                                   \
    NEED(phi);                                                                                                 \
    NEED(S);                                                                                                   \
    NEED(F);                                                                                                   \
    NEED(x);                                                                                                   \
    NEED(y);                                                                                                   \
    NEED(z);                                                                                                   \
    KP kp = make_kp(p);                                                                                        \
    kp.fvm = EIG ? F3D_FVM_EIGEN : F3D_FVM_CONSER;                                                             \
    return launch_fvm_face<T>(kp, D, Q, U, phi, S, nullptr, nullptr, nullptr, F, x, y, z, nullptr,             \
                              (hipStream_t)s);                                                                 \
  }
// Face geometry arrays of the fused finite-volume entries (FVM.jl:16-31 per face, once per mesh).
#define GEO_IMPL(SUF, T)                                                                                     \
  F3D_API int f3d_fvm_geometry_##SUF(const f3d_params* p, const double* x, const double* y, const double* z, \
                                     T* const* geo, f3d_stream_t s) {                                        \
    CHECK_P();                                                                                               \
    NEED(x);                                                                                                 \
    NEED(y);                                                                                                 \
    NEED(z);                                                                                                 \
    NEED(geo);                                                                                               \
    for (int d = 0; d < 3; ++d)                                                                              \
      if (!geo[d]) {                                                                                         \
        set_error("fvm_geometry: geo[%d] is NULL", d);                                                       \
        return F3D_E_NULL;                                                                                   \
      }                                                                                                      \
    return launch_fvm_geometry<T>(make_kp(p), x, y, z, geo, (hipStream_t)s);                                 \
  }
GEO_IMPL(f64, double)
GEO_IMPL(f32b, float)

FVM_IMPL(eigen_reconstruct_i, f64, double, 0, 1)
FVM_IMPL(eigen_reconstruct_j, f64, double, 1, 1)
FVM_IMPL(eigen_reconstruct_k, f64, double, 2, 1)
FVM_IMPL(conser_reconstruct_i, f64, double, 0, 0)
FVM_IMPL(conser_reconstruct_j, f64, double, 1, 0)
FVM_IMPL(conser_reconstruct_k, f64, double, 2, 0)
FVM_IMPL(eigen_reconstruct_i, f32b, float, 0, 1)
FVM_IMPL(eigen_reconstruct_j, f32b, float, 1, 1)
FVM_IMPL(eigen_reconstruct_k, f32b, float, 2, 1)
FVM_IMPL(conser_reconstruct_i, f32b, float, 0, 0)
FVM_IMPL(conser_reconstruct_j, f32b, float, 1, 0)
FVM_IMPL(conser_reconstruct_k, f32b, float, 2, 0)
