/*
 * f3d.h - C-ABI of the MI355X (gfx950) Flame3D RHS + RK3 stage library (libf3d_hip.so).
 *
 * Replaces the finite-difference right-hand-side path of Flame3D.jl (rygan-code/Flame3D-MHD):
 * shock sensor, Steger-Warming flux splitting, hybrid UP7/WENO7/WENO5/NND characteristic
 * reconstruction, laminar viscous flux on J^-1-scaled curvilinear metrics, conservative flux
 * divergence, SSP-RK3 blending, conserved<->primitive conversion and ghost-cell filling.
 *
 * Array layout (identical to the reference's CuArrays, solver.jl:294-308): Julia column-major,
 * element (i,j,k,n) (0-based here) at offset  i + Nxt*(j + Nyt*(k + Nzt*n)),
 * Nxt = Nxp+2*NG etc.  Cell arrays carry NG=4 ghost layers per side.
 *   Q   [Nxt,Nyt,Nzt,6]  primitive (rho,u,v,w,p,T)          U, Un [Nxt,Nyt,Nzt,5] conserved
 *   phi, s1..s3, 9 metrics, J, x,y,z  [Nxt,Nyt,Nzt]
 *   Fp, Fm [Nxt,Nyt,Nzt,5]   Fx [Nxp+1,Nyp,Nzp,5]  Fy [Nxp,Nyp+1,Nzp,5]  Fz [Nxp,Nyp,Nzp+1,5]
 *   Fv_x, Fv_y, Fv_z [Nxp+NG,Nyp+NG,Nzp+NG,4]
 *   mask_x [Nxp+1,Nyp,Nzp], mask_y [Nxp,Nyp+1,Nzp], mask_z [Nxp,Nyp,Nzp+1]  (uint8)
 *
 * Two numeric forms, selected by the suffix:
 *   _f64  : every array is double.
 *   _f32b : "FP32-bulk": U, Un, J, dt and the RK3 accumulation are double; Q, phi, Fp, Fm, Fx..Fz,
 *           Fv_*, s1..s3 and the 9 metrics are float, and splitting, reconstruction and the viscous
 *           flux are evaluated in float with float accumulation.
 *
 * Every pointer argument is a device pointer unless stated otherwise.  Every call is asynchronous
 * on `stream` (NULL = the null stream); the caller synchronises.  Return value: 0 on success,
 * F3D_E_* (<0) for an invalid argument, or the positive hipError_t of a failed HIP call.
 * f3d_last_error() returns a thread-local message for the last failure.  Nothing here throws.
 * The library never frees caller memory; hot calls never allocate (workspaces are created once).
 */
#ifndef F3D_H
#define F3D_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* f3d_stream_t; /* == hipStream_t */

#define F3D_VERSION 10000

/* status codes */
#define F3D_OK 0
#define F3D_E_NULL (-1)     /* required pointer is NULL */
#define F3D_E_DIMS (-2)     /* bad Nxp/Nyp/Nzp */
#define F3D_E_NG (-3)       /* NG != 4 */
#define F3D_E_ARG (-4)      /* other invalid argument (BC kind, NV, KRK, precision ...) */
#define F3D_E_NAN (-5)      /* NaN/Inf detected by f3d_nan_check_* (solver.jl:481-486) */
#define F3D_E_WS (-6)       /* workspace does not match params / precision */
#define F3D_E_IO (-7)       /* file I/O failed (f3d_h5.h) */

/* boundary-condition kinds, one per domain face (xlo,xhi,ylo,yhi,zlo,zhi) */
#define F3D_BC_PERIODIC 0      /* single-rank exchange_ghost semantics (mpi.jl:2-103) + fill_z (boundary.jl:123-148) */
#define F3D_BC_DIRICHLET 1     /* freestream q_inf in every ghost layer (boundary.jl:12-38, 98-118) */
#define F3D_BC_ZEROGRAD 2      /* copy of the last interior layer (boundary.jl:42-51) */
#define F3D_BC_WALL_ADIABATIC 3 /* mirror, velocity negated, rho/p/T copied (boundary.jl:67-93) */
#define F3D_BC_WALL_ISOTHERMAL 4 /* mirror, velocity negated, p copied, T = 2*T_wall - T_mirror, rho = p/(Rg T) (no reference counterpart) */
#define F3D_BC_INFLOW 5        /* Dirichlet from a caller profile inflow[j,k,6] (xlo only; no reference counterpart) */

#define F3D_PREC_F64 0
#define F3D_PREC_F32B 1

/* Scheme selected per face by the shock-sensor mask (schemes.jl:731, 813, 931, 1000). */
#define F3D_SCHEME_UP7 0
#define F3D_SCHEME_WENO7 1
#define F3D_SCHEME_WENO5 2
#define F3D_SCHEME_NND2 3

/* Inviscid discretisation (run.jl:64-65 finite_volume / eigen_reconstruction; solver.jl:21-31).
 * F3D_FDM: Steger-Warming flux-vector splitting with hybrid characteristic reconstruction of the
 * split fluxes (schemes.jl, split.jl).  F3D_FVM_*: reconstruction of the conserved state U to both
 * sides of each face, HLLC flux times the face area from the node coordinates (FVM.jl), with the
 * same sensor thresholds (the non-UP7 branches use WENO epsilons 1e-10 / 1e-8, FVM.jl:91). */
/* Flux-vector splitting of the finite-difference branch (split.jl; solver.jl:134-144). */
#define F3D_SPLIT_SW 0    /* Steger-Warming, split.jl:2-68 */
#define F3D_SPLIT_LF 1    /* local Lax-Friedrichs, split.jl:70-108 */
#define F3D_SPLIT_VL 2    /* van Leer, split.jl:110-183 */
#define F3D_SPLIT_AUSM 3  /* AUSM, split.jl:185-255 */

/* Subgrid model of the viscous flux (viscous.jl:66-106). */
#define F3D_LES_NONE 0
#define F3D_LES_SMAG 1
#define F3D_LES_WALE 2

#define F3D_FDM 0
#define F3D_FVM_EIGEN 1   /* characteristic (Roe-frame) reconstruction, FVM.jl:6-670 */
#define F3D_FVM_CONSER 2  /* component-wise reconstruction of U, FVM.jl:672-1222 */

/* Runtime form of the reference's compile-time constants (run.jl:8-13, 57-77, 80-86). */
typedef struct f3d_params {
  int32_t Nxp, Nyp, Nzp;   /* interior cells (run.jl:84-86) */
  int32_t NG;              /* ghost layers; must be 4 (run.jl:80; viscous.jl:119 and div.jl:14-39 assume it) */
  int32_t viscous;         /* run.jl:61 */
  int32_t bc[6];           /* F3D_BC_* for xlo,xhi,ylo,yhi,zlo,zhi */
  int32_t fvm;             /* F3D_FDM / F3D_FVM_EIGEN / F3D_FVM_CONSER (run.jl:64-65) */
  double gamma, Rg, Cp, C_s, T_s, Pr;          /* run.jl:8-13 */
  double hybrid_phi1, hybrid_phi2, hybrid_phi3; /* run.jl:70-72 */
  double Linear_phi;                            /* run.jl:73: Linear = UP7*Linear_phi + CD6*(1-Linear_phi) */
  double weno7_eps, weno5_eps;                  /* schemes.jl:712-713 */
  double q_inf[6];   /* Dirichlet state (rho,u,v,w,p,T); reference values boundary.jl:14-19 */
  double T_wall;     /* F3D_BC_WALL_ISOTHERMAL wall temperature */
  /* flag-selected alternatives of run.jl (f3d_params_init sets the reference's values) */
  int32_t split;     /* F3D_SPLIT_* (run.jl:69 splitMethod) */
  int32_t character; /* 1: characteristic reconstruction advect_*c, 0: component-wise advect_* (run.jl:68) */
  int32_t les;       /* F3D_LES_NONE / F3D_LES_SMAG / F3D_LES_WALE (run.jl:4-5) */
  int32_t lts;       /* local time stepping (run.jl:21): per-cell dt of compute_dt, div_LTS */
  int32_t filtering, filtering_nonlinear, filtering_interval; /* run.jl:50-52 */
  int32_t pad_;
  double lts_cfl;                    /* run.jl:22 */
  double filtering_rth, filtering_s0; /* run.jl:53-54 */
} f3d_params;

int f3d_version(void);
const char* f3d_last_error(void);

/* Fill *p with the reference's run.jl constants and boundary set (inlet Dirichlet, outlet
 * zero-gradient, adiabatic wall below, Dirichlet farfield above, periodic span). Host-only. */
void f3d_params_init(f3d_params* p, int32_t Nxp, int32_t Nyp, int32_t Nzp);
/* Validate params; 0 or F3D_E_*. Host-only, no GPU call. */
int f3d_params_check(const f3d_params* p);
/* Sizes in elements of the arrays above (n_cell = Nxt*Nyt*Nzt, etc.). Host-only. */
size_t f3d_cell_count(const f3d_params* p);

/* ---------------- kernel-by-kernel drop-ins (the reference's per-stage launches) ---------- */

/* shockSensor(phi, Q)  schemes.jl:1-24; writes phi on [2, N_tot-1]^3 (1-based) */
int f3d_shock_sensor_f64(const f3d_params* p, double* phi, const double* Q, f3d_stream_t s);
int f3d_shock_sensor_f32b(const f3d_params* p, float* phi, const float* Q, f3d_stream_t s);

/* Face scheme mask, the branch advect_{x,y,z}c takes on every face (schemes.jl:724-731, 813, 931,
 * 1000; y 1038-1045; z 1352-1359).  No reference counterpart as a kernel: it is the integer the
 * reference computes implicitly, materialised for bit-exact checking. */
int f3d_scheme_mask_f64(const f3d_params* p, uint8_t* mask_x, uint8_t* mask_y, uint8_t* mask_z,
                        const double* phi, f3d_stream_t s);
int f3d_scheme_mask_f32b(const f3d_params* p, uint8_t* mask_x, uint8_t* mask_y, uint8_t* mask_z,
                         const float* phi, f3d_stream_t s);

/* fluxSplit_SW(Q, Fp, Fm, S, Ax, Ay, Az)  split.jl:2-68 */
int f3d_flux_split_sw_f64(const f3d_params* p, const double* Q, double* Fp, double* Fm,
                          const double* S, const double* Ax, const double* Ay, const double* Az,
                          f3d_stream_t s);
int f3d_flux_split_sw_f32b(const f3d_params* p, const float* Q, float* Fp, float* Fm,
                           const float* S, const float* Ax, const float* Ay, const float* Az,
                           f3d_stream_t s);
/* fluxSplit_LF / fluxSplit_VL / fluxSplit_AUSM(Q, Fp, Fm, S, Ax, Ay, Az)  split.jl:70 / 110 / 185:
 * local Lax-Friedrichs, van Leer and AUSM splittings, every cell, from rho, u, v, w, p. */
int f3d_flux_split_lf_f64(const f3d_params* p, const double* Q, double* Fp, double* Fm,
                          const double* S, const double* Ax, const double* Ay, const double* Az,
                          f3d_stream_t s);
int f3d_flux_split_lf_f32b(const f3d_params* p, const float* Q, float* Fp, float* Fm,
                           const float* S, const float* Ax, const float* Ay, const float* Az,
                           f3d_stream_t s);
int f3d_flux_split_vl_f64(const f3d_params* p, const double* Q, double* Fp, double* Fm,
                          const double* S, const double* Ax, const double* Ay, const double* Az,
                          f3d_stream_t s);
int f3d_flux_split_vl_f32b(const f3d_params* p, const float* Q, float* Fp, float* Fm,
                           const float* S, const float* Ax, const float* Ay, const float* Az,
                           f3d_stream_t s);
int f3d_flux_split_ausm_f64(const f3d_params* p, const double* Q, double* Fp, double* Fm,
                            const double* S, const double* Ax, const double* Ay, const double* Az,
                            f3d_stream_t s);
int f3d_flux_split_ausm_f32b(const f3d_params* p, const float* Q, float* Fp, float* Fm,
                             const float* S, const float* Ax, const float* Ay, const float* Az,
                             f3d_stream_t s);

/* advect_x / advect_y / advect_z(F, phi, S, Fp, Fm, NV)  schemes.jl:86, 291, 496: the component-wise
 * reconstruction (no characteristic projection) of the first NV (1..5) split-flux components. */
int f3d_advect_x_f64(const f3d_params* p, double* F, const double* phi, const double* S,
                     const double* Fp, const double* Fm, int32_t NV, f3d_stream_t s);
int f3d_advect_y_f64(const f3d_params* p, double* F, const double* phi, const double* S,
                     const double* Fp, const double* Fm, int32_t NV, f3d_stream_t s);
int f3d_advect_z_f64(const f3d_params* p, double* F, const double* phi, const double* S,
                     const double* Fp, const double* Fm, int32_t NV, f3d_stream_t s);
int f3d_advect_x_f32b(const f3d_params* p, float* F, const float* phi, const float* S,
                      const float* Fp, const float* Fm, int32_t NV, f3d_stream_t s);
int f3d_advect_y_f32b(const f3d_params* p, float* F, const float* phi, const float* S,
                      const float* Fp, const float* Fm, int32_t NV, f3d_stream_t s);
int f3d_advect_z_f32b(const f3d_params* p, float* F, const float* phi, const float* S,
                      const float* Fp, const float* Fm, int32_t NV, f3d_stream_t s);

/* advect_xc / advect_yc / advect_zc(F, phi, S, Fp, Fm, Q, Ax, Ay, Az)  schemes.jl:700, 1014, 1330 */
int f3d_advect_xc_f64(const f3d_params* p, double* F, const double* phi, const double* S,
                      const double* Fp, const double* Fm, const double* Q, const double* Ax,
                      const double* Ay, const double* Az, f3d_stream_t s);
int f3d_advect_yc_f64(const f3d_params* p, double* F, const double* phi, const double* S,
                      const double* Fp, const double* Fm, const double* Q, const double* Ax,
                      const double* Ay, const double* Az, f3d_stream_t s);
int f3d_advect_zc_f64(const f3d_params* p, double* F, const double* phi, const double* S,
                      const double* Fp, const double* Fm, const double* Q, const double* Ax,
                      const double* Ay, const double* Az, f3d_stream_t s);
int f3d_advect_xc_f32b(const f3d_params* p, float* F, const float* phi, const float* S,
                       const float* Fp, const float* Fm, const float* Q, const float* Ax,
                       const float* Ay, const float* Az, f3d_stream_t s);
int f3d_advect_yc_f32b(const f3d_params* p, float* F, const float* phi, const float* S,
                       const float* Fp, const float* Fm, const float* Q, const float* Ax,
                       const float* Ay, const float* Az, f3d_stream_t s);
int f3d_advect_zc_f32b(const f3d_params* p, float* F, const float* phi, const float* S,
                       const float* Fp, const float* Fm, const float* Q, const float* Ax,
                       const float* Ay, const float* Az, f3d_stream_t s);

/* viscousFlux(Fv_x, Fv_y, Fv_z, Q, dxi_dx..dzeta_dz, J)  viscous.jl:1-134 (laminar) */
int f3d_viscous_flux_f64(const f3d_params* p, double* Fv_x, double* Fv_y, double* Fv_z,
                         const double* Q, const double* dxidx, const double* dxidy,
                         const double* dxidz, const double* detadx, const double* detady,
                         const double* detadz, const double* dzetadx, const double* dzetady,
                         const double* dzetadz, const double* J, f3d_stream_t s);
int f3d_viscous_flux_f32b(const f3d_params* p, float* Fv_x, float* Fv_y, float* Fv_z,
                          const float* Q, const float* dxidx, const float* dxidy,
                          const float* dxidz, const float* detadx, const float* detady,
                          const float* detadz, const float* dzetadx, const float* dzetady,
                          const float* dzetadz, const double* J, f3d_stream_t s);

/* div(U, Fx, Fy, Fz, Fv_x, Fv_y, Fv_z, dt, J)  div.jl:1-50 : U += J*dt*(-dF + dFv) on the interior */
int f3d_div_f64(const f3d_params* p, double* U, const double* Fx, const double* Fy,
                const double* Fz, const double* Fv_x, const double* Fv_y, const double* Fv_z,
                double dt, const double* J, f3d_stream_t s);
int f3d_div_f32b(const f3d_params* p, double* U, const float* Fx, const float* Fy,
                 const float* Fz, const float* Fv_x, const float* Fv_y, const float* Fv_z,
                 double dt, const double* J, f3d_stream_t s);

/* div_LTS(U, Fx, Fy, Fz, Fv_x, Fv_y, Fv_z, dt, J)  div.jl:52-101 : div with the per-cell dt
 * (cell-shaped, double in both forms; the interior cells are read) in place of the scalar dt */
int f3d_div_lts_f64(const f3d_params* p, double* U, const double* Fx, const double* Fy,
                    const double* Fz, const double* Fv_x, const double* Fv_y, const double* Fv_z,
                    const double* dt, const double* J, f3d_stream_t s);
int f3d_div_lts_f32b(const f3d_params* p, double* U, const float* Fx, const float* Fy,
                     const float* Fz, const float* Fv_x, const float* Fv_y, const float* Fv_z,
                     const double* dt, const double* J, f3d_stream_t s);

/* compute_dt(dt, Q, J, S1, S2, S3)  utils.jl:73-100 : the local time step of every interior cell,
 * dt = lts_cfl * min(dx/(u+c), dy/(v+c), dz/(w+c)), dx = (1/J)/S1, dy = (1/J)/S2, dz = (1/J)/S3
 * (u+c as the reference writes it, not |u|+c).  dt is double in both forms (the reference's
 * LTS_dt is Float32); ghost cells of dt are not written.  The fused stage entries compute it
 * themselves at KRK = 1 when p->lts (sizes from the metric rows) and then ignore their scalar dt. */
int f3d_compute_dt_f64(const f3d_params* p, double* dt, const double* Q, const double* J,
                       const double* S1, const double* S2, const double* S3, f3d_stream_t s);
int f3d_compute_dt_f32b(const f3d_params* p, double* dt, const float* Q, const double* J,
                        const float* S1, const float* S2, const float* S3, f3d_stream_t s);

/* Filters  utils.jl:102-316 (U, Un double in both forms).
 * pre_x/_y/_z(Q, sc, rth)  102-169 : the nonlinear filter's shock sensor along x/y/z from the pressure,
 *   sc = max(0, 1 - rth/r), r = ((dp0-dp1)^2 + (dp0-dp2)^2)/(2 p^2) + 1e-16 of the second differences
 *   dp of p at offsets -2..2.  sc is cell-shaped double and written on the interior and on the first
 *   ghost layer each side along the direction (the reference's sc is (Nxp, Nyp, Nzp) Float32 indexed
 *   with the ghost-offset interior range, so its boundary neighbours alias other cells or lie outside
 *   the allocation; here the boundary faces read the ghost cell's own sensor).
 * filter_x/_y/_z(U, Un, sc, s0)  171-241 : U = Un - s0 (sc+ D+ - sc- D-) on the interior (conservative).
 * linearFilter_x/_y/_z(U, Un, s0)  243-316 : 9-point linear filter of the interior. */
int f3d_pre_x_f64(const f3d_params* p, double* sc, const double* Q, double rth, f3d_stream_t s);
int f3d_pre_y_f64(const f3d_params* p, double* sc, const double* Q, double rth, f3d_stream_t s);
int f3d_pre_z_f64(const f3d_params* p, double* sc, const double* Q, double rth, f3d_stream_t s);
int f3d_pre_x_f32b(const f3d_params* p, double* sc, const float* Q, double rth, f3d_stream_t s);
int f3d_pre_y_f32b(const f3d_params* p, double* sc, const float* Q, double rth, f3d_stream_t s);
int f3d_pre_z_f32b(const f3d_params* p, double* sc, const float* Q, double rth, f3d_stream_t s);
int f3d_filter_x(const f3d_params* p, double* U, const double* Un, const double* sc, double s0,
                 f3d_stream_t s);
int f3d_filter_y(const f3d_params* p, double* U, const double* Un, const double* sc, double s0,
                 f3d_stream_t s);
int f3d_filter_z(const f3d_params* p, double* U, const double* Un, const double* sc, double s0,
                 f3d_stream_t s);
int f3d_linear_filter_x(const f3d_params* p, double* U, const double* Un, double s0, f3d_stream_t s);
int f3d_linear_filter_y(const f3d_params* p, double* U, const double* Un, double s0, f3d_stream_t s);
int f3d_linear_filter_z(const f3d_params* p, double* U, const double* Un, double s0, f3d_stream_t s);

/* linComb(U, Un, NV, a, b)  utils.jl:57-70 : U = a*U + b*Un on the interior (both forms: double) */
int f3d_lin_comb(const f3d_params* p, double* U, const double* Un, int32_t NV, double a,
                 double b, f3d_stream_t s);

/* c2Prim(U, Q)  utils.jl:2-30 (interior)   prim2c(U, Q)  utils.jl:33-54 (whole array) */
int f3d_c2prim_f64(const f3d_params* p, const double* U, double* Q, f3d_stream_t s);
int f3d_c2prim_f32b(const f3d_params* p, const double* U, float* Q, f3d_stream_t s);
int f3d_prim2c_f64(const f3d_params* p, double* U, const double* Q, f3d_stream_t s);
int f3d_prim2c_f32b(const f3d_params* p, double* U, const float* Q, f3d_stream_t s);

/* exchange_ghost(Q, NV, comm_cart, ...)  mpi.jl:2-103 with one rank: for every periodic dimension
 * (x, then y, then z) the NG-thick ghost slabs of Q[..., 1:NV] are copied from the opposite
 * interior slabs, over the full extent of the other two dimensions.  Host staging and MPI are gone. */
int f3d_exchange_ghost_f64(const f3d_params* p, double* Q, int32_t NV, f3d_stream_t s);
int f3d_exchange_ghost_f32b(const f3d_params* p, float* Q, int32_t NV, f3d_stream_t s);

/* fillGhost(Q, U, rankx, ranky)  boundary.jl:151-156 : fill_x, fill_y, fill_z in that order with
 * the face kinds of p->bc.  `inflow` [Nyt,Nzt,6] (device) is read only by F3D_BC_INFLOW, else NULL.
 * U may be NULL (ghost U is not read on the finite-difference path). */
int f3d_fill_ghost_f64(const f3d_params* p, double* Q, double* U, const double* inflow,
                       f3d_stream_t s);
int f3d_fill_ghost_f32b(const f3d_params* p, float* Q, double* U, const float* inflow,
                        f3d_stream_t s);

/* ---------------- finite-volume branch (FVM.jl; selected by p->fvm, solver.jl:21-31) ----------
 * One entry per reference kernel: face fluxes F (Fx [Nxp+1,Nyp,Nzp,5], Fy, Fz as for advect_*) =
 * HLLC flux (FVM.jl:1224-1358) of the states reconstructed to both sides of the face, times the
 * face area.  U: conserved state with its ghost cells filled (fillGhost fills U as well as Q,
 * boundary.jl:1-156); Q: primitive state (the Roe-averaged frame of the eigen variant); phi: shock
 * sensor; S: |grad xi_d| of the direction (s1, s2 or s3).  x, y, z: the mesh points [Nxt,Nyt,Nzt]
 * the metrics are computed from, double in both numeric forms: the face normal and area are formed
 * in FP64 and rounded to the bulk type (the near-wall spacing of a stretched mesh is below float's
 * resolution of the coordinates themselves).
 *   eigen:  characteristic reconstruction in the Roe frame of cells i, i+1   FVM.jl:6-670
 *   conser: component-wise reconstruction of U                               FVM.jl:672-1222 */
int f3d_eigen_reconstruct_i_f64(const f3d_params* p, const double* Q, const double* U,
                                const double* phi, const double* S, double* Fx, const double* x,
                                const double* y, const double* z, f3d_stream_t s);
int f3d_eigen_reconstruct_j_f64(const f3d_params* p, const double* Q, const double* U,
                                const double* phi, const double* S, double* Fy, const double* x,
                                const double* y, const double* z, f3d_stream_t s);
int f3d_eigen_reconstruct_k_f64(const f3d_params* p, const double* Q, const double* U,
                                const double* phi, const double* S, double* Fz, const double* x,
                                const double* y, const double* z, f3d_stream_t s);
int f3d_conser_reconstruct_i_f64(const f3d_params* p, const double* Q, const double* U,
                                 const double* phi, const double* S, double* Fx, const double* x,
                                 const double* y, const double* z, f3d_stream_t s);
int f3d_conser_reconstruct_j_f64(const f3d_params* p, const double* Q, const double* U,
                                 const double* phi, const double* S, double* Fy, const double* x,
                                 const double* y, const double* z, f3d_stream_t s);
int f3d_conser_reconstruct_k_f64(const f3d_params* p, const double* Q, const double* U,
                                 const double* phi, const double* S, double* Fz, const double* x,
                                 const double* y, const double* z, f3d_stream_t s);
int f3d_eigen_reconstruct_i_f32b(const f3d_params* p, const float* Q, const double* U,
                                 const float* phi, const float* S, float* Fx, const double* x,
                                 const double* y, const double* z, f3d_stream_t s);
int f3d_eigen_reconstruct_j_f32b(const f3d_params* p, const float* Q, const double* U,
                                 const float* phi, const float* S, float* Fy, const double* x,
                                 const double* y, const double* z, f3d_stream_t s);
int f3d_eigen_reconstruct_k_f32b(const f3d_params* p, const float* Q, const double* U,
                                 const float* phi, const float* S, float* Fz, const double* x,
                                 const double* y, const double* z, f3d_stream_t s);
int f3d_conser_reconstruct_i_f32b(const f3d_params* p, const float* Q, const double* U,
                                  const float* phi, const float* S, float* Fx, const double* x,
                                  const double* y, const double* z, f3d_stream_t s);
int f3d_conser_reconstruct_j_f32b(const f3d_params* p, const float* Q, const double* U,
                                  const float* phi, const float* S, float* Fy, const double* x,
                                  const double* y, const double* z, f3d_stream_t s);
int f3d_conser_reconstruct_k_f32b(const f3d_params* p, const float* Q, const double* U,
                                  const float* phi, const float* S, float* Fz, const double* x,
                                  const double* y, const double* z, f3d_stream_t s);

/* flowAdvance(U, Q, Fp, Fm, Fx, Fy, Fz, Fv_x, Fv_y, Fv_z, s1, s2, s3, 9 metrics, J, x, y, z, phi)
 * solver.jl:19.  p->fvm == F3D_FDM: the finite-difference branch solver.jl:133-188 (per direction the
 * splitting p->split, then advect_*c when p->character, else advect_*; x, y, z unused); otherwise the finite-volume branch solver.jl:21-31 (eigen or conser
 * reconstruct i, j, k; Fp, Fm unused).  Then viscousFlux when p->viscous.  x, y, z are double in
 * both forms (see the finite-volume entries). */
int f3d_flow_advance_f64(const f3d_params* p, double* U, double* Q, double* Fp, double* Fm,
                         double* Fx, double* Fy, double* Fz, double* Fv_x, double* Fv_y,
                         double* Fv_z, const double* s1, const double* s2, const double* s3,
                         const double* dxidx, const double* dxidy, const double* dxidz,
                         const double* detadx, const double* detady, const double* detadz,
                         const double* dzetadx, const double* dzetady, const double* dzetadz,
                         const double* J, const double* x, const double* y, const double* z,
                         double* phi, f3d_stream_t s);
int f3d_flow_advance_f32b(const f3d_params* p, double* U, float* Q, float* Fp, float* Fm,
                          float* Fx, float* Fy, float* Fz, float* Fv_x, float* Fv_y,
                          float* Fv_z, const float* s1, const float* s2, const float* s3,
                          const float* dxidx, const float* dxidy, const float* dxidz,
                          const float* detadx, const float* detady, const float* detadz,
                          const float* dzetadx, const float* dzetady, const float* dzetadz,
                          const double* J, const double* x, const double* y, const double* z,
                          float* phi, f3d_stream_t s);

/* ---------------- fused path (the MI355X hot path) ------------------------------------------ */

/* Opaque persistent workspace (RHS accumulator, phi, viscous-flux scratch, reduction partials),
 * allocated once for one params/precision pair.  Not thread-safe; re-entrant across workspaces. */
typedef struct f3d_workspace f3d_workspace;
int f3d_workspace_create(const f3d_params* p, int32_t precision, f3d_workspace** ws);
int f3d_workspace_destroy(f3d_workspace* ws);
/* Bytes of device memory held by a workspace. Host-only. */
size_t f3d_workspace_bytes(const f3d_workspace* ws);

/* Metric layout for the fused entries: metrics[9] = device pointers in the reference's order
 * (dxi_dx, dxi_dy, dxi_dz, deta_dx, deta_dy, deta_dz, dzeta_dx, dzeta_dy, dzeta_dz), each
 * [Nxt,Nyt,Nzt]; s1..s3 are not needed (recomputed as |grad xi_d| on the fly). */

/* RHS: R = J * ( -(Fx[i+1]-Fx[i]) - ... + dFv_x/dxi + dFv_y/deta + dFv_z/dzeta ) on interior cells,
 * from primitive Q with ghosts already filled; R has U's layout (ghost entries untouched).
 * Equals the reference's div increment per unit dt (div.jl:40-48) after flowAdvance.
 * `phi` (optional, may be NULL) receives the shock sensor; `mask_*` (optional) the face schemes. */
int f3d_rhs_f64(const f3d_params* p, f3d_workspace* ws, double* R, const double* Q,
                const double* const* metrics, const double* J, double* phi, f3d_stream_t s);
int f3d_rhs_f32b(const f3d_params* p, f3d_workspace* ws, double* R, const float* Q,
                 const float* const* metrics, const double* J, float* phi, f3d_stream_t s);

/* f3d_rhs_f64 without a workspace and without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  The scratch (interior RHS accumulator, heavy-face runs, their corrections and
 * per-cell marks, the sensor classes, and phi when `phi` is NULL) comes from one static 128 MiB device array
 * per GPU, held until the process ends: grids whose scratch exceeds it (about 200 bytes per interior cell:
 * some 0.66M cells, e.g. 87^3) are rejected with F3D_E_DIMS before any launch.
 * Every call on one GPU shares that array: calls that overlap on one GPU (two streams, two threads)
 * race.  Finite-difference branch, fused engine with the viscous fluxes inside the update march only
 * (LES models and the F3D_RHS_ENGINE / F3D_VISC_FUSED switches give F3D_E_ARG). */
int f3d_rhs_noalloc_f64(const f3d_params* p, double* R, const double* Q, const double* const* metrics,
                        const double* J, double* phi, f3d_stream_t s);

/* One SSP-RK3 stage KRK in {1,2,3} of the loop body solver.jl:416-444: (KRK==1: Un <- U), sensor,
 * RHS, U += dt*R, (KRK==2: U = U/4 + 3Un/4, KRK==3: U = 2U/3 + Un/3), c2Prim, ghost fill of Q.
 * p->lts: KRK==1 also computes the per-cell dt (compute_dt) from Q into the workspace, and every
 * stage updates with it (div_LTS) instead of `dt`, which is then ignored; stages 2-3 use the per-cell
 * dt of the workspace's last stage 1. */
int f3d_rk3_stage_f64(const f3d_params* p, f3d_workspace* ws, int32_t KRK, double* U,
                      double* Un, double* Q, const double* const* metrics, const double* J,
                      double dt, const double* inflow, f3d_stream_t s);
int f3d_rk3_stage_f32b(const f3d_params* p, f3d_workspace* ws, int32_t KRK, double* U,
                       double* Un, float* Q, const float* const* metrics, const double* J,
                       double dt, const float* inflow, f3d_stream_t s);

/* `nsteps` full RK3 steps with a fixed dt (time_step's loop, solver.jl:410-444, without I/O);
 * p->lts: a per-cell dt recomputed at the start of every step, as above.  p->filtering: after the t-th
 * step of the call (t = 1..nsteps) with t % filtering_interval == 0, the filtering block
 * (f3d_filter_step_*); a caller that steps one call at a time filters with f3d_filter_step_* when its
 * own step counter reaches a multiple of the interval. */
int f3d_time_step_f64(const f3d_params* p, f3d_workspace* ws, int32_t nsteps, double* U,
                      double* Un, double* Q, const double* const* metrics, const double* J,
                      double dt, const double* inflow, f3d_stream_t s);
int f3d_time_step_f32b(const f3d_params* p, f3d_workspace* ws, int32_t nsteps, double* U,
                       double* Un, float* Q, const float* const* metrics, const double* J,
                       double dt, const float* inflow, f3d_stream_t s);

/* The filtering block of the time loop  solver.jl:446-470 : for x, y, z in turn, Un <- U, then the
 * nonlinear filter with its sensor from Q (p->filtering_nonlinear, threshold filtering_rth) or the
 * linear filter, strength filtering_s0.  Q is not updated (as in the reference).  U's ghost cells
 * must hold the boundary procedure's values: the stage entries refill them when p->filtering. */
int f3d_filter_step_f64(const f3d_params* p, f3d_workspace* ws, double* U, double* Un, const double* Q,
                        f3d_stream_t s);
int f3d_filter_step_f32b(const f3d_params* p, f3d_workspace* ws, double* U, double* Un, const float* Q,
                         f3d_stream_t s);

/* Face geometry of the finite-volume branch, once per mesh (as f3d_metrics_f64 for the metrics):
 * geo[d] (d = 0, 1, 2) = [face extents of Fx / Fy / Fz, 4] holding (nx, ny, nz, Area) of every
 * face, n the unit normal pointing along +d (FVM.jl:16-31, 250-264, 465-479), formed in FP64 from the
 * mesh points and stored in the bulk type.  The fused entries below read it instead of forming the
 * geometry from the 24 mesh points around each face on every stage. */
int f3d_fvm_geometry_f64(const f3d_params* p, const double* x, const double* y, const double* z,
                         double* const* geo, f3d_stream_t s);
int f3d_fvm_geometry_f32b(const f3d_params* p, const double* x, const double* y, const double* z,
                          float* const* geo, f3d_stream_t s);

/* Finite-volume form of the three entries above (p->fvm = F3D_FVM_EIGEN or F3D_FVM_CONSER; the
 * entries without _fvm reject finite-volume params and these reject F3D_FDM).  Extra arguments: the
 * conserved state U (its ghost cells are read by the reconstruction; the stage drivers refill them
 * together with Q's, boundary.jl:1-156) and geo[3] from f3d_fvm_geometry_*.  f3d_rhs_fvm_* reads U
 * and does not modify it. */
int f3d_rhs_fvm_f64(const f3d_params* p, f3d_workspace* ws, double* R, const double* Q,
                    const double* U, const double* const* metrics, const double* J,
                    const double* const* geo, double* phi, f3d_stream_t s);
int f3d_rhs_fvm_f32b(const f3d_params* p, f3d_workspace* ws, double* R, const float* Q,
                     const double* U, const float* const* metrics, const double* J,
                     const float* const* geo, float* phi, f3d_stream_t s);
int f3d_rk3_stage_fvm_f64(const f3d_params* p, f3d_workspace* ws, int32_t KRK, double* U,
                          double* Un, double* Q, const double* const* metrics, const double* J,
                          const double* const* geo, double dt, const double* inflow,
                          f3d_stream_t s);
int f3d_rk3_stage_fvm_f32b(const f3d_params* p, f3d_workspace* ws, int32_t KRK, double* U,
                           double* Un, float* Q, const float* const* metrics, const double* J,
                           const float* const* geo, double dt, const float* inflow,
                           f3d_stream_t s);
int f3d_time_step_fvm_f64(const f3d_params* p, f3d_workspace* ws, int32_t nsteps, double* U,
                          double* Un, double* Q, const double* const* metrics, const double* J,
                          const double* const* geo, double dt, const double* inflow,
                          f3d_stream_t s);
int f3d_time_step_fvm_f32b(const f3d_params* p, f3d_workspace* ws, int32_t nsteps, double* U,
                           double* Un, float* Q, const float* const* metrics, const double* J,
                           const float* const* geo, double dt, const float* inflow,
                           f3d_stream_t s);

/* Global CFL time step (no reference counterpart as a reduction; per-cell formula of compute_dt,
 * utils.jl:73-100, with |contravariant velocity| + c):
 *   dt = cfl * min over interior cells of  1 / ( J * max_d (|A_d . u| + c*|A_d|) ),  c = sqrt(gamma Rg T).
 * Result written to *dt_out (device double). */
int f3d_cfl_dt_f64(const f3d_params* p, f3d_workspace* ws, double* dt_out, const double* Q,
                   const double* const* metrics, const double* J, double cfl, f3d_stream_t s);
int f3d_cfl_dt_f32b(const f3d_params* p, f3d_workspace* ws, double* dt_out, const float* Q,
                    const float* const* metrics, const double* J, double cfl, f3d_stream_t s);

/* any(isnan, U) of solver.jl:481 over the whole U array (NaN or Inf); writes 0/1 to *flag_out
 * (device int32).  Synchronous variant below returns F3D_E_NAN when set. */
int f3d_nan_check(const f3d_params* p, f3d_workspace* ws, int32_t* flag_out, const double* U,
                  int32_t NV, f3d_stream_t s);
int f3d_nan_check_sync(const f3d_params* p, f3d_workspace* ws, const double* U, int32_t NV,
                       f3d_stream_t s);

/* Metrics from node coordinates (parse_mesh.jl:139-248), always FP64: CD6 interior, one-sided
 * second order at the first/last three indices, J = 1/det, nine J^-1-scaled metrics in the order
 * of `metrics`; s_out[3] (optional, may hold NULLs) receive |grad xi_d| (solver.jl:286-288). */
int f3d_metrics_f64(const f3d_params* p, const double* x, const double* y, const double* z,
                    double* const* metrics, double* J, double* const* s_out, f3d_stream_t s);

/* ---------------- adjoint (reverse mode, FP64; no reference counterpart) ---------------------
 * Hand-written vector-Jacobian products of the path above with the scheme mask frozen at the
 * primal state (the sensor, face schemes, minmod selections and tangent-vector branch are not
 * differentiated; |x| uses the subgradient sign(0) = 0; a WENO smoothness indicator below the
 * rounding floor of its own evaluation -- a stencil flat to the last bit -- passes no weight
 * derivative).  They back a ChainRulesCore.rrule of the RHS on the Julia side.  The adjoint
 * workspace holds its own FP64 forward workspace.  Finite-difference branch only: the workspace,
 * RHS/stage VJPs and the multi-step driver reject p->fvm != F3D_FDM (the ghost and c2Prim VJPs
 * are the same for both branches). */
typedef struct f3d_adj_workspace f3d_adj_workspace;
int f3d_adjoint_workspace_create(const f3d_params* p, f3d_adj_workspace** aws);
int f3d_adjoint_workspace_destroy(f3d_adj_workspace* aws);
size_t f3d_adjoint_workspace_bytes(const f3d_adj_workspace* aws); /* host-only */

/* lamQ = (dR/dQ)^T lamR for R of f3d_rhs_f64 (interior, U layout).  lamQ [Nxt,Nyt,Nzt,6] is
 * written for every cell, ghosts included (ghost adjoints are folded by f3d_ghost_vjp_f64); its
 * p entries are zero (the RHS reads rho,u,v,w,T). */
int f3d_rhs_vjp_f64(const f3d_params* p, f3d_adj_workspace* aws, const double* lamR, const double* Q,
                    const double* const* metrics, const double* J, double* lamQ, f3d_stream_t s);

/* Reverse of the stage ghost procedure (exchange_ghost then fillGhost, solver.jl:439-443) on the
 * filled Q: ghost adjoints of lamQ are added into their source cells and cleared; Dirichlet,
 * inflow and isothermal-wall ghost values add to the parameter adjoints
 *   lam_inflow [Nyt,Nzt,6] (may be NULL), lam_bc[7] = (q_inf[0..5], T_wall) (may be NULL). */
int f3d_ghost_vjp_f64(const f3d_params* p, double* lamQ, const double* Q, double* lam_inflow,
                      double* lam_bc, f3d_stream_t s);

/* lamU = c2Prim^T lamQ at U on interior cells (utils.jl:2-30; eps floors pass nothing). */
int f3d_c2prim_vjp_f64(const f3d_params* p, double* lamU, const double* lamQ, const double* U,
                       f3d_stream_t s);

/* Reverse of f3d_rk3_stage_f64 stage KRK taken at its input (Us, Qs filled): from the adjoint of
 * the stage output lam_out, lam_s = adjoint of Us (may alias lam_out);  lamUn (adjoint of Un) is
 * set to lam_out/3 by KRK 3, incremented by 3 lam_out/4 by KRK 2 and added into lam_s by KRK 1
 * (Un = U^n), so calling KRK 3, 2, 1 in turn reverses one RK3 step.  lam_inflow/lam_bc accumulate. */
int f3d_rk3_stage_vjp_f64(const f3d_params* p, f3d_adj_workspace* aws, int32_t KRK,
                          const double* lam_out, double* lam_s, double* lamUn, const double* Us,
                          const double* Qs, const double* const* metrics, const double* J,
                          double dt, double* lam_inflow, double* lam_bc, f3d_stream_t s);

/* Device bytes of the checkpoint buffer f3d_adjoint_f64 needs for `nsteps` with a checkpoint
 * every `every` steps: ceil(nsteps/every) + every + 2*(every == 1 ? nsteps : every) conserved
 * states (checkpoints, the recomputed states of a segment, the final state, and the two stage
 * states of each step of a segment); with every == 1 (and the fused engine) also the primitive states
 * (6 values, ghosts filled) of the three stage inputs of every step, which the forward sweep keeps so
 * that the reverse sweep forms none; an `every` above nsteps counts as nsteps (one segment).
 * Host-only. */
size_t f3d_adjoint_ckpt_bytes(const f3d_params* p, int32_t nsteps, int32_t every);

/* Checkpointed multi-step adjoint.  Forward: Q0 = ghosts(c2Prim(U0)), `nsteps` RK3 steps of
 * f3d_rk3_stage_f64 with a fixed dt, obj[n] = dt * sum over interior cells of W . Q^{n+1}
 * (W [Nxt,Nyt,Nzt,6], e.g. wall-face area on p of the first cell row above a wall).  Reverse:
 * U^n checkpointed every `every` steps; each segment recomputed once from its checkpoint, keeping
 * the states and the stage states U^1, U^2 of its steps (every == 1: kept by the forward sweep);
 * stage VJPs 3,2,1 at (U_s, c2Prim + ghosts of U_s).  Outputs: lamU0 = d(sum obj)/dU0
 * (interior; ghosts 0), lam_inflow / lam_bc accumulate d(sum obj)/d(inflow, q_inf, T_wall). */
int f3d_adjoint_f64(const f3d_params* p, f3d_adj_workspace* aws, int32_t nsteps, int32_t every,
                    const double* U0, const double* const* metrics, const double* J, double dt,
                    const double* inflow, const double* W, double* ckpt, size_t ckpt_bytes,
                    double* obj, double* lamU0, double* lam_inflow, double* lam_bc,
                    f3d_stream_t s);

/* ---------------- profiling: hipEvent timing of the library's own kernel launches -------------
 * No reference counterpart (the reference prints wall-clock time only, solver.jl:472-479).
 * When enabled, every kernel the library launches is bracketed by a pair of hipEvents recorded on
 * the launching stream; queries synchronise on those events.  Process-global, not thread-safe. */
int f3d_timing_enable(int32_t on);
/* Restrict the bracketing to the launches named `name` (NULL: every launch), so that a timed
 * region can carry the events of one kernel only. */
int f3d_timing_only(const char* name);
int f3d_timing_reset(void);
/* Sum of elapsed ms and number of launches of the kernels whose name equals `name`. */
int f3d_timing_query(const char* name, double* total_ms, int64_t* launches);
/* Comma-separated distinct kernel names timed since the last reset (truncated to len-1 chars). */
int f3d_timing_names(char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* F3D_H */
