// C entry points of the CPU oracle (liboracle_f3d.so), loaded through ctypes by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  TEST INFRASTRUCTURE ONLY: the
// product path (flame3d-mhd_amd/, libf3d_hip.so) never links, loads or calls this file.
//
// What pins each oracle function apart from GPU parity (test file : test name, under tests/):
//   c2prim / prim2c (utils.jl:2-54)   round trip, eps floors, hand-evaluated cell
//                                     (test_oracle_invariants: test_c2prim_prim2c_round_trip_and_floors).
//   shock_sensor (schemes.jl:1-24)    constant p -> 0 and a numpy evaluation (test_sensor_constant_and_numpy).
//   scheme_mask (schemes.jl:724-1009) band pattern at distance <= 3 of a spike (test_mask_band_around_spike).
//   flux_split_sw (split.jl:2-68)     Fp+Fm = Euler flux on the metric row; Fm = 0 when supersonic
//                                     (test_split_sums_to_euler_flux_and_supersonic_limit).
//   advect_c UP7 (schemes.jl:731-754) 7th-order flux difference (test_up7_seventh_order); uniform state
//                                     (test_uniform_state_reconstructs_exactly, both tangent branches:
//                                     test_uniform_state_affine_metrics_both_tangent_branches).
//   WENO7 / WENO5 (schemes.jl:813-999, FVM.jl:116-201) equal to an exact rational evaluation of their
//                                     definition (polynomial candidates, Jiang-Shu/Balsara-Shu integral
//                                     indicators, derived linear weights; FP64, long double, dual) on
//                                     random stencils with eps negligible, dominant and comparable
//                                     (test_oracle_weno: test_weno_equals_definition); ENO selection at a
//                                     step (test_eno_selection_at_a_step); order 7 / 5 on smooth data,
//                                     scalar and through advect_c and the finite-volume reconstructions
//                                     with the band forced (test_weno_order_on_smooth_data,
//                                     test_advect_c_forced_band_order, test_fvm_reconstruct_forced_band_order);
//                                     linear data -> linear schemes (test_weno_linear_data_reduces_to_linear_schemes).
//   NND2 (schemes.jl:1000-1009)       minmod truth table (test_nnd_minmod_truth_table).
//   viscous_flux + div (viscous.jl, div.jl) linear velocity at constant T -> zero momentum RHS and the
//                                     dissipation in energy (test_viscous_linear_velocity_closed_form);
//                                     affine fields incl. grad T on sheared meshes against the closed-form
//                                     flux (test_oracle_les: test_les_general_gradient[LES_NONE]); periodic
//                                     telescoping sum(dU/J) = 0 (test_periodic_conservation_telescopes);
//                                     free stream -> 0 (test_free_stream_rhs_is_zero).
//   exchange_ghost / fill_ghost       index identities on a labelled array (test_ghost_fill_index_identities,
//                                     test_exchange_periodic_slabs).
//   rk3 stage blends (solver.jl:416-444) third order in dt on a smooth case, for the finite-difference,
//                                     finite-volume and forward-mode steps (test_oracle_weno:
//                                     test_rk3_third_order_in_time).
//   cfl_dt                            uniform flow closed form (test_cfl_dt_uniform_flow).
//   metrics (parse_mesh.jl:139-248)   affine grid -> the cofactor matrix (test_metrics_affine_grid_exact).
//   fvm geometry (FVM.jl:16-31)       affine mesh -> n*Area = the J^-1-scaled metric row of the direction,
//                                     curved mesh -> second order (test_oracle_fvm).
//   hllc_flux (FVM.jl:1224-1358)      UL = UR -> the Euler flux; mirror and rotation symmetry; contact and
//                                     shear exact; the p*-form of the star flux (test_oracle_fvm).
//   fvm reconstruction (FVM.jl:33-226, 707-846) cell averages of polynomials -> the exact face value
//                                     (UP7 degree 6, WENO7 3, WENO5 2, minmod-left 1), eigen and conser;
//                                     the right minmod state as written (test_oracle_fvm).
//   forward mode (f3d_dual.hpp)       central finite differences (test_oracle_jvp).
// No reference test, fixture or golden vector exists for this path (the reference has no tests);
// the CUDA.jl fastmath Float32 outputs of the reference itself are unpinned (no Julia here).
#include "f3d_oracle.hpp"
#include "f3d_oracle_fvm.hpp"
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace f3d_oracle;

#define OR_API extern "C" __attribute__((visibility("default")))

namespace {
template <typename T>
struct M9 {
  const T* m[9];
  M9(const void* const* p) {
    for (int i = 0; i < 9; ++i) m[i] = (const T*)p[i];
  }
};
}  // namespace

OR_API int oracle_version(void) { return F3D_VERSION; }

OR_API void oracle_c2prim_f64(const f3d_params* p, const double* U, double* Q) { c2prim<double>(p, U, Q); }
OR_API void oracle_c2prim_f32b(const f3d_params* p, const double* U, float* Q) { c2prim<float>(p, U, Q); }
OR_API void oracle_prim2c_f64(const f3d_params* p, double* U, const double* Q) { prim2c<double>(p, U, Q); }
OR_API void oracle_prim2c_f32b(const f3d_params* p, double* U, const float* Q) { prim2c<float>(p, U, Q); }
OR_API void oracle_lin_comb(const f3d_params* p, double* U, const double* Un, int NV, double a, double b) {
  lin_comb(p, U, Un, NV, a, b);
}
OR_API void oracle_shock_sensor_f64(const f3d_params* p, double* phi, const double* Q) { shock_sensor<double>(p, phi, Q); }
OR_API void oracle_shock_sensor_f32b(const f3d_params* p, float* phi, const float* Q) { shock_sensor<float>(p, phi, Q); }
OR_API void oracle_scheme_mask_f64(const f3d_params* p, uint8_t* mx, uint8_t* my, uint8_t* mz, const double* phi) {
  scheme_mask<double>(p, mx, my, mz, phi);
}
OR_API void oracle_scheme_mask_f32b(const f3d_params* p, uint8_t* mx, uint8_t* my, uint8_t* mz, const float* phi) {
  scheme_mask<float>(p, mx, my, mz, phi);
}
OR_API void oracle_flux_split_sw_f64(const f3d_params* p, const double* Q, double* Fp, double* Fm, const double* S,
                                     const double* Ax, const double* Ay, const double* Az) {
  flux_split_sw<double>(p, Q, Fp, Fm, S, Ax, Ay, Az);
}
OR_API void oracle_flux_split_sw_f32b(const f3d_params* p, const float* Q, float* Fp, float* Fm, const float* S,
                                      const float* Ax, const float* Ay, const float* Az) {
  flux_split_sw<float>(p, Q, Fp, Fm, S, Ax, Ay, Az);
}
OR_API void oracle_advect_c_f64(const f3d_params* p, int d, double* F, const double* phi, const double* S,
                                const double* Fp, const double* Fm, const double* Q, const double* Ax,
                                const double* Ay, const double* Az) {
  advect_c<double>(p, d, F, phi, S, Fp, Fm, Q, Ax, Ay, Az);
}
OR_API void oracle_advect_c_f32b(const f3d_params* p, int d, float* F, const float* phi, const float* S,
                                 const float* Fp, const float* Fm, const float* Q, const float* Ax,
                                 const float* Ay, const float* Az) {
  advect_c<float>(p, d, F, phi, S, Fp, Fm, Q, Ax, Ay, Az);
}
OR_API void oracle_flux_split_f64(const f3d_params* p, int method, const double* Q, double* Fp, double* Fm,
                                  const double* S, const double* Ax, const double* Ay, const double* Az) {
  if (method == F3D_SPLIT_SW)
    flux_split_sw<double>(p, Q, Fp, Fm, S, Ax, Ay, Az);
  else
    flux_split_other<double>(p, method, Q, Fp, Fm, S, Ax, Ay, Az);
}
OR_API void oracle_flux_split_f32b(const f3d_params* p, int method, const float* Q, float* Fp, float* Fm,
                                   const float* S, const float* Ax, const float* Ay, const float* Az) {
  if (method == F3D_SPLIT_SW)
    flux_split_sw<float>(p, Q, Fp, Fm, S, Ax, Ay, Az);
  else
    flux_split_other<float>(p, method, Q, Fp, Fm, S, Ax, Ay, Az);
}
OR_API void oracle_advect_f64(const f3d_params* p, int d, double* F, const double* phi, const double* S,
                              const double* Fp, const double* Fm, int NV) {
  advect_comp<double>(p, d, F, phi, S, Fp, Fm, NV);
}
OR_API void oracle_advect_f32b(const f3d_params* p, int d, float* F, const float* phi, const float* S,
                               const float* Fp, const float* Fm, int NV) {
  advect_comp<float>(p, d, F, phi, S, Fp, Fm, NV);
}
OR_API void oracle_viscous_flux_f64(const f3d_params* p, double* Fvx, double* Fvy, double* Fvz, const double* Q,
                                    const void* const* m, const double* J) {
  viscous_flux<double>(p, Fvx, Fvy, Fvz, Q, M9<double>(m).m, J);
}
OR_API void oracle_viscous_flux_f32b(const f3d_params* p, float* Fvx, float* Fvy, float* Fvz, const float* Q,
                                     const void* const* m, const double* J) {
  viscous_flux<float>(p, Fvx, Fvy, Fvz, Q, M9<float>(m).m, J);
}
OR_API void oracle_div_f64(const f3d_params* p, double* U, const double* Fx, const double* Fy, const double* Fz,
                           const double* Fvx, const double* Fvy, const double* Fvz, double dt, const double* J) {
  div<double>(p, U, Fx, Fy, Fz, Fvx, Fvy, Fvz, dt, J);
}
OR_API void oracle_div_f32b(const f3d_params* p, double* U, const float* Fx, const float* Fy, const float* Fz,
                            const float* Fvx, const float* Fvy, const float* Fvz, double dt, const double* J) {
  div<float>(p, U, Fx, Fy, Fz, Fvx, Fvy, Fvz, dt, J);
}
OR_API void oracle_exchange_ghost_f64(const f3d_params* p, double* Q, int NV) { exchange_ghost<double>(p, Q, NV); }
OR_API void oracle_exchange_ghost_f32b(const f3d_params* p, float* Q, int NV) { exchange_ghost<float>(p, Q, NV); }
OR_API void oracle_fill_ghost_f64(const f3d_params* p, double* Q, double* U, const double* inflow) {
  fill_ghost<double>(p, Q, U, inflow);
}
OR_API void oracle_fill_ghost_f32b(const f3d_params* p, float* Q, double* U, const float* inflow) {
  fill_ghost<float>(p, Q, U, inflow);
}
OR_API void oracle_metric_norms_f64(const f3d_params* p, const void* const* m, double* s1, double* s2, double* s3) {
  metric_norms<double>(p, M9<double>(m).m, s1, s2, s3);
}
OR_API void oracle_metric_norms_f32b(const f3d_params* p, const void* const* m, float* s1, float* s2, float* s3) {
  metric_norms<float>(p, M9<float>(m).m, s1, s2, s3);
}

namespace {
// RHS per unit dt exactly as the reference produces it: sensor, flowAdvance, then div into a
// zeroed U with dt = 1 (div.jl:40-48).
template <typename T>
void rhs_ref(const f3d_params* p, double* R, const T* Q, const void* const* mv, const double* J, T* phi_out,
             uint8_t* mx, uint8_t* my, uint8_t* mz) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  Scratch<T> sc(p);
  M9<T> m(mv);
  std::vector<T> s1(nc), s2(nc), s3(nc);
  metric_norms<T>(p, m.m, s1.data(), s2.data(), s3.data());
  std::vector<T> Qc(Q, Q + nc * 6);
  shock_sensor<T>(p, sc.phi.data(), Qc.data());
  flow_advance<T>(p, Qc.data(), sc.Fp.data(), sc.Fm.data(), sc.Fx.data(), sc.Fy.data(), sc.Fz.data(),
                  sc.Fvx.data(), sc.Fvy.data(), sc.Fvz.data(), s1.data(), s2.data(), s3.data(), m.m, J,
                  sc.phi.data());
  std::fill(R, R + nc * 5, 0.0);
  div<T>(p, R, sc.Fx.data(), sc.Fy.data(), sc.Fz.data(), sc.Fvx.data(), sc.Fvy.data(), sc.Fvz.data(), 1.0, J);
  if (phi_out) std::copy(sc.phi.begin(), sc.phi.end(), phi_out);
  if (mx || my || mz) scheme_mask<T>(p, mx, my, mz, sc.phi.data());
}

template <typename T>
void time_step_ref(const f3d_params* p, int nsteps, double* U, double* Un, T* Q, const void* const* mv,
                   const double* J, double dt, const T* inflow) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  Scratch<T> sc(p);
  M9<T> m(mv);
  std::vector<T> s1(nc), s2(nc), s3(nc);
  metric_norms<T>(p, m.m, s1.data(), s2.data(), s3.data());
  for (int t = 1; t <= nsteps; ++t) {
    for (int KRK = 1; KRK <= 3; ++KRK)
      rk3_stage<T>(p, sc, KRK, U, Un, Q, s1.data(), s2.data(), s3.data(), m.m, J, dt, inflow);
    // solver.jl:446: filtering after the steps whose index is a multiple of filtering_interval
    if (p->filtering && t % p->filtering_interval == 0) filter_block<T>(p, U, Un, Q, sc.fsc.data());
  }
}
}  // namespace

// compute_dt (utils.jl:73-100) into dt (cell-shaped, double; interior cells written) and div_LTS
// (div.jl:52-101) with such a per-cell dt.
OR_API void oracle_compute_dt_f64(const f3d_params* p, double* dt, const double* Q, const double* J, const double* S1,
                                  const double* S2, const double* S3) {
  compute_dt<double>(p, dt, Q, J, S1, S2, S3);
}
OR_API void oracle_compute_dt_f32b(const f3d_params* p, double* dt, const float* Q, const double* J, const float* S1,
                                   const float* S2, const float* S3) {
  compute_dt<float>(p, dt, Q, J, S1, S2, S3);
}
OR_API void oracle_div_lts_f64(const f3d_params* p, double* U, const double* Fx, const double* Fy, const double* Fz,
                               const double* Fvx, const double* Fvy, const double* Fvz, const double* dt,
                               const double* J) {
  div<double>(p, U, Fx, Fy, Fz, Fvx, Fvy, Fvz, 0.0, J, dt);
}
OR_API void oracle_div_lts_f32b(const f3d_params* p, double* U, const float* Fx, const float* Fy, const float* Fz,
                                const float* Fvx, const float* Fvy, const float* Fvz, const double* dt,
                                const double* J) {
  div<float>(p, U, Fx, Fy, Fz, Fvx, Fvy, Fvz, 0.0, J, dt);
}

// Filters (utils.jl:102-316): the sensor along d, the nonlinear and linear filters along d, and the
// whole filtering block of the time loop (solver.jl:446-470).
OR_API void oracle_pre_sensor_f64(const f3d_params* p, int d, double* sc, const double* Q, double rth) {
  pre_sensor<double>(p, d, sc, Q, rth);
}
OR_API void oracle_pre_sensor_f32b(const f3d_params* p, int d, double* sc, const float* Q, double rth) {
  pre_sensor<float>(p, d, sc, Q, rth);
}
OR_API void oracle_filter_nonlinear(const f3d_params* p, int d, double* U, const double* Un, const double* sc,
                                    double s0) {
  filter_nonlinear(p, d, U, Un, sc, s0);
}
OR_API void oracle_filter_linear(const f3d_params* p, int d, double* U, const double* Un, double s0) {
  filter_linear(p, d, U, Un, s0);
}
OR_API void oracle_filter_block_f64(const f3d_params* p, double* U, double* Un, const double* Q, double* sc) {
  filter_block<double>(p, U, Un, Q, sc);
}
OR_API void oracle_filter_block_f32b(const f3d_params* p, double* U, double* Un, const float* Q, double* sc) {
  filter_block<float>(p, U, Un, Q, sc);
}

OR_API void oracle_rhs_f64(const f3d_params* p, double* R, const double* Q, const void* const* m, const double* J,
                           double* phi, uint8_t* mx, uint8_t* my, uint8_t* mz) {
  rhs_ref<double>(p, R, Q, m, J, phi, mx, my, mz);
}
OR_API void oracle_rhs_f32b(const f3d_params* p, double* R, const float* Q, const void* const* m, const double* J,
                            float* phi, uint8_t* mx, uint8_t* my, uint8_t* mz) {
  rhs_ref<float>(p, R, Q, m, J, phi, mx, my, mz);
}
OR_API void oracle_time_step_f64(const f3d_params* p, int nsteps, double* U, double* Un, double* Q,
                                 const void* const* m, const double* J, double dt, const double* inflow) {
  time_step_ref<double>(p, nsteps, U, Un, Q, m, J, dt, inflow);
}
OR_API void oracle_time_step_f32b(const f3d_params* p, int nsteps, double* U, double* Un, float* Q,
                                  const void* const* m, const double* J, double dt, const float* inflow) {
  time_step_ref<float>(p, nsteps, U, Un, Q, m, J, dt, inflow);
}
OR_API double oracle_cfl_dt_f64(const f3d_params* p, const double* Q, const void* const* m, const double* J,
                                double cfl) {
  return cfl_dt<double>(p, Q, M9<double>(m).m, J, cfl);
}
OR_API double oracle_cfl_dt_f32b(const f3d_params* p, const float* Q, const void* const* m, const double* J,
                                 double cfl) {
  return cfl_dt<float>(p, Q, M9<float>(m).m, J, cfl);
}
OR_API void oracle_metrics_f64(const f3d_params* p, const double* x, const double* y, const double* z,
                               void* const* m, double* J) {
  double* mm[9];
  for (int i = 0; i < 9; ++i) mm[i] = (double*)m[i];
  metrics(p, x, y, z, mm, J);
}
namespace {
// Viscous part of rhs_absterms: J * sum_d (|c1|(|Fv_-2|+|Fv_+2|) + |c2|(|Fv_-1|+|Fv_+1|)) (div.jl:14-48).
template <typename T>
void viscous_absterms(const f3d_params* p, double* A, const T* Q, const T* const m[9], const double* J) {
  Grid g(p);
  const long NG = g.NG;
  Scratch<T> sc(p);
  viscous_flux<T>(p, sc.Fvx.data(), sc.Fvy.data(), sc.Fvz.data(), Q, m, J);
  const double c1 = 1.0 / 12, c2 = 2.0 / 3;
  for (long k = 1; k <= g.Nzp; ++k)
    for (long j = 1; j <= g.Nyp; ++j)
      for (long i = 1; i <= g.Nxp; ++i) {
        const double Jc = std::fabs(J[g.cell.at(i + NG, j + NG, k + NG)]);
        for (int v = 1; v <= 4; ++v) {
          auto a = [&](const std::vector<T>& F, long ii, long jj, long kk) { return std::fabs((double)F[g.fv.at(ii, jj, kk, v)]); };
          double s = c1 * (a(sc.Fvx, i, j + 2, k + 2) + a(sc.Fvx, i + 4, j + 2, k + 2)) + c2 * (a(sc.Fvx, i + 1, j + 2, k + 2) + a(sc.Fvx, i + 3, j + 2, k + 2));
          s += c1 * (a(sc.Fvy, i + 2, j, k + 2) + a(sc.Fvy, i + 2, j + 4, k + 2)) + c2 * (a(sc.Fvy, i + 2, j + 1, k + 2) + a(sc.Fvy, i + 2, j + 3, k + 2));
          s += c1 * (a(sc.Fvz, i + 2, j + 2, k) + a(sc.Fvz, i + 2, j + 2, k + 4)) + c2 * (a(sc.Fvz, i + 2, j + 2, k + 1) + a(sc.Fvz, i + 2, j + 2, k + 3));
          A[g.cell.at(i + NG, j + NG, k + NG, v + 1)] += Jc * s;
        }
      }
}

// Magnitude of the terms the RHS is summed from, per cell and variable (the scale the parity bar is
// relative to).  Inviscid part: each face flux is a sum over its stencil of c_q*Fp and c_q*Fm
// (schemes.jl:741-753), so its terms have magnitude t = sum_q |c_q| (|Fp_q| + |Fm_q|) -- the split
// fluxes cancel wherever the face velocity is small (|Fp|, |Fm| ~ rho c |A| while Fp+Fm ~ rho u |A|).
// A = J*(t_left + t_right summed over the three directions) + J*sum_d(|c1|(|Fv_-2|+|Fv_+2|) +
// |c2|(|Fv_-1|+|Fv_+1|)) for the viscous part (div.jl:14-48).
template <typename T>
void rhs_absterms(const f3d_params* p, double* A, const T* Q, const void* const* mv, const double* J) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  Scratch<T> sc(p);
  M9<T> m(mv);
  std::vector<T> s1(nc), s2(nc), s3(nc);
  metric_norms<T>(p, m.m, s1.data(), s2.data(), s3.data());
  const T* S[3] = {s1.data(), s2.data(), s3.data()};
  std::fill(A, A + nc * 5, 0.0);
  const long NG = g.NG;
  T cl[7];
  linear_coeffs<T>(p, cl);
  for (int d = 0; d < 3; ++d) {
    if (p->split == F3D_SPLIT_SW)
      flux_split_sw<T>(p, Q, sc.Fp.data(), sc.Fm.data(), S[d], m.m[3 * d], m.m[3 * d + 1], m.m[3 * d + 2]);
    else
      flux_split_other<T>(p, p->split, Q, sc.Fp.data(), sc.Fm.data(), S[d], m.m[3 * d], m.m[3 * d + 1],
                          m.m[3 * d + 2]);
    Dir D = make_dir(d);
    for (long k = NG + 1; k <= g.Nzp + NG; ++k)
      for (long j = NG + 1; j <= g.Nyp + NG; ++j)
        for (long i = NG + 1; i <= g.Nxp + NG; ++i) {
          const double Jc = std::fabs(J[g.cell.at(i, j, k)]);
          for (long n = 1; n <= 5; ++n) {
            double t = 0;
            for (int side = 0; side <= 1; ++side) {  // left face (cell i) and right face (cell i+1)
              const long fi = i + side * D.di, fj = j + side * D.dj, fk = k + side * D.dk;
              for (int q = 0; q < 7; ++q) {
                const long op = q - 4, om = 3 - q;
                t += std::fabs((double)cl[q]) *
                     (std::fabs((double)sc.Fp[g.cell.at(fi + op * D.di, fj + op * D.dj, fk + op * D.dk, n)]) +
                      std::fabs((double)sc.Fm[g.cell.at(fi + om * D.di, fj + om * D.dj, fk + om * D.dk, n)]));
              }
            }
            A[g.cell.at(i, j, k, n)] += Jc * t;
          }
        }
  }
  if (p->viscous) viscous_absterms<T>(p, A, Q, m.m, J);
}

// advect_c evaluated in long double from double inputs (conditioning reference for the tests).
void advect_ld(const f3d_params* p, int d, const double* phi, const double* S, const double* Fp, const double* Fm,
               const double* Q, const double* Ax, const double* Ay, const double* Az, double* F) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  std::vector<long double> phiL(phi, phi + nc), SL(S, S + nc), FpL(Fp, Fp + 5 * nc), FmL(Fm, Fm + 5 * nc),
      QL(Q, Q + 6 * nc), AxL(Ax, Ax + nc), AyL(Ay, Ay + nc), AzL(Az, Az + nc);
  size_t nf = (size_t)(g.Nxp + (d == 0)) * (g.Nyp + (d == 1)) * (g.Nzp + (d == 2)) * 5;
  std::vector<long double> FL(nf);
  advect_c<long double>(p, d, FL.data(), phiL.data(), SL.data(), FpL.data(), FmL.data(), QL.data(), AxL.data(),
                        AyL.data(), AzL.data());
  for (size_t i = 0; i < nf; ++i) F[i] = (double)FL[i];
}
// RHS in long double from double inputs, with the scheme mask frozen from the double sensor
// (rounding-floor reference for the FP64 parity bar).
void rhs_ld(const f3d_params* p, double* Rout, const double* Q, const void* const* mv, const double* J) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  std::vector<double> phid(nc, 0.0);
  shock_sensor<double>(p, phid.data(), Q);
  typedef long double L;
  std::vector<L> QL(Q, Q + 6 * nc), phiL(phid.begin(), phid.end());
  std::vector<std::vector<L>> mL(9);
  const L* mp[9];
  for (int q = 0; q < 9; ++q) {
    const double* src = (const double*)mv[q];
    mL[q].assign(src, src + nc);
    mp[q] = mL[q].data();
  }
  Scratch<L> sc(p);
  std::vector<L> s1(nc), s2(nc), s3(nc);
  metric_norms<L>(p, mp, s1.data(), s2.data(), s3.data());
  flow_advance<L>(p, QL.data(), sc.Fp.data(), sc.Fm.data(), sc.Fx.data(), sc.Fy.data(), sc.Fz.data(), sc.Fvx.data(),
                  sc.Fvy.data(), sc.Fvz.data(), s1.data(), s2.data(), s3.data(), mp, J, phiL.data());
  std::fill(Rout, Rout + nc * 5, 0.0);
  const long NG = g.NG;
  const L c1 = L(1) / L(12), c2 = L(-2) / L(3);
  for (long k = 1; k <= g.Nzp; ++k)
    for (long j = 1; j <= g.Nyp; ++j)
      for (long i = 1; i <= g.Nxp; ++i) {
        const L Jc = J[g.cell.at(i + NG, j + NG, k + NG)];
        for (long n = 1; n <= 5; ++n) {
          L s = sc.Fx[g.fx.at(i, j, k, n)] - sc.Fx[g.fx.at(i + 1, j, k, n)] + sc.Fy[g.fy.at(i, j, k, n)] -
                sc.Fy[g.fy.at(i, j + 1, k, n)] + sc.Fz[g.fz.at(i, j, k, n)] - sc.Fz[g.fz.at(i, j, k + 1, n)];
          if (n >= 2 && p->viscous) {
            const int v = (int)n - 1;
            auto F = [&](const std::vector<L>& A, long ii, long jj, long kk) { return A[g.fv.at(ii, jj, kk, v)]; };
            s += c1 * (F(sc.Fvx, i, j + 2, k + 2) - F(sc.Fvx, i + 4, j + 2, k + 2)) + c2 * (F(sc.Fvx, i + 1, j + 2, k + 2) - F(sc.Fvx, i + 3, j + 2, k + 2));
            s += c1 * (F(sc.Fvy, i + 2, j, k + 2) - F(sc.Fvy, i + 2, j + 4, k + 2)) + c2 * (F(sc.Fvy, i + 2, j + 1, k + 2) - F(sc.Fvy, i + 2, j + 3, k + 2));
            s += c1 * (F(sc.Fvz, i + 2, j + 2, k) - F(sc.Fvz, i + 2, j + 2, k + 4)) + c2 * (F(sc.Fvz, i + 2, j + 2, k + 1) - F(sc.Fvz, i + 2, j + 2, k + 3));
          }
          Rout[g.cell.at(i + NG, j + NG, k + NG, n)] = (double)(s * Jc);
        }
      }
}
// Forward-mode (dual-number) pieces of the FP64 path with the scheme mask frozen at the primal
// state: the CPU forward-mode reference the adjoint (VJP) is checked against.
struct DualMetrics {
  std::vector<std::vector<Dual>> a;
  const Dual* mp[9];
  std::vector<Dual> s1, s2, s3;
  DualMetrics(const f3d_params* p, const void* const* mv) : a(9) {
    Grid g(p);
    size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
    for (int q = 0; q < 9; ++q) {
      const double* src = (const double*)mv[q];
      a[q].assign(src, src + nc);
      mp[q] = a[q].data();
    }
    s1.resize(nc); s2.resize(nc); s3.resize(nc);
    metric_norms<Dual>(p, mp, s1.data(), s2.data(), s3.data());
  }
};

// R = J * (sum_d dF_d + viscous divergence) on the interior (U layout, ghosts 0), in Dual; the
// sensor is evaluated on the value part only.
void rhs_dual(const f3d_params* p, const DualMetrics& dm, Dual* R, const Dual* QD, const double* J) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  std::vector<double> qv(6 * nc), phid(nc, 0.0);
  for (size_t i = 0; i < 6 * nc; ++i) qv[i] = QD[i].v;
  shock_sensor<double>(p, phid.data(), qv.data());
  std::vector<Dual> phiD(phid.begin(), phid.end());
  std::vector<Dual> Qc(QD, QD + 6 * nc);
  Scratch<Dual> sc(p);
  flow_advance<Dual>(p, Qc.data(), sc.Fp.data(), sc.Fm.data(), sc.Fx.data(), sc.Fy.data(), sc.Fz.data(),
                     sc.Fvx.data(), sc.Fvy.data(), sc.Fvz.data(), dm.s1.data(), dm.s2.data(), dm.s3.data(), dm.mp, J,
                     phiD.data());
  std::fill(R, R + nc * 5, Dual(0.0));
  const long NG = g.NG;
  const double c1 = 1.0 / 12, c2 = -2.0 / 3;
  for (long k = 1; k <= g.Nzp; ++k)
    for (long j = 1; j <= g.Nyp; ++j)
      for (long i = 1; i <= g.Nxp; ++i) {
        const double Jc = J[g.cell.at(i + NG, j + NG, k + NG)];
        for (long n = 1; n <= 5; ++n) {
          Dual s = sc.Fx[g.fx.at(i, j, k, n)] - sc.Fx[g.fx.at(i + 1, j, k, n)] + sc.Fy[g.fy.at(i, j, k, n)] -
                   sc.Fy[g.fy.at(i, j + 1, k, n)] + sc.Fz[g.fz.at(i, j, k, n)] - sc.Fz[g.fz.at(i, j, k + 1, n)];
          if (n >= 2 && p->viscous) {
            const int v = (int)n - 1;
            auto F = [&](const std::vector<Dual>& A, long ii, long jj, long kk) { return A[g.fv.at(ii, jj, kk, v)]; };
            s += c1 * (F(sc.Fvx, i, j + 2, k + 2) - F(sc.Fvx, i + 4, j + 2, k + 2)) +
                 c2 * (F(sc.Fvx, i + 1, j + 2, k + 2) - F(sc.Fvx, i + 3, j + 2, k + 2));
            s += c1 * (F(sc.Fvy, i + 2, j, k + 2) - F(sc.Fvy, i + 2, j + 4, k + 2)) +
                 c2 * (F(sc.Fvy, i + 2, j + 1, k + 2) - F(sc.Fvy, i + 2, j + 3, k + 2));
            s += c1 * (F(sc.Fvz, i + 2, j + 2, k) - F(sc.Fvz, i + 2, j + 2, k + 4)) +
                 c2 * (F(sc.Fvz, i + 2, j + 2, k + 1) - F(sc.Fvz, i + 2, j + 2, k + 3));
          }
          R[g.cell.at(i + NG, j + NG, k + NG, n)] = s * Jc;
        }
      }
}

std::vector<Dual> duals(const double* v, const double* d, size_t n) {
  std::vector<Dual> out(n);
  for (size_t i = 0; i < n; ++i) out[i] = Dual(v[i], d ? d[i] : 0.0);
  return out;
}

void split(const std::vector<Dual>& a, double* v, double* d) {
  for (size_t i = 0; i < a.size(); ++i) {
    if (v) v[i] = a[i].v;
    if (d) d[i] = a[i].d;
  }
}

// Ghost procedure of an RK stage (exchange, then fill; solver.jl:439-443) in Dual.
void ghosts_dual(const f3d_params* p, Dual* QD, const Dual* inflowD, const Dual* qinfD, const Dual* TwD) {
  exchange_ghost<Dual>(p, QD, 6);
  fill_ghost<Dual, Dual>(p, QD, nullptr, inflowD, qinfD, TwD);
}

// Jacobian-vector product of the RHS with respect to Q (all cells, ghosts included), mask frozen.
void rhs_jvp(const f3d_params* p, double* R, double* dR, const double* Q, const double* dQ, const void* const* mv,
             const double* J) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  DualMetrics dm(p, mv);
  std::vector<Dual> QD = duals(Q, dQ, 6 * nc), RD(5 * nc);
  rhs_dual(p, dm, RD.data(), QD.data(), J);
  split(RD, R, dR);
}

// nsteps RK3 steps of the FP64 path (solver.jl:416-444) in Dual, from Q0 = ghosts(c2Prim(U0)),
// with the objective sum_n dt <W, Q^{n+1}> over interior cells and all 6 primitive variables.
// Tangents: dU (interior), dinflow (inflow layout), dqinf[6], dTw; any may be null (zero).
void time_step_jvp(const f3d_params* p, int nsteps, double* U, double* dU, const double* inflow,
                   const double* dinflow, const double* dqinf, const double* dTw, const void* const* mv,
                   const double* J, double dt, const double* W, double* obj, double* dobj) {
  Grid g(p);
  const size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  DualMetrics dm(p, mv);
  std::vector<Dual> UD = duals(U, dU, 5 * nc), UnD(5 * nc), QD(6 * nc), RD(5 * nc);
  const size_t nin = (size_t)g.Nyt * g.Nzt * 6;
  std::vector<Dual> inflowD;
  if (inflow) inflowD = duals(inflow, dinflow, nin);
  Dual qinfD[6], TwD(p->T_wall, dTw ? *dTw : 0.0);
  for (int n = 0; n < 6; ++n) qinfD[n] = Dual(p->q_inf[n], dqinf ? dqinf[n] : 0.0);
  const Dual* inP = inflow ? inflowD.data() : nullptr;
  c2prim<Dual, Dual>(p, UD.data(), QD.data());
  ghosts_dual(p, QD.data(), inP, qinfD, &TwD);
  Dual J_obj(0.0);
  const long NG = g.NG;
  for (int t = 0; t < nsteps; ++t) {
    UnD = UD;
    for (int KRK = 1; KRK <= 3; ++KRK) {
      rhs_dual(p, dm, RD.data(), QD.data(), J);
      const double a = KRK == 1 ? 1.0 : (KRK == 2 ? 0.25 : 2.0 / 3.0);
      const double b = KRK == 1 ? 0.0 : (KRK == 2 ? 0.75 : 1.0 / 3.0);
      for (long n = 1; n <= 5; ++n)
        for (long k = NG + 1; k <= g.Nzp + NG; ++k)
          for (long j = NG + 1; j <= g.Nyp + NG; ++j)
            for (long i = NG + 1; i <= g.Nxp + NG; ++i) {
              const size_t o = g.cell.at(i, j, k, n);
              const Dual u1 = UD[o] + RD[o] * dt;
              UD[o] = KRK == 1 ? u1 : u1 * a + UnD[o] * b;
            }
      c2prim<Dual, Dual>(p, UD.data(), QD.data());
      ghosts_dual(p, QD.data(), inP, qinfD, &TwD);
    }
    if (W) {
      Dual acc(0.0);
      for (long n = 1; n <= 6; ++n)
        for (long k = NG + 1; k <= g.Nzp + NG; ++k)
          for (long j = NG + 1; j <= g.Nyp + NG; ++j)
            for (long i = NG + 1; i <= g.Nxp + NG; ++i) {
              const size_t o = g.cell.at(i, j, k, n);
              if (W[o] != 0.0) acc += QD[o] * W[o];
            }
      J_obj += acc * dt;
    }
  }
  split(UD, U, dU);
  if (obj) *obj = J_obj.v;
  if (dobj) *dobj = J_obj.d;
}
}  // namespace

OR_API void oracle_rhs_jvp_f64(const f3d_params* p, double* R, double* dR, const double* Q, const double* dQ,
                               const void* const* m, const double* J) {
  rhs_jvp(p, R, dR, Q, dQ, m, J);
}

OR_API void oracle_time_step_jvp_f64(const f3d_params* p, int nsteps, double* U, double* dU, const double* inflow,
                                     const double* dinflow, const double* dqinf, const double* dTw,
                                     const void* const* m, const double* J, double dt, const double* W, double* obj,
                                     double* dobj) {
  time_step_jvp(p, nsteps, U, dU, inflow, dinflow, dqinf, dTw, m, J, dt, W, obj, dobj);
}

// Forward-mode c2Prim (interior) and stage ghost procedure (exchange + fill), FP64.
OR_API void oracle_c2prim_jvp_f64(const f3d_params* p, const double* U, const double* dU, double* Q, double* dQ) {
  Grid g(p);
  const size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  std::vector<Dual> UD = duals(U, dU, 5 * nc), QD = duals(Q, dQ, 6 * nc);
  c2prim<Dual, Dual>(p, UD.data(), QD.data());
  split(QD, Q, dQ);
}

OR_API void oracle_ghosts_jvp_f64(const f3d_params* p, double* Q, double* dQ, const double* inflow,
                                  const double* dinflow, const double* dqinf, const double* dTw) {
  Grid g(p);
  const size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  std::vector<Dual> QD = duals(Q, dQ, 6 * nc), inflowD;
  if (inflow) inflowD = duals(inflow, dinflow, (size_t)g.Nyt * g.Nzt * 6);
  Dual qinfD[6], TwD(p->T_wall, dTw ? *dTw : 0.0);
  for (int n = 0; n < 6; ++n) qinfD[n] = Dual(p->q_inf[n], dqinf ? dqinf[n] : 0.0);
  ghosts_dual(p, QD.data(), inflow ? inflowD.data() : nullptr, qinfD, &TwD);
  split(QD, Q, dQ);
}

// advect_c in Dual: face fluxes and their tangents from tangents of Fp, Fm and Q (mask from phi).
OR_API void oracle_advect_c_jvp_f64(const f3d_params* p, int d, double* F, double* dF, const double* phi,
                                    const double* S, const double* Fp, const double* dFp, const double* Fm,
                                    const double* dFm, const double* Q, const double* dQ, const double* Ax,
                                    const double* Ay, const double* Az) {
  Grid g(p);
  const size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  const Ext& fe = d == 0 ? g.fx : (d == 1 ? g.fy : g.fz);
  const size_t nf = (size_t)fe.n1 * fe.n2 * fe.n3 * 5;
  std::vector<Dual> FpD = duals(Fp, dFp, 5 * nc), FmD = duals(Fm, dFm, 5 * nc), QD = duals(Q, dQ, 6 * nc);
  std::vector<Dual> phiD = duals(phi, nullptr, nc), SD = duals(S, nullptr, nc), AxD = duals(Ax, nullptr, nc),
                    AyD = duals(Ay, nullptr, nc), AzD = duals(Az, nullptr, nc), FD(nf);
  advect_c<Dual>(p, d, FD.data(), phiD.data(), SD.data(), FpD.data(), FmD.data(), QD.data(), AxD.data(), AyD.data(),
                 AzD.data());
  split(FD, F, dF);
}

OR_API void oracle_rhs_ld(const f3d_params* p, double* R, const double* Q, const void* const* m, const double* J) {
  rhs_ld(p, R, Q, m, J);
}

OR_API void oracle_rhs_absterms_f64(const f3d_params* p, double* A, const double* Q, const void* const* m,
                                    const double* J) {
  rhs_absterms<double>(p, A, Q, m, J);
}
OR_API void oracle_rhs_absterms_f32b(const f3d_params* p, double* A, const float* Q, const void* const* m,
                                     const double* J) {
  rhs_absterms<float>(p, A, Q, m, J);
}
OR_API void oracle_advect_c_ld(const f3d_params* p, int d, double* F, const double* phi, const double* S,
                               const double* Fp, const double* Fm, const double* Q, const double* Ax,
                               const double* Ay, const double* Az) {
  advect_ld(p, d, phi, S, Fp, Fm, Q, Ax, Ay, Az, F);
}

// ---------------------------------------------------------------- finite-volume branch (FVM.jl)
namespace {
struct Coords {
  const double* x;
  const double* y;
  const double* z;
  explicit Coords(const void* const* c) : x((const double*)c[0]), y((const double*)c[1]), z((const double*)c[2]) {}
};

template <typename T>
void rhs_fvm_ref(const f3d_params* p, double* R, const T* Q, const double* U, const void* const* mv, const double* J,
                 const void* const* cv, T* phi_out) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  Scratch<T> sc(p);
  M9<T> m(mv);
  Coords c(cv);
  std::vector<T> s1(nc), s2(nc), s3(nc);
  metric_norms<T>(p, m.m, s1.data(), s2.data(), s3.data());
  shock_sensor<T>(p, sc.phi.data(), Q);
  flow_advance_fvm<T>(p, U, Q, sc.Fx.data(), sc.Fy.data(), sc.Fz.data(), sc.Fvx.data(), sc.Fvy.data(), sc.Fvz.data(),
                      s1.data(), s2.data(), s3.data(), m.m, J, c.x, c.y, c.z, sc.phi.data());
  std::fill(R, R + nc * 5, 0.0);
  div<T>(p, R, sc.Fx.data(), sc.Fy.data(), sc.Fz.data(), sc.Fvx.data(), sc.Fvy.data(), sc.Fvz.data(), 1.0, J);
  if (phi_out) std::copy(sc.phi.begin(), sc.phi.end(), phi_out);
}

template <typename T>
void time_step_fvm_ref(const f3d_params* p, int nsteps, double* U, double* Un, T* Q, const void* const* mv,
                       const double* J, const void* const* cv, double dt, const T* inflow) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  Scratch<T> sc(p);
  M9<T> m(mv);
  Coords c(cv);
  std::vector<T> s1(nc), s2(nc), s3(nc);
  metric_norms<T>(p, m.m, s1.data(), s2.data(), s3.data());
  for (int t = 1; t <= nsteps; ++t) {
    for (int KRK = 1; KRK <= 3; ++KRK)
      rk3_stage_fvm<T>(p, sc, KRK, U, Un, Q, s1.data(), s2.data(), s3.data(), m.m, J, c.x, c.y, c.z, dt, inflow);
    if (p->filtering && t % p->filtering_interval == 0) filter_block<T>(p, U, Un, Q, sc.fsc.data());
  }
}

// Finite-volume RHS in long double from double inputs, the scheme mask taken from the double sensor
// (rounding-floor reference of the FP64 parity bar, as rhs_ld).
void rhs_fvm_ld(const f3d_params* p, double* Rout, const double* Q, const double* U, const void* const* mv,
                const double* J, const void* const* cv) {
  typedef long double L;
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  std::vector<double> phid(nc, 0.0);
  shock_sensor<double>(p, phid.data(), Q);
  std::vector<L> QL(Q, Q + 6 * nc), phiL(phid.begin(), phid.end());
  std::vector<std::vector<L>> mL(9);
  const L* mp[9];
  for (int q = 0; q < 9; ++q) {
    const double* src = (const double*)mv[q];
    mL[q].assign(src, src + nc);
    mp[q] = mL[q].data();
  }
  Coords c(cv);
  Scratch<L> sc(p);
  std::vector<L> s1(nc), s2(nc), s3(nc);
  metric_norms<L>(p, mp, s1.data(), s2.data(), s3.data());
  flow_advance_fvm<L>(p, U, QL.data(), sc.Fx.data(), sc.Fy.data(), sc.Fz.data(), sc.Fvx.data(), sc.Fvy.data(),
                      sc.Fvz.data(), s1.data(), s2.data(), s3.data(), mp, J, c.x, c.y, c.z, phiL.data());
  std::fill(Rout, Rout + nc * 5, 0.0);
  const long NG = g.NG;
  const L c1 = L(1) / L(12), c2 = L(-2) / L(3);
  for (long k = 1; k <= g.Nzp; ++k)
    for (long j = 1; j <= g.Nyp; ++j)
      for (long i = 1; i <= g.Nxp; ++i) {
        const L Jc = J[g.cell.at(i + NG, j + NG, k + NG)];
        for (long n = 1; n <= 5; ++n) {
          L s = sc.Fx[g.fx.at(i, j, k, n)] - sc.Fx[g.fx.at(i + 1, j, k, n)] + sc.Fy[g.fy.at(i, j, k, n)] -
                sc.Fy[g.fy.at(i, j + 1, k, n)] + sc.Fz[g.fz.at(i, j, k, n)] - sc.Fz[g.fz.at(i, j, k + 1, n)];
          if (n >= 2 && p->viscous) {
            const int v = (int)n - 1;
            auto F = [&](const std::vector<L>& A, long ii, long jj, long kk) { return A[g.fv.at(ii, jj, kk, v)]; };
            s += c1 * (F(sc.Fvx, i, j + 2, k + 2) - F(sc.Fvx, i + 4, j + 2, k + 2)) + c2 * (F(sc.Fvx, i + 1, j + 2, k + 2) - F(sc.Fvx, i + 3, j + 2, k + 2));
            s += c1 * (F(sc.Fvy, i + 2, j, k + 2) - F(sc.Fvy, i + 2, j + 4, k + 2)) + c2 * (F(sc.Fvy, i + 2, j + 1, k + 2) - F(sc.Fvy, i + 2, j + 3, k + 2));
            s += c1 * (F(sc.Fvz, i + 2, j + 2, k) - F(sc.Fvz, i + 2, j + 2, k + 4)) + c2 * (F(sc.Fvz, i + 2, j + 2, k + 1) - F(sc.Fvz, i + 2, j + 2, k + 3));
          }
          Rout[g.cell.at(i + NG, j + NG, k + NG, n)] = (double)(s * Jc);
        }
      }
}

// Magnitude of the terms of the finite-volume RHS per cell and variable: J * sum_d (|F_d| on the
// cell's two faces) for the inviscid part (each face flux enters one cell difference), plus the
// viscous terms of rhs_absterms.
template <typename T>
void rhs_absterms_fvm(const f3d_params* p, double* A, const T* Q, const double* U, const void* const* mv,
                      const double* J, const void* const* cv) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  Scratch<T> sc(p);
  M9<T> m(mv);
  Coords c(cv);
  std::vector<T> s1(nc), s2(nc), s3(nc);
  metric_norms<T>(p, m.m, s1.data(), s2.data(), s3.data());
  shock_sensor<T>(p, sc.phi.data(), Q);
  const bool eigen = p->fvm == F3D_FVM_EIGEN;
  fvm_reconstruct<T>(p, 0, eigen, Q, U, sc.phi.data(), s1.data(), sc.Fx.data(), c.x, c.y, c.z);
  fvm_reconstruct<T>(p, 1, eigen, Q, U, sc.phi.data(), s2.data(), sc.Fy.data(), c.x, c.y, c.z);
  fvm_reconstruct<T>(p, 2, eigen, Q, U, sc.phi.data(), s3.data(), sc.Fz.data(), c.x, c.y, c.z);
  std::fill(A, A + nc * 5, 0.0);
  const long NG = g.NG;
  for (long k = 1; k <= g.Nzp; ++k)
    for (long j = 1; j <= g.Nyp; ++j)
      for (long i = 1; i <= g.Nxp; ++i) {
        const double Jc = std::fabs(J[g.cell.at(i + NG, j + NG, k + NG)]);
        for (long n = 1; n <= 5; ++n) {
          auto a = [](T v) { return std::fabs((double)v); };
          const double t = a(sc.Fx[g.fx.at(i, j, k, n)]) + a(sc.Fx[g.fx.at(i + 1, j, k, n)]) +
                           a(sc.Fy[g.fy.at(i, j, k, n)]) + a(sc.Fy[g.fy.at(i, j + 1, k, n)]) +
                           a(sc.Fz[g.fz.at(i, j, k, n)]) + a(sc.Fz[g.fz.at(i, j, k + 1, n)]);
          A[g.cell.at(i + NG, j + NG, k + NG, n)] = Jc * t;
        }
      }
  if (p->viscous) viscous_absterms<T>(p, A, Q, m.m, J);
}
}  // namespace

OR_API void oracle_fvm_reconstruct_f64(const f3d_params* p, int d, int eigen, const double* Q, const double* U,
                                       const double* phi, const double* S, double* F, const double* x,
                                       const double* y, const double* z, double* UL, double* UR, double* geo) {
  fvm_reconstruct<double>(p, d, eigen != 0, Q, U, phi, S, F, x, y, z, UL, UR, geo);
}
// Face fluxes in long double from double inputs (rounding-floor reference of the FP64 face bar).
OR_API void oracle_fvm_reconstruct_ld(const f3d_params* p, int d, int eigen, const double* Q, const double* U,
                                      const double* phi, const double* S, double* F, const double* x,
                                      const double* y, const double* z) {
  Grid g(p);
  size_t nc = (size_t)g.Nxt * g.Nyt * g.Nzt;
  std::vector<long double> QL(Q, Q + 6 * nc), phiL(phi, phi + nc), SL(S, S + nc);
  size_t nf = (size_t)(g.Nxp + (d == 0)) * (g.Nyp + (d == 1)) * (g.Nzp + (d == 2)) * 5;
  std::vector<long double> FL(nf);
  fvm_reconstruct<long double>(p, d, eigen != 0, QL.data(), U, phiL.data(), SL.data(), FL.data(), x, y, z);
  for (size_t i = 0; i < nf; ++i) F[i] = (double)FL[i];
}
OR_API void oracle_fvm_reconstruct_f32b(const f3d_params* p, int d, int eigen, const float* Q, const double* U,
                                        const float* phi, const float* S, float* F, const double* x,
                                        const double* y, const double* z, float* UL, float* UR, float* geo) {
  fvm_reconstruct<float>(p, d, eigen != 0, Q, U, phi, S, F, x, y, z, UL, UR, geo);
}
// Face geometry only (nx, ny, nz, Area in the face layout of direction d).
OR_API void oracle_fvm_geometry_f64(const f3d_params* p, int d, const double* x, const double* y, const double* z,
                                    double* geo) {
  Grid g(p);
  const Dir D = make_dir(d);
  const Ext& fe = d == 0 ? g.fx : (d == 1 ? g.fy : g.fz);
  const double* const X[3] = {x, y, z};
  for (long k = g.NG + 1 - D.dk; k <= g.Nzp + g.NG; ++k)
    for (long j = g.NG + 1 - D.dj; j <= g.Nyp + g.NG; ++j)
      for (long i = g.NG + 1 - D.di; i <= g.Nxp + g.NG; ++i) {
        double n[3], A;
        fvm_face_geometry<double>(g, d, X, i, j, k, n, &A);
        const long fi = i - g.NG + D.di, fj = j - g.NG + D.dj, fk = k - g.NG + D.dk;
        for (int m = 0; m < 3; ++m) geo[fe.at(fi, fj, fk, m + 1)] = n[m];
        geo[fe.at(fi, fj, fk, 4)] = A;
      }
}
// HLLC flux of nf independent faces: UL, UR [5][nf], n [3][nf], F [5][nf] (struct of arrays).
OR_API void oracle_hllc_f64(const f3d_params* p, long nf, const double* UL, const double* UR, const double* n,
                            double* F) {
  for (long f = 0; f < nf; ++f) {
    double l[5], r[5], o[5];
    for (int q = 0; q < 5; ++q) {
      l[q] = UL[q * nf + f];
      r[q] = UR[q * nf + f];
    }
    hllc_flux<double>(p, l, r, n[f], n[nf + f], n[2 * nf + f], o);
    for (int q = 0; q < 5; ++q) F[q * nf + f] = o[q];
  }
}
OR_API void oracle_rhs_fvm_f64(const f3d_params* p, double* R, const double* Q, const double* U, const void* const* m,
                               const double* J, const void* const* coords, double* phi) {
  rhs_fvm_ref<double>(p, R, Q, U, m, J, coords, phi);
}
OR_API void oracle_rhs_fvm_f32b(const f3d_params* p, double* R, const float* Q, const double* U, const void* const* m,
                                const double* J, const void* const* coords, float* phi) {
  rhs_fvm_ref<float>(p, R, Q, U, m, J, coords, phi);
}
OR_API void oracle_rhs_fvm_ld(const f3d_params* p, double* R, const double* Q, const double* U, const void* const* m,
                              const double* J, const void* const* coords) {
  rhs_fvm_ld(p, R, Q, U, m, J, coords);
}
OR_API void oracle_rhs_absterms_fvm_f64(const f3d_params* p, double* A, const double* Q, const double* U,
                                        const void* const* m, const double* J, const void* const* coords) {
  rhs_absterms_fvm<double>(p, A, Q, U, m, J, coords);
}
OR_API void oracle_rhs_absterms_fvm_f32b(const f3d_params* p, double* A, const float* Q, const double* U,
                                         const void* const* m, const double* J, const void* const* coords) {
  rhs_absterms_fvm<float>(p, A, Q, U, m, J, coords);
}
OR_API void oracle_time_step_fvm_f64(const f3d_params* p, int nsteps, double* U, double* Un, double* Q,
                                     const void* const* m, const double* J, const void* const* coords, double dt,
                                     const double* inflow) {
  time_step_fvm_ref<double>(p, nsteps, U, Un, Q, m, J, coords, dt, inflow);
}
OR_API void oracle_time_step_fvm_f32b(const f3d_params* p, int nsteps, double* U, double* Un, float* Q,
                                      const void* const* m, const double* J, const void* const* coords, double dt,
                                      const float* inflow) {
  time_step_fvm_ref<float>(p, nsteps, U, Un, Q, m, J, coords, dt, inflow);
}

// WENO7 smoothness indicators of one 7-point stencil: form 0 the reference's quadratic forms in V
// (schemes.jl:848-869, FVM.jl:127-135), form 1 the first-difference forms of the float path.
OR_API void oracle_weno7_is_f64(int form, const double* V, double* Is) {
  if (form == 1) {
    weno7_is_diff<double>(V, Is);
    return;
  }
  const double V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3], V5 = V[4], V6 = V[5], V7 = V[6];
  Is[0] = V1 * (547 * V1 - 3882 * V2 + 4642 * V3 - 1854 * V4) + V2 * (7043 * V2 - 17246 * V3 + 7042 * V4) +
          V3 * (11003 * V3 - 9402 * V4) + V4 * (2107 * V4);
  Is[1] = V2 * (267 * V2 - 1642 * V3 + 1602 * V4 - 494 * V5) + V3 * (2843 * V3 - 5966 * V4 + 1922 * V5) +
          V4 * (3443 * V4 - 2522 * V5) + V5 * (547 * V5);
  Is[2] = V3 * (547 * V3 - 2522 * V4 + 1922 * V5 - 494 * V6) + V4 * (3443 * V4 - 5966 * V5 + 1602 * V6) +
          V5 * (2843 * V5 - 1642 * V6) + V6 * (267 * V6);
  Is[3] = V4 * (2107 * V4 - 9402 * V5 + 7042 * V6 - 1854 * V7) + V5 * (11003 * V5 - 17246 * V6 + 4642 * V7) +
          V6 * (7043 * V6 - 3882 * V7) + V7 * (547 * V7);
}

// One scalar WENO reconstruction, exactly as the face loops call it: kind 0 the finite-difference
// WENO7 weno7(V, s0, eps) (schemes.jl:847-879), 1 the finite-difference WENO5 weno5(V[0..4], s0, eps)
// (schemes.jl:958-972), 2 the finite-volume WENO7 fvm_weno7(V, ss) (FVM.jl:116-147, its own eps
// 1e-10; `eps` unused), 3 the finite-volume WENO5 fvm_weno5(V, ss) on V[1..5] (FVM.jl:151-201, eps
// 1e-8).  form 0 double, 1 long double, 2 dual number (value part; kinds 0-1 only), 3 float.  The
// finite-difference kinds return the weighted sum of the unscaled candidates (the face loops multiply
// by 1/12 or 1/6 afterwards); the finite-volume kinds return the face value.
OR_API double oracle_weno_scalar(int kind, int form, const double* V, double s0, double eps) {
  auto run = [&](auto tag) -> double {
    using T = decltype(tag);
    T v[7];
    for (int s = 0; s < 7; ++s) v[s] = (T)V[s];
    switch (kind) {
      case 0: return (double)weno7<T>(v, (T)s0, (T)eps);
      case 1: return (double)weno5<T>(v, (T)s0, (T)eps);
      default: return std::nan("");
    }
  };
  auto run_fvm = [&](auto tag) -> double {
    using T = decltype(tag);
    T v[7];
    for (int s = 0; s < 7; ++s) v[s] = (T)V[s];
    return (double)(kind == 2 ? fvm_weno7<T>(v, (T)s0) : fvm_weno5<T>(v, (T)s0));
  };
  if (kind < 0 || kind > 3) return std::nan("");
  if (kind >= 2) {
    if (form == 0) return run_fvm(0.0);
    if (form == 1) return run_fvm((long double)0);
    if (form == 3) return run_fvm(0.f);
    return std::nan("");
  }
  if (form == 0) return run(0.0);
  if (form == 1) return run((long double)0);
  if (form == 2) return run(Dual(0.0));
  if (form == 3) return run(0.f);
  return std::nan("");
}

OR_API int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
