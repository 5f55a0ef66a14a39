// Host-side internals shared by the translation units of libf3d_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "f3d_device.hpp"

namespace f3d {

template <typename T>
struct MetricPtrs {
  const T* m[9];
};

// Thread-local message of the last failure (f3d_last_error).
void set_error(const char* fmt, ...);
const char* last_error();

// Status of the launches just queued (hipGetLastError), as an f3d status code.
int check_launch(const char* what);

#define F3D_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::f3d::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                       __LINE__);                                                       \
      return (int)e_;                                                                   \
    }                                                                                   \
  } while (0)

KP make_kp(const f3d_params* p);

// RAII bracket of one kernel launch with hipEvents when timing is enabled (f3d_timing_*).
class TimedScope {
 public:
  TimedScope(const char* name, hipStream_t s);
  ~TimedScope();

 private:
  int slot_;
  hipStream_t s_;
};

// kernel-by-kernel launchers (f3d_kernels.hip)
template <typename T> int launch_shock_sensor(const KP&, T* phi, const T* Q, hipStream_t);
template <typename T> int launch_scheme_mask(const KP&, uint8_t*, uint8_t*, uint8_t*, const T* phi, hipStream_t);
template <typename T>
int launch_flux_split(const KP&, const T* Q, T* Fp, T* Fm, const T* S, const T* Ax, const T* Ay, const T* Az,
                      hipStream_t);
// split method kp.split; advect: characteristic when kp.character, else component-wise on NV fluxes
template <typename T>
int launch_advect(const KP&, int d, T* F, const T* phi, const T* S, const T* Fp, const T* Fm, const T* Q,
                  const T* Ax, const T* Ay, const T* Az, hipStream_t, int NV = 5);
template <typename T>
int launch_viscous(const KP&, T* Fvx, T* Fvy, T* Fvz, const T* Q, const MetricPtrs<T>& m, const double* J,
                   hipStream_t);
template <typename T>
int launch_div(const KP&, double* U, const T* Fx, const T* Fy, const T* Fz, const T* Fvx, const T* Fvy,
               const T* Fvz, double dt, const double* J, hipStream_t, const double* ldt = nullptr);
// Filters (utils.jl:102-316): the sensor along d (interior + one ghost layer each side along d), one
// direction's filter of the interior (sc != nullptr: the nonlinear filter, else the linear one), and
// the filtering block of the time loop (solver.jl:446-470: per direction Un <- U, sensor, filter).
template <typename T>
int launch_pre_sensor(const KP&, int d, double* sc, const T* Q, double rth, hipStream_t);
int launch_filter(const KP&, int d, double* U, const double* Un, const double* sc, double s0, hipStream_t);
template <typename T>
int launch_filter_block(const KP&, double* U, double* Un, const T* Q, double* sc, int nonlinear, double rth,
                        double s0, hipStream_t);
// compute_dt (utils.jl:73-100) into the interior cells of dt; m != nullptr: the sizes from the metric
// rows instead of S1..S3.
template <typename T>
int launch_compute_dt(const KP&, double* dt, const T* Q, const double* J, const T* S1, const T* S2, const T* S3,
                      const MetricPtrs<T>* m, hipStream_t);
int launch_lin_comb(const KP&, double* U, const double* Un, int NV, double a, double b, hipStream_t);
template <typename T> int launch_c2prim(const KP&, const double* U, T* Q, hipStream_t);
template <typename T> int launch_prim2c(const KP&, double* U, const T* Q, hipStream_t);
template <typename T> int launch_exchange(const KP&, T* Q, int NV, hipStream_t);
template <typename T> int launch_fill(const KP&, T* Q, double* U, const T* inflow, hipStream_t);

// finite-volume face fluxes of direction d (f3d_fvm.hip), variant kp.fvm; S == nullptr: |grad xi_d|
// from the metric row A (Ax, Ay, Az); geo == nullptr: face geometry from the mesh points x, y, z,
// else from the face geometry array of direction d
template <typename T>
int launch_fvm_face(const KP&, int d, const T* Q, const double* U, const T* phi, const T* S, const T* Ax,
                    const T* Ay, const T* Az, T* F, const double* x, const double* y, const double* z, const T* geo,
                    hipStream_t);
// face geometry arrays (nx, ny, nz, Area) of the three directions
template <typename T>
int launch_fvm_geometry(const KP&, const double* x, const double* y, const double* z, T* const* geo, hipStream_t);

}  // namespace f3d
