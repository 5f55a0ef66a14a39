// Fused RHS / RK3-stage path (placeholder until the sweep kernels land: the chain serves).
#include "f3d_fused.hpp"

namespace f3d {

bool use_fused() { return false; }

template <typename T>
int fused_rhs(const KP&, const WsView<T>&, const T*, const MetricPtrs<T>&, const double*, T*, double*, int, double,
              double*, hipStream_t) {
  set_error("fused RHS not available");
  return F3D_E_ARG;
}
template <typename T>
int fused_stage(const KP&, const WsView<T>&, int, double*, double*, T*, const MetricPtrs<T>&, const double*, double,
                hipStream_t) {
  set_error("fused stage not available");
  return F3D_E_ARG;
}
template int fused_rhs<double>(const KP&, const WsView<double>&, const double*, const MetricPtrs<double>&,
                               const double*, double*, double*, int, double, double*, hipStream_t);
template int fused_rhs<float>(const KP&, const WsView<float>&, const float*, const MetricPtrs<float>&, const double*,
                              float*, double*, int, double, double*, hipStream_t);
template int fused_stage<double>(const KP&, const WsView<double>&, int, double*, double*, double*,
                                 const MetricPtrs<double>&, const double*, double, hipStream_t);
template int fused_stage<float>(const KP&, const WsView<float>&, int, double*, double*, float*,
                                const MetricPtrs<float>&, const double*, double, hipStream_t);
}  // namespace f3d
