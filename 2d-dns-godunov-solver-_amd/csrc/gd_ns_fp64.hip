// ns-scheme kernels instantiated for double face arithmetic (FP64 step; built with -ffp-contract=off).
#include "gd_ns_impl.hpp"

namespace gd {

static ns::NsKArgs<double> cast(const NsStepArgs& a) { return ns::kargs<double>(a); }

hipError_t launch_ns_prims_f64(int64_t nc, double gamma, const double* U, void* Q, hipStream_t s) {
  return ns::launch_prims<double>(nc, gamma, U, static_cast<double*>(Q), s);
}
hipError_t launch_ns_grad_f64(const NsStepArgs& a, hipStream_t s) { return ns::launch_grad<double>(cast(a), s); }
hipError_t launch_ns_face_f64(const NsStepArgs& a, hipStream_t s) { return ns::launch_face<double>(cast(a), s); }
hipError_t launch_ns_update_f64(const NsStepArgs& a, hipStream_t s) { return ns::launch_update<double>(cast(a), s); }

}  // namespace gd
