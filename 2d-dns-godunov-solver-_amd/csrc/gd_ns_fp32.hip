// ns-scheme kernels instantiated for float face arithmetic (FP32-mixed step).
#include "gd_ns_impl.hpp"

namespace gd {

static ns::NsKArgs<float> cast(const NsStepArgs& a) { return ns::kargs<float>(a); }

hipError_t launch_ns_prims_f32(int64_t nc, double gamma, const double* U, void* Q, hipStream_t s) {
  return ns::launch_prims<float>(nc, gamma, U, static_cast<float*>(Q), s);
}
hipError_t launch_ns_grad_f32(const NsStepArgs& a, hipStream_t s) { return ns::launch_grad<float>(cast(a), s); }
hipError_t launch_ns_face_f32(const NsStepArgs& a, hipStream_t s) { return ns::launch_face<float>(cast(a), s); }
hipError_t launch_ns_update_f32(const NsStepArgs& a, hipStream_t s) { return ns::launch_update<float>(cast(a), s); }

}  // namespace gd
