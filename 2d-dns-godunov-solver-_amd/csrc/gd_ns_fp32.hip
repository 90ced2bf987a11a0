// ns-scheme kernels instantiated for float face arithmetic (FP32-mixed step).
#include "gd_ns_impl.hpp"

namespace gd {

static ns::NsKArgs<float> cast(const NsStepArgs& a) {
  ns::NsKArgs<float> k;
  k.m = a.m;
  k.prm = a.prm;
  k.Uo = a.Uo;
  k.Un = a.Un;
  k.Q = static_cast<const float*>(a.Q);
  k.Qw = static_cast<float*>(a.Qw);
  k.rec = static_cast<float*>(a.rec);
  k.fflux = static_cast<float*>(a.fflux);
  k.acc = a.acc;
  k.part = a.part;
  k.step = a.step;
  k.dt_policy = a.dt_policy;
  k.dt = a.dt;
  k.cfl = a.cfl;
  k.wall_f = a.wall_f;
  k.wall_code = a.wall_code;
  return k;
}

hipError_t launch_ns_prims_f32(int64_t nc, double gamma, const double* U, void* Q, hipStream_t s) {
  return ns::launch_prims<float>(nc, gamma, U, static_cast<float*>(Q), s);
}
hipError_t launch_ns_grad_f32(const NsStepArgs& a, hipStream_t s) { return ns::launch_grad<float>(cast(a), s); }
hipError_t launch_ns_flux_f32(const NsStepArgs& a, hipStream_t s) {
  hipError_t e = ns::launch_face<float>(cast(a), s);
  return e != hipSuccess ? e : ns::launch_update<float>(cast(a), s);
}
hipError_t launch_ns_face_f32(const NsStepArgs& a, hipStream_t s) { return ns::launch_face<float>(cast(a), s); }
hipError_t launch_ns_update_f32(const NsStepArgs& a, hipStream_t s) { return ns::launch_update<float>(cast(a), s); }

}  // namespace gd
