// ns-scheme kernels instantiated for double face arithmetic (FP64 step; built with -ffp-contract=off).
#include "gd_ns_impl.hpp"

namespace gd {

static ns::NsKArgs<double> cast(const NsStepArgs& a) {
  ns::NsKArgs<double> k;
  k.m = a.m;
  k.prm = a.prm;
  k.Uo = a.Uo;
  k.Un = a.Un;
  k.Q = static_cast<const double*>(a.Q);
  k.Qw = static_cast<double*>(a.Qw);
  k.rec = static_cast<double*>(a.rec);
  k.fflux = static_cast<double*>(a.fflux);
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

hipError_t launch_ns_prims_f64(int64_t nc, double gamma, const double* U, void* Q, hipStream_t s) {
  return ns::launch_prims<double>(nc, gamma, U, static_cast<double*>(Q), s);
}
hipError_t launch_ns_grad_f64(const NsStepArgs& a, hipStream_t s) { return ns::launch_grad<double>(cast(a), s); }
hipError_t launch_ns_flux_f64(const NsStepArgs& a, hipStream_t s) {
  hipError_t e = ns::launch_face<double>(cast(a), s);
  return e != hipSuccess ? e : ns::launch_update<double>(cast(a), s);
}
hipError_t launch_ns_face_f64(const NsStepArgs& a, hipStream_t s) { return ns::launch_face<double>(cast(a), s); }
hipError_t launch_ns_update_f64(const NsStepArgs& a, hipStream_t s) { return ns::launch_update<double>(cast(a), s); }

}  // namespace gd
