// ns-scheme kernels (placeholder until the ns scheme lands)
#include "gd_internal.hpp"
namespace gd {
hipError_t launch_ns_grad(const NsArgs&, hipStream_t) { return hipErrorNotSupported; }
hipError_t launch_ns_flux(const NsArgs&, hipStream_t) { return hipErrorNotSupported; }
}  // namespace gd
