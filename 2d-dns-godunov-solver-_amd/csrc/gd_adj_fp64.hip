// Discrete-adjoint kernels of the ns step instantiated for double face arithmetic.
#include "gd_adj_impl.hpp"

namespace gd {

hipError_t launch_adj_slot_geo_f64(const MeshView& m, void* sgeo, hipStream_t s) {
  return adj::launch_slot_geo<double>(m, static_cast<double*>(sgeo), s);
}
hipError_t launch_adj_cell_f64(const AdjStepArgs& a, hipStream_t s) {
  return adj::launch_cell<double>(adj::kargs<double>(a), s);
}
hipError_t launch_adj_state_f64(const AdjStepArgs& a, bool gather, hipStream_t s) {
  return adj::launch_state<double>(adj::kargs<double>(a), gather, s);
}

}  // namespace gd
