// Discrete-adjoint kernels of the ns step instantiated for float face arithmetic.
#include "gd_adj_impl.hpp"

namespace gd {

hipError_t launch_adj_slot_geo_f32(const MeshView& m, void* sgeo, hipStream_t s) {
  return adj::launch_slot_geo<float>(m, static_cast<float*>(sgeo), s);
}
hipError_t launch_adj_cell_f32(const AdjStepArgs& a, hipStream_t s) {
  return adj::launch_cell<float>(adj::kargs<float>(a), s);
}
hipError_t launch_adj_state_f32(const AdjStepArgs& a, bool gather, hipStream_t s) {
  return adj::launch_state<float>(adj::kargs<float>(a), gather, s);
}

}  // namespace gd
