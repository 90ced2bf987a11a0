// calcDiffTerm on caller-owned device arrays with no device allocation: its
// intermediates live in one file-scope device array, `scratch`, sized once for
// all meshes up to a fixed maximum.  For callers that cannot allocate device
// memory (work recorded into HIP graphs, tight time loops); the state-based
// gd_calc_diff_term allocates its intermediates with the state instead.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "gd_internal.hpp"

namespace {

// node cluster means [nn][4], first Green-Gauss gradients [nc][6], their node
// means [nn][6]: 48 nc + 80 nn bytes (each part 256 B aligned)
constexpr size_t kScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // the array's address on each device, looked up once

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// The device address of `scratch` on the current device (its host-side name
// is not a device address), looked up on the first call on each device.
int scratch_on_current_device(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return gd::set_error(GD_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return gd::set_error(GD_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return gd::set_error(GD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return GD_OK;
}

}  // namespace

size_t gd_diff_term_scratch_bytes(int64_t n_cells, int64_t n_nodes) {
  return align256((size_t)n_nodes * 4 * sizeof(double)) + align256((size_t)n_cells * 6 * sizeof(double)) +
         align256((size_t)n_nodes * 6 * sizeof(double));
}

int gd_calc_diff_term_noalloc(const gd_mesh* mesh, const gd_params* params, const double* d_U, const double* d_Q,
                              double* d_Udiff, void* stream) {
  if (!mesh || !params || !d_U || !d_Q || !d_Udiff) return gd::set_error(GD_ERR_INVALID, "null argument");
  const gd::Mesh& M = *reinterpret_cast<const gd::Mesh*>(mesh);
  int rc = gd::check_params(&M, params);
  if (rc) return rc;
  if (params->scheme != GD_SCHEME_REF)
    return gd::set_error(GD_ERR_UNSUPPORTED, "calcDiffTerm belongs to the reference scheme");
  const gd::MeshView& v = M.v;
  const size_t need = gd_diff_term_scratch_bytes(v.nc, v.nn);
  if (need > kScratchBytes)
    return gd::set_error(GD_ERR_INVALID, "mesh too large for gd_calc_diff_term_noalloc: needs " +
                                             std::to_string(need) + " bytes of scratch, the maximum is " +
                                             std::to_string(kScratchBytes));
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != M.device) return gd::set_error(GD_ERR_INVALID, "the current device is not the mesh's device");
  void* base = nullptr;
  if ((rc = scratch_on_current_device(&base))) return rc;
  unsigned char* b = static_cast<unsigned char*>(base);
  double* phiN = reinterpret_cast<double*>(b);
  double* g6 = reinterpret_cast<double*>(b + align256((size_t)v.nn * 4 * sizeof(double)));
  double* g6N = reinterpret_cast<double*>(b + align256((size_t)v.nn * 4 * sizeof(double)) +
                                          align256((size_t)v.nc * 6 * sizeof(double)));
  hipStream_t s = static_cast<hipStream_t>(stream);
  // every intermediate is written in full before it is read: nothing to clear
  hipError_t e;
  if ((e = gd::launch_ref_phi(v, d_U, phiN, s)) != hipSuccess ||
      (e = gd::launch_ref_gg1(v, phiN, g6, s)) != hipSuccess ||
      (e = gd::launch_ref_gg_nodes(v, g6, g6N, s)) != hipSuccess ||
      (e = gd::launch_ref_diff_term(v, *params, d_Q, g6N, d_Udiff, s)) != hipSuccess)
    return gd::set_error(GD_ERR_HIP, std::string("calcDiffTerm launch: ") + hipGetErrorString(e));
  return GD_OK;
}
