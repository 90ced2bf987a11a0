// C ABI of libgodunov: mesh upload and reordering, state management, the
// reference operators and the device-resident time loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <numeric>
#include <string>
#include <vector>

#include "gd_internal.hpp"

using namespace gd;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define GD_HIP(expr)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) return fail(GD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int dev_alloc(std::vector<void*>& allocs, T** p, size_t n, int64_t* bytes = nullptr) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return fail(GD_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  allocs.push_back((void*)*p);
  if (bytes) *bytes += (int64_t)(n * sizeof(T));
  return GD_OK;
}

template <class T>
int upload(std::vector<void*>& allocs, const T** dst, const std::vector<T>& src, int64_t* bytes) {
  T* p = nullptr;
  int rc = dev_alloc(allocs, &p, src.size(), bytes);
  if (rc) return rc;
  if (!src.empty()) {
    hipError_t e = hipMemcpy(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(GD_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  }
  *dst = p;
  return GD_OK;
}

// 1-based column-major reads of the caller's Julia arrays
inline int32_t CI(const int32_t* a, int64_t n, int64_t i0, int k1) { return a[(int64_t)(k1 - 1) * n + i0]; }
inline double CD(const double* a, int64_t n, int64_t i0, int k1) { return a[(int64_t)(k1 - 1) * n + i0]; }

// Reverse Cuthill-McKee on the cell face graph (Cuthill & McKee 1969; George &
// Liu): BFS from a low-degree peripheral cell, neighbours visited in order of
// increasing degree, final order reversed.
std::vector<int64_t> rcm_order(int64_t n, const std::vector<int64_t>& xadj, const std::vector<int64_t>& adj) {
  std::vector<int64_t> deg(n), order;
  order.reserve(n);
  for (int64_t i = 0; i < n; ++i) deg[i] = xadj[i + 1] - xadj[i];
  std::vector<char> seen(n, 0);
  std::vector<int64_t> level(n, -1);
  auto bfs_last = [&](int64_t s) {  // farthest low-degree cell from s
    std::deque<int64_t> q{s};
    std::vector<int64_t> touched{s};
    level[s] = 0;
    int64_t best = s;
    while (!q.empty()) {
      int64_t c = q.front(); q.pop_front();
      if (level[c] > level[best] || (level[c] == level[best] && deg[c] < deg[best])) best = c;
      for (int64_t e = xadj[c]; e < xadj[c + 1]; ++e) {
        int64_t d = adj[e];
        if (level[d] < 0) { level[d] = level[c] + 1; q.push_back(d); touched.push_back(d); }
      }
    }
    for (int64_t t : touched) level[t] = -1;
    return best;
  };
  std::vector<int64_t> by_deg(n);
  std::iota(by_deg.begin(), by_deg.end(), 0);
  std::stable_sort(by_deg.begin(), by_deg.end(), [&](int64_t a, int64_t b) { return deg[a] < deg[b]; });
  std::vector<int64_t> nbrs;
  for (int64_t s0 : by_deg) {
    if (seen[s0]) continue;
    const int64_t s = bfs_last(bfs_last(s0));
    std::deque<int64_t> q{s};
    seen[s] = 1;
    while (!q.empty()) {
      int64_t c = q.front(); q.pop_front();
      order.push_back(c);
      nbrs.clear();
      for (int64_t e = xadj[c]; e < xadj[c + 1]; ++e)
        if (!seen[adj[e]]) { nbrs.push_back(adj[e]); seen[adj[e]] = 1; }
      std::stable_sort(nbrs.begin(), nbrs.end(), [&](int64_t a, int64_t b) { return deg[a] < deg[b]; });
      for (int64_t d : nbrs) q.push_back(d);
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// Hilbert-curve order of the cell centroids: the bounding square is divided
// into 2^21 x 2^21 squares and cells are sorted by the position of theirs
// along the curve (the iterative quadrant / rotate-and-reflect construction of
// the curve's index; ties by caller id).  Consecutive cells then form compact
// patches at every scale, so a block of cells and the faces it owns touch few
// cache lines outside the block.
std::vector<int64_t> hilbert_order(int64_t n, const double* xy /* n x 2, column-major */) {
  double lo[2] = {xy[0], xy[n]}, hi[2] = {xy[0], xy[n]};
  for (int q = 0; q < 2; ++q)
    for (int64_t i = 0; i < n; ++i) {
      lo[q] = std::min(lo[q], xy[q * n + i]);
      hi[q] = std::max(hi[q], xy[q * n + i]);
    }
  constexpr uint64_t side = uint64_t(1) << 21;
  const double span = std::max(hi[0] - lo[0], hi[1] - lo[1]);
  const double scale = span > 0 ? (double)(side - 1) / span : 0.0;
  std::vector<uint64_t> key(n);
  for (int64_t i = 0; i < n; ++i) {
    uint64_t x = (uint64_t)std::min((double)(side - 1), std::max(0.0, (xy[i] - lo[0]) * scale));
    uint64_t y = (uint64_t)std::min((double)(side - 1), std::max(0.0, (xy[n + i] - lo[1]) * scale));
    uint64_t d = 0;
    for (uint64_t s = side >> 1; s > 0; s >>= 1) {
      const uint64_t rx = (x & s) ? 1 : 0, ry = (y & s) ? 1 : 0;
      d += s * s * ((3 * rx) ^ ry);
      if (ry == 0) {  // enter the quadrant in the curve's orientation
        if (rx == 1) { x = side - 1 - x; y = side - 1 - y; }
        std::swap(x, y);
      }
    }
    key[i] = d;
  }
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return key[a] != key[b] ? key[a] < key[b] : a < b; });
  return order;
}

}  // namespace

struct gd_mesh : gd::Mesh {};
struct gd_state : gd::State {
  std::vector<hipEvent_t> ev_pool;
};
static int materialize(gd_state* S);

// The definitions below take C linkage from their declarations in godunov.h.

const char* gd_version(void) { return "libgodunov 0.1 (gfx950)"; }

int gd_last_error(char* buf, size_t n) {
  if (!buf || n == 0) return GD_ERR_INVALID;
  std::snprintf(buf, n, "%s", g_err.c_str());
  return GD_OK;
}

int gd_device_count(int32_t* n) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  if (n) *n = c;
  return GD_OK;
}

void gd_params_default(gd_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->scheme = GD_SCHEME_REF;
  p->precision = GD_FP64;
  p->flux = GD_FLUX_AUSM_PLUS;
  p->limiter = GD_LIM_REF_HARMONIC;
  p->limiter_eps = 1.0e-12;
  p->venkat_K = 5.0;
  for (int v = 0; v < 4; ++v) p->venkat_scale[v] = 1.0;
  p->viscous = 0;
  p->energy_law = GD_ENERGY_AIR_POLY;
  p->mu_const = 1.0e-2;
  p->k_const = 0.0;
  p->RGAS = 287.058;
  p->Gamma = 1.4;
  p->mu_ref = 1.8476e-5;
  p->T_ref = 300.0;
  p->S_suth = 116.0;
  p->Pr = 0.72;
  for (int i = 0; i <= GD_MAX_BC; ++i) p->bc[i].kind = GD_BC_EXTRAPOLATE;
}

int gd_mesh_create(const gd_mesh_desc* d, int32_t order, const int64_t* perm, gd_mesh** out) {
  if (!d || !out) return fail(GD_ERR_INVALID, "null argument");
  *out = nullptr;
  const int64_t nc = d->n_cells, nn = d->n_nodes;
  const int32_t ncol = d->n_neib_cells;
  if (nc <= 0 || nn <= 0 || ncol <= 0 || nc >= (int64_t)INT32_MAX || nn >= (int64_t)INT32_MAX)
    return fail(GD_ERR_INVALID, "bad mesh sizes");
  if (!d->mesh_connectivity || !d->cell_stiffness || !d->cell_edges_Nx || !d->cell_edges_Ny ||
      !d->cell_edges_length || !d->cell_areas || !d->cell_clusters || !d->node_stencils || !d->node2cellsL2up ||
      !d->node2cellsL2down || !d->cells2nodes || !d->cell_mid_points || !d->cell_nodes_X || !d->cell_nodes_Y ||
      !d->xNodes || !d->yNodes)
    return fail(GD_ERR_INVALID, "missing mesh array");
  if (order != GD_ORDER_AS_GIVEN && order != GD_ORDER_RCM && order != GD_ORDER_PERMUTATION && order != GD_ORDER_HILBERT)
    return fail(GD_ERR_INVALID, "unknown cell order");
  // --- validate and collect topology in the caller's numbering (0-based); all
  // of the host-side construction runs (and is checked) before any device call
  std::vector<int32_t> nv(nc);
  int32_t has_quads = 0;
  for (int64_t i = 0; i < nc; ++i) {
    nv[i] = CI(d->mesh_connectivity, nc, i, 3);
    if (nv[i] != 3 && nv[i] != 4) return fail(GD_ERR_INVALID, "cell with neither 3 nor 4 vertices");
    if (nv[i] == 4) has_quads = 1;
  }
  const int nf = has_quads ? 4 : 3;
  std::vector<int64_t> xadj(nc + 1, 0), adj;
  adj.reserve(nc * 3);
  int64_t nbf = 0;
  for (int64_t i = 0; i < nc; ++i) {
    for (int k = 1; k <= nv[i]; ++k) {
      const int32_t e = CI(d->cell_stiffness, nc, i, k);
      if (e >= 1 && e <= nc) adj.push_back(e - 1);
      else if (e > nc) return fail(GD_ERR_INVALID, "neighbour id out of range");
      else ++nbf;
    }
    xadj[i + 1] = (int64_t)adj.size();
  }
  for (int64_t i = 0; i < nc; ++i)
    for (int k = 1; k <= nv[i]; ++k) {
      const int64_t vv[4] = {CI(d->node2cellsL2down, nc, i, 2 * k - 1), CI(d->node2cellsL2up, nc, i, 2 * k - 1),
                             CI(d->cells2nodes, nc, i, 2 * k - 1), CI(d->cells2nodes, nc, i, 2 * k)};
      const bool interior = CI(d->cell_stiffness, nc, i, k) >= 1;
      if (vv[2] < 1 || vv[2] > nn || vv[3] < 1 || vv[3] > nn) return fail(GD_ERR_INVALID, "cells2nodes out of range");
      if (interior && (vv[0] < 1 || vv[0] > nn || vv[1] < 1 || vv[1] > nn))
        return fail(GD_ERR_INVALID, "node2cellsL2up/down out of range");
      if (interior && nv[i] == 4) {
        const int64_t a = CI(d->node2cellsL2down, nc, i, 2 * k), b = CI(d->node2cellsL2up, nc, i, 2 * k);
        if (a < 1 || a > nn || b < 1 || b > nn) return fail(GD_ERR_INVALID, "quad stencil node out of range");
      }
    }

  gd_mesh* M = new gd_mesh();
  (void)hipGetDevice(&M->device);
  M->nv_caller = nv;
  // --- cell order
  std::vector<int64_t>& cp = M->cperm;
  if (order == GD_ORDER_RCM) {
    cp = rcm_order(nc, xadj, adj);
  } else if (order == GD_ORDER_HILBERT) {
    cp = hilbert_order(nc, d->cell_mid_points);
  } else if (order == GD_ORDER_PERMUTATION) {
    if (!perm) { delete M; return fail(GD_ERR_INVALID, "GD_ORDER_PERMUTATION without a permutation"); }
    cp.assign(perm, perm + nc);
  } else {
    cp.resize(nc);
    std::iota(cp.begin(), cp.end(), 0);
  }
  M->ciperm.assign(nc, -1);
  for (int64_t c = 0; c < nc; ++c) {
    if (cp[c] < 0 || cp[c] >= nc || M->ciperm[cp[c]] != -1) { delete M; return fail(GD_ERR_INVALID, "not a permutation"); }
    M->ciperm[cp[c]] = c;
  }
  // node order: as given, or first touch in the new cell order
  std::vector<int64_t>& np = M->nperm;
  M->niperm.assign(nn, -1);
  if (order == GD_ORDER_AS_GIVEN) {
    np.resize(nn);
    std::iota(np.begin(), np.end(), 0);
    for (int64_t j = 0; j < nn; ++j) M->niperm[j] = j;
  } else {
    np.reserve(nn);
    for (int64_t c = 0; c < nc; ++c) {
      const int64_t o = cp[c];
      for (int v = 0; v < nv[o]; ++v) {
        const int64_t j = CI(d->mesh_connectivity, nc, o, 4 + v) - 1;
        if (j >= 0 && j < nn && M->niperm[j] < 0) { M->niperm[j] = (int64_t)np.size(); np.push_back(j); }
      }
    }
    for (int64_t j = 0; j < nn; ++j)
      if (M->niperm[j] < 0) { M->niperm[j] = (int64_t)np.size(); np.push_back(j); }
  }
  M->cells_identity = M->nodes_identity = true;
  for (int64_t c = 0; c < nc; ++c) M->cells_identity = M->cells_identity && cp[c] == c;
  for (int64_t j = 0; j < nn; ++j) M->nodes_identity = M->nodes_identity && np[j] == j;
  const auto& ci = M->ciperm;
  const auto& ni = M->niperm;
  auto new_node = [&](int64_t id1) -> int32_t { return (id1 >= 1 && id1 <= nn) ? (int32_t)ni[id1 - 1] : 0; };

  // --- per-cell arrays in device order
  std::vector<int32_t> h_nv(nc), h_nb(nf * nc), h_dn1(nf * nc, 0), h_dn2(nf * nc, 0), h_up1(nf * nc, 0),
      h_up2(nf * nc, 0), h_ea(nf * nc, 0), h_eb(nf * nc, 0), h_wid(nf * nc, -1);
  std::vector<double> h_nx(nf * nc, 0.0), h_ny(nf * nc, 0.0), h_len(nf * nc, 0.0), h_area(nc), h_rarea(nc), h_hx(nc),
      h_hmin(nc), h_xc(nc), h_yc(nc);
  std::vector<double> h_lsq(nf * 2 * nc, 0.0), h_rf(nf * 2 * nc, 0.0), h_dc(nf * 2 * nc, 0.0),
      h_rn(nf * 2 * nc, 0.0);
  std::vector<int32_t> h_kn(nf * nc, -1);
  double max_hx = 0.0;
  int64_t bandwidth = 0;
  int32_t n_wall = 0;
  // face midpoint relative to the centroid, from the cell's own vertex coordinates
  auto rface = [&](int64_t o, int k, double* r) {
    const int32_t a = CI(d->cells2nodes, nc, o, 2 * k - 1), b = CI(d->cells2nodes, nc, o, 2 * k);
    double xa = 0, ya = 0, xb = 0, yb = 0;
    for (int v = 0; v < nv[o]; ++v) {
      const int32_t id = CI(d->mesh_connectivity, nc, o, 4 + v);
      if (id == a) { xa = CD(d->cell_nodes_X, nc, o, v + 1); ya = CD(d->cell_nodes_Y, nc, o, v + 1); }
      if (id == b) { xb = CD(d->cell_nodes_X, nc, o, v + 1); yb = CD(d->cell_nodes_Y, nc, o, v + 1); }
    }
    r[0] = 0.5 * (xa + xb) - CD(d->cell_mid_points, nc, o, 1);
    r[1] = 0.5 * (ya + yb) - CD(d->cell_mid_points, nc, o, 2);
  };
  for (int64_t c = 0; c < nc; ++c) {
    const int64_t o = cp[c];
    h_nv[c] = nv[o];
    const double A = d->cell_areas[o];
    if (!(A > 0.0)) { delete M; return fail(GD_ERR_INVALID, "non-positive cell area"); }
    h_area[c] = A;
    h_rarea[c] = 1.0 / A;
    h_hx[c] = d->HX ? d->HX[o] : 0.0;
    max_hx = std::max(max_hx, h_hx[c]);
    h_xc[c] = CD(d->cell_mid_points, nc, o, 1);
    h_yc[c] = CD(d->cell_mid_points, nc, o, 2);
    double perim = 0.0;
    double M11 = 0, M12 = 0, M22 = 0;
    double dcx[4] = {0, 0, 0, 0}, dcy[4] = {0, 0, 0, 0}, wk[4] = {0, 0, 0, 0};
    for (int k = 1; k <= nf; ++k) {
      const int64_t s = (int64_t)(k - 1) * nc + c;
      if (k > nv[o]) { h_nb[s] = -1; continue; }
      const int32_t e = CI(d->cell_stiffness, nc, o, k);
      h_nx[s] = CD(d->cell_edges_Nx, nc, o, k);
      h_ny[s] = CD(d->cell_edges_Ny, nc, o, k);
      h_len[s] = CD(d->cell_edges_length, nc, o, k);
      perim += h_len[s];
      h_ea[s] = new_node(CI(d->cells2nodes, nc, o, 2 * k - 1));
      h_eb[s] = new_node(CI(d->cells2nodes, nc, o, 2 * k));
      double r[2];
      rface(o, k, r);
      h_rf[((int64_t)(k - 1) * 2 + 0) * nc + c] = r[0];
      h_rf[((int64_t)(k - 1) * 2 + 1) * nc + c] = r[1];
      if (e >= 1) {
        h_nb[s] = (int32_t)ci[e - 1];
        bandwidth = std::max<int64_t>(bandwidth, (int64_t)std::llabs((long long)(ci[e - 1] - c)));
        h_dn1[s] = new_node(CI(d->node2cellsL2down, nc, o, 2 * k - 1));
        h_up1[s] = new_node(CI(d->node2cellsL2up, nc, o, 2 * k - 1));
        if (nv[o] == 4) {
          h_dn2[s] = new_node(CI(d->node2cellsL2down, nc, o, 2 * k));
          h_up2[s] = new_node(CI(d->node2cellsL2up, nc, o, 2 * k));
        }
        // neighbour centroid offset through the shared face (periodic safe)
        const int64_t oj = e - 1;
        int kj = 0;
        for (int q = 1; q <= nv[oj]; ++q)
          if (CI(d->cell_stiffness, nc, oj, q) == o + 1) kj = q;
        if (kj == 0) { delete M; return fail(GD_ERR_INVALID, "neighbour relation is not symmetric"); }
        double rj[2];
        rface(oj, kj, rj);
        dcx[k - 1] = r[0] - rj[0];
        dcy[k - 1] = r[1] - rj[1];
        h_kn[s] = kj - 1;
        h_rn[((int64_t)(k - 1) * 2 + 0) * nc + c] = rj[0];
        h_rn[((int64_t)(k - 1) * 2 + 1) * nc + c] = rj[1];
      } else {
        h_nb[s] = e - 1;  // code - 1 < 0
        h_wid[s] = n_wall++;
        M->wall_codes.push_back(e);
        // ghost centroid: the own centroid mirrored across the face
        const double rn = r[0] * h_nx[s] + r[1] * h_ny[s];
        dcx[k - 1] = 2.0 * rn * h_nx[s];
        dcy[k - 1] = 2.0 * rn * h_ny[s];
      }
      const double w = 1.0 / (dcx[k - 1] * dcx[k - 1] + dcy[k - 1] * dcy[k - 1]);
      wk[k - 1] = w;
      M11 += w * dcx[k - 1] * dcx[k - 1];
      M12 += w * dcx[k - 1] * dcy[k - 1];
      M22 += w * dcy[k - 1] * dcy[k - 1];
      h_dc[((int64_t)(k - 1) * 2 + 0) * nc + c] = dcx[k - 1];
      h_dc[((int64_t)(k - 1) * 2 + 1) * nc + c] = dcy[k - 1];
    }
    h_hmin[c] = 2.0 * A / perim;
    // inverse-distance-squared weighted least squares: grad = M^-1 sum_k w_k d_k dq_k
    const double det = M11 * M22 - M12 * M12;
    for (int k = 1; k <= nv[o]; ++k) {
      const double ix = (M22 * dcx[k - 1] - M12 * dcy[k - 1]) / det;
      const double iy = (-M12 * dcx[k - 1] + M11 * dcy[k - 1]) / det;
      h_lsq[((int64_t)(k - 1) * 2 + 0) * nc + c] = wk[k - 1] * ix;
      h_lsq[((int64_t)(k - 1) * 2 + 1) * nc + c] = wk[k - 1] * iy;
    }
  }
  // --- clusters: keep slot order, drop empty slots, width = widest node
  int32_t ncl = 0;
  for (int64_t j = 0; j < nn; ++j) {
    int32_t cnt = 0;
    for (int q = 1; q <= ncol; ++q) {
      const int32_t c = CI(d->cell_clusters, nn, j, q);
      if (c != 0) {
        if (c < 0 || c > nc) { delete M; return fail(GD_ERR_INVALID, "cell_clusters id out of range"); }
        ++cnt;
      }
    }
    ncl = std::max(ncl, cnt);
  }
  if (ncl == 0) ncl = 1;
  std::vector<int32_t> h_cl((int64_t)ncl * nn, -1);
  std::vector<double> h_w((int64_t)ncl * nn, 0.0), h_xn(nn), h_yn(nn);
  for (int64_t jn = 0; jn < nn; ++jn) {
    const int64_t j = np[jn];
    int32_t s = 0;
    for (int q = 1; q <= ncol; ++q) {
      const int32_t c = CI(d->cell_clusters, nn, j, q);
      if (c != 0) {
        h_cl[(int64_t)s * nn + jn] = (int32_t)ci[c - 1];
        h_w[(int64_t)s * nn + jn] = CD(d->node_stencils, nn, j, q);
        ++s;
      }
    }
    h_xn[jn] = d->xNodes[j];
    h_yn[jn] = d->yNodes[j];
  }

  // --- face list of the ns flux pass: each interior face once, owned by the
  // cell with the lower caller id; faces numbered in owner order
  std::vector<int32_t> h_fc0, h_fc1, h_fwall, h_cface(nf * nc, -1);
  std::vector<double> h_fnx, h_fny, h_flen, h_fr0[2], h_fr1[2];
  for (int64_t c = 0; c < nc; ++c) {
    for (int k = 0; k < h_nv[c]; ++k) {
      const int64_t s = (int64_t)k * nc + c;
      const int32_t j = h_nb[s];
      if (j >= 0 && cp[j] < cp[c]) continue;  // the neighbour owns it
      const int32_t f = (int32_t)h_fc0.size();
      h_fc0.push_back((int32_t)c);
      h_fc1.push_back(j);
      h_fwall.push_back(j >= 0 ? -1 : h_wid[s]);
      h_cface[s] = 2 * f;
      if (j >= 0) h_cface[(int64_t)h_kn[s] * nc + j] = 2 * f + 1;
      h_fnx.push_back(h_nx[s]);
      h_fny.push_back(h_ny[s]);
      h_flen.push_back(h_len[s]);
      for (int q = 0; q < 2; ++q) {
        h_fr0[q].push_back(h_rf[((int64_t)k * 2 + q) * nc + c]);
        h_fr1[q].push_back(h_rn[((int64_t)k * 2 + q) * nc + c]);
      }
    }
  }
  const int64_t nfaces = (int64_t)h_fc0.size();
  if (2 * nfaces + 1 > (int64_t)INT32_MAX) {  // cface holds 2 f + side in int32
    delete M;
    return fail(GD_ERR_INVALID, "too many faces");
  }
  for (int64_t c = 0; c < nc; ++c)
    for (int k = 0; k < h_nv[c]; ++k)
      if (h_cface[(int64_t)k * nc + c] < 0) {
        delete M;
        return fail(GD_ERR_INVALID, "a cell lists the same neighbour on two faces");
      }
  // --- face list of the reference scheme's face pass (see MeshView::rcf)
  std::vector<int32_t> h_rfc, h_rfs, h_rcf(nf * nc, -1), h_rfnb, h_rfdn1, h_rfup1;
  std::vector<double> h_rfnx, h_rfny, h_rflen;
  for (int64_t c = 0; c < nc; ++c) {
    if (h_nv[c] != 3) continue;
    for (int k = 0; k < 3; ++k) {
      const int64_t s = (int64_t)k * nc + c;
      const int32_t j = h_nb[s];
      if (j < 0 || h_nv[j] != 3 || cp[j] < cp[c]) continue;  // boundary, quad, or the neighbour owns it
      const int64_t sj = (int64_t)h_kn[s] * nc + j;
      const bool mirror = h_nx[sj] == -h_nx[s] && h_ny[sj] == -h_ny[s] && h_len[sj] == h_len[s] &&
                          h_dn1[sj] == h_up1[s] && h_up1[sj] == h_dn1[s];
      if (!mirror) continue;
      const int32_t f = (int32_t)h_rfc.size();
      h_rfc.push_back((int32_t)c);
      h_rfs.push_back(k);
      h_rcf[s] = 2 * f;
      h_rcf[sj] = 2 * f + 1;
      h_rfnb.push_back(j);
      h_rfdn1.push_back(h_dn1[s]);
      h_rfup1.push_back(h_up1[s]);
      h_rfnx.push_back(h_nx[s]);
      h_rfny.push_back(h_ny[s]);
      h_rflen.push_back(h_len[s]);
    }
  }
  const int64_t n_rmirror = (int64_t)h_rfc.size();
  if (2 * (int64_t)nf * nc + 1 > (int64_t)INT32_MAX) {  // rcf holds 2 f + side in int32
    delete M;
    return fail(GD_ERR_INVALID, "too many face slots");
  }
  for (int64_t c = 0; c < nc; ++c)  // one-sided entries: every other face slot, in cell order
    for (int k = 0; k < h_nv[c]; ++k) {
      const int64_t s = (int64_t)k * nc + c;
      if (h_rcf[s] >= 0) continue;
      h_rcf[s] = 2 * (int32_t)h_rfc.size();
      h_rfc.push_back((int32_t)c);
      h_rfs.push_back(k);
    }
  auto pair =[&](const std::vector<double>* a) {
    std::vector<double> o(2 * nfaces);
    std::copy(a[0].begin(), a[0].end(), o.begin());
    std::copy(a[1].begin(), a[1].end(), o.begin() + nfaces);
    return o;
  };
  auto to_float = [](const std::vector<double>& a) {
    std::vector<float> o(a.size());
    for (size_t q = 0; q < a.size(); ++q) o[q] = (float)a[q];
    return o;
  };
  const std::vector<double> h_fr0p = pair(h_fr0), h_fr1p = pair(h_fr1);

  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) {
    delete M;
    return fail(GD_ERR_NO_DEVICE, "no HIP device");
  }
  {
    auto to32 = [](const std::vector<int64_t>& a) { return std::vector<int32_t>(a.begin(), a.end()); };
    int rcp = 0;
    rcp |= upload(M->allocs, &M->d_cperm, to32(M->cperm), &M->device_bytes);
    rcp |= upload(M->allocs, &M->d_ciperm, to32(M->ciperm), &M->device_bytes);
    rcp |= upload(M->allocs, &M->d_nperm, to32(M->nperm), &M->device_bytes);
    rcp |= upload(M->allocs, &M->d_niperm, to32(M->niperm), &M->device_bytes);
    if (rcp) { std::string msg = g_err; gd_mesh_destroy(M); return fail(rcp, msg); }
  }
  MeshView& v = M->v;
  v.nc = nc; v.nn = nn; v.nf = nf; v.ncl = ncl; v.has_quads = has_quads; v.n_wall = n_wall; v.max_hx = max_hx;
  int64_t* B = &M->device_bytes;
  int rc = 0;
  auto& al = M->allocs;
  rc |= upload(al, &v.nv, h_nv, B);
  rc |= upload(al, &v.nb, h_nb, B);
  rc |= upload(al, &v.fnx, h_nx, B);
  rc |= upload(al, &v.fny, h_ny, B);
  rc |= upload(al, &v.flen, h_len, B);
  rc |= upload(al, &v.dn1, h_dn1, B);
  rc |= upload(al, &v.dn2, h_dn2, B);
  rc |= upload(al, &v.up1, h_up1, B);
  rc |= upload(al, &v.up2, h_up2, B);
  rc |= upload(al, &v.ea, h_ea, B);
  rc |= upload(al, &v.eb, h_eb, B);
  rc |= upload(al, &v.area, h_area, B);
  rc |= upload(al, &v.rarea, h_rarea, B);
  rc |= upload(al, &v.hx, h_hx, B);
  rc |= upload(al, &v.hmin, h_hmin, B);
  rc |= upload(al, &v.xc, h_xc, B);
  rc |= upload(al, &v.yc, h_yc, B);
  rc |= upload(al, &v.lsq, h_lsq, B);
  rc |= upload(al, &v.rf, h_rf, B);
  rc |= upload(al, &v.wall_id, h_wid, B);
  rc |= upload(al, &v.kn, h_kn, B);
  {
    std::vector<float> f_nx(h_nx.size()), f_ny(h_ny.size()), f_lsq(h_lsq.size()), f_rf(h_rf.size());
    for (size_t q = 0; q < h_nx.size(); ++q) { f_nx[q] = (float)h_nx[q]; f_ny[q] = (float)h_ny[q]; }
    for (size_t q = 0; q < h_lsq.size(); ++q) { f_lsq[q] = (float)h_lsq[q]; f_rf[q] = (float)h_rf[q]; }
    rc |= upload(al, &v.fnx32, f_nx, B);
    rc |= upload(al, &v.fny32, f_ny, B);
    rc |= upload(al, &v.lsq32, f_lsq, B);
    rc |= upload(al, &v.rf32, f_rf, B);
  }
  v.n_faces = nfaces;
  rc |= upload(al, &v.fc0, h_fc0, B);
  rc |= upload(al, &v.fc1, h_fc1, B);
  rc |= upload(al, &v.fwall, h_fwall, B);
  rc |= upload(al, &v.f_nx, h_fnx, B);
  rc |= upload(al, &v.f_ny, h_fny, B);
  rc |= upload(al, &v.f_len, h_flen, B);
  rc |= upload(al, &v.f_r0, h_fr0p, B);
  rc |= upload(al, &v.f_r1, h_fr1p, B);
  rc |= upload(al, &v.f_nx32, to_float(h_fnx), B);
  rc |= upload(al, &v.f_ny32, to_float(h_fny), B);
  rc |= upload(al, &v.f_r032, to_float(h_fr0p), B);
  rc |= upload(al, &v.f_r132, to_float(h_fr1p), B);
  rc |= upload(al, &v.f_len32, to_float(h_flen), B);
  rc |= upload(al, &v.hmin32, to_float(h_hmin), B);
  rc |= upload(al, &v.cface, h_cface, B);
  v.n_rfaces = (int64_t)h_rfc.size();
  v.n_rmirror = n_rmirror;
  rc |= upload(al, &v.rf_cell, h_rfc, B);
  rc |= upload(al, &v.rf_slot, h_rfs, B);
  rc |= upload(al, &v.rcf, h_rcf, B);
  rc |= upload(al, &v.rf_nb, h_rfnb, B);
  rc |= upload(al, &v.rf_dn1, h_rfdn1, B);
  rc |= upload(al, &v.rf_up1, h_rfup1, B);
  rc |= upload(al, &v.rf_nx, h_rfnx, B);
  rc |= upload(al, &v.rf_ny, h_rfny, B);
  rc |= upload(al, &v.rf_len, h_rflen, B);
  rc |= upload(al, &v.cl, h_cl, B);
  rc |= upload(al, &v.w, h_w, B);
  rc |= upload(al, &v.xn, h_xn, B);
  rc |= upload(al, &v.yn, h_yn, B);
  if (rc) {
    std::string msg = g_err;
    gd_mesh_destroy(M);
    return fail(rc, msg);
  }
  M->n_boundary_faces = nbf;
  M->bandwidth = bandwidth;
  M->max_cluster = ncl;
  *out = M;
  return GD_OK;
}

int gd_mesh_info_get(const gd_mesh* m, gd_mesh_info* info) {
  if (!m || !info) return fail(GD_ERR_INVALID, "null argument");
  info->n_cells = m->v.nc;
  info->n_nodes = m->v.nn;
  info->n_boundary_faces = m->n_boundary_faces;
  info->has_quads = m->v.has_quads;
  info->max_cluster = m->max_cluster;
  info->device_bytes = m->device_bytes;
  info->bandwidth = m->bandwidth;
  return GD_OK;
}

int gd_mesh_cell_order(const gd_mesh* m, int64_t* new_to_old) {
  if (!m || !new_to_old) return fail(GD_ERR_INVALID, "null argument");
  std::copy(m->cperm.begin(), m->cperm.end(), new_to_old);
  return GD_OK;
}

int gd_mesh_destroy(gd_mesh* m) {
  if (!m) return GD_OK;
  for (void* p : m->allocs) (void)hipFree(p);
  delete m;
  return GD_OK;
}

namespace gd {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
int check_params(const Mesh* m, const gd_params* p);
}  // namespace gd

int gd::check_params(const gd::Mesh* m, const gd_params* p) {
  if (p->scheme != GD_SCHEME_REF && p->scheme != GD_SCHEME_NS) return fail(GD_ERR_INVALID, "unknown scheme");
  if (p->precision != GD_FP64 && p->precision != GD_FP32_MIXED) return fail(GD_ERR_INVALID, "unknown precision");
  if (p->scheme == GD_SCHEME_REF && p->precision != GD_FP64)
    return fail(GD_ERR_UNSUPPORTED, "the reference scheme runs in FP64 only");
  if (p->flux < GD_FLUX_AUSM_PLUS || p->flux > GD_FLUX_AUSM_PLUS_M2020) return fail(GD_ERR_INVALID, "unknown flux");
  if (p->scheme == GD_SCHEME_NS && p->flux == GD_FLUX_AUSM_PLUS_M2020)
    return fail(GD_ERR_UNSUPPORTED, "AUSM+M2020 is available in the reference scheme only");
  if (p->scheme == GD_SCHEME_REF && (p->limiter < GD_LIM_REF_HARMONIC || p->limiter > GD_LIM_MINMOD))
    return fail(GD_ERR_INVALID, "reference scheme limiter must be one of the reference limiters");
  if (p->scheme == GD_SCHEME_NS && p->limiter != GD_LIM_VENKATAKRISHNAN)
    return fail(GD_ERR_UNSUPPORTED, "the ns scheme uses the Venkatakrishnan limiter");
  if (!(p->Gamma > 1.0) || !(p->RGAS > 0.0)) return fail(GD_ERR_INVALID, "bad gas constants");
  for (int i = 0; i <= GD_MAX_BC; ++i)
    if (p->bc[i].kind < GD_BC_NONE || p->bc[i].kind > GD_BC_ISOTHERMAL_WALL) return fail(GD_ERR_INVALID, "bad BC kind");
  (void)m;
  return GD_OK;
}

int gd_state_create(gd_mesh* m, const gd_params* p, gd_state** out) {
  if (!m || !p || !out) return fail(GD_ERR_INVALID, "null argument");
  *out = nullptr;
  int rc = check_params(m, p);
  if (rc) return rc;
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != m->device) return fail(GD_ERR_INVALID, "the current device is not the mesh's device");
  gd_state* S = new gd_state();
  S->mesh = m;
  S->prm = *p;
  const int64_t nc = m->v.nc, nn = m->v.nn;
  hipError_t e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete S; return fail(GD_ERR_HIP, "hipStreamCreate failed"); }
  auto& al = S->allocs;
  rc |= dev_alloc(al, &S->U[0], 4 * nc);
  rc |= dev_alloc(al, &S->U[1], 4 * nc);
  rc |= dev_alloc(al, &S->Q[0], 4 * nc);
  rc |= dev_alloc(al, &S->Q[1], 4 * nc);
  rc |= dev_alloc(al, &S->aS, nc);
  rc |= dev_alloc(al, &S->vmax, nc);
  rc |= dev_alloc(al, &S->Qn, 4 * nn);
  rc |= dev_alloc(al, &S->Un, 4 * nn);
  rc |= dev_alloc(al, &S->Ud, 4 * nc);
  rc |= dev_alloc(al, &S->F, 4 * nc);
  rc |= dev_alloc(al, &S->Dl, 4 * nc);
  rc |= dev_alloc(al, &S->phiN, 4 * nn);
  rc |= dev_alloc(al, &S->g6, 6 * nc);
  rc |= dev_alloc(al, &S->g6N, 6 * nn);
  double* rbuf = nullptr;
  rc |= dev_alloc(al, &rbuf, 16 * nc);
  S->rec = rbuf;
  double* fbuf = nullptr;
  // the ns face fluxes, or the reference loop's face-pass fluxes
  rc |= dev_alloc(al, &fbuf, 4 * (size_t)std::max<int64_t>(std::max(m->v.n_faces, m->v.n_rfaces), 1));
  S->fflux = fbuf;
  double* qt0 = nullptr;
  double* qt1 = nullptr;
  rc |= dev_alloc(al, &qt0, 4 * nc);
  rc |= dev_alloc(al, &qt1, 4 * nc);
  S->Qt[0] = qt0;
  S->Qt[1] = qt1;
  rc |= dev_alloc(al, &S->wall_f, 2 * (size_t)std::max<int32_t>(m->v.n_wall, 1));
  rc |= dev_alloc(al, &S->hdr, 1);
  rc |= dev_alloc(al, &S->part, 3 * ((nc + kThreads - 1) / kThreads) + 3);
  rc |= dev_alloc(al, &S->red, 4 * ((nc + kThreads - 1) / kThreads) + 4);
  if (!rc) {
    S->acc = S->hdr->acc;
    S->drag = S->hdr->drag;
    if (hipHostMalloc((void**)&S->hdr_host, sizeof(RunHdr), hipHostMallocDefault) != hipSuccess) {
      S->hdr_host = nullptr;
      rc = fail(GD_ERR_NOMEM, "hipHostMalloc failed");
    }
  }
  rc |= dev_alloc(al, &S->xfer, 4 * (size_t)std::max(nc, nn));
  if (rc) {
    std::string msg = g_err;
    gd_state_destroy(S);
    return fail(rc, msg);
  }
  // zero everything once so that reads of never-written arrays are defined
  (void)hipMemsetAsync(S->U[0], 0, 4 * nc * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->U[1], 0, 4 * nc * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->Q[0], 0, 4 * nc * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->Q[1], 0, 4 * nc * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->aS, 0, nc * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->vmax, 0, nc * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->Qn, 0, 4 * nn * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->Un, 0, 4 * nn * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->Ud, 0, 4 * nc * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->F, 0, 4 * nc * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->Dl, 0, 4 * nc * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->wall_f, 0, 2 * std::max<int32_t>(m->v.n_wall, 1) * sizeof(double), S->stream);
  (void)hipMemsetAsync(S->hdr, 0, sizeof(RunHdr), S->stream);
  e = hipStreamSynchronize(S->stream);
  if (e != hipSuccess) { gd_state_destroy(S); return fail(GD_ERR_HIP, "state init failed"); }
  *out = S;
  return GD_OK;
}

int gd_state_destroy(gd_state* S) {
  if (!S) return GD_OK;
  if (S->stream) (void)hipStreamSynchronize(S->stream);
  for (hipEvent_t ev : S->ev_pool) (void)hipEventDestroy(ev);
  for (void* p : S->allocs) (void)hipFree(p);
  if (S->adj_ws) (void)hipFree(S->adj_ws);
  if (S->hdr_host) (void)hipHostFree(S->hdr_host);
  if (S->stream) (void)hipStreamDestroy(S->stream);
  delete S;
  return GD_OK;
}

int gd_state_set_params(gd_state* S, const gd_params* p) {
  if (!S || !p) return fail(GD_ERR_INVALID, "null argument");
  int rcm = materialize(S);
  if (rcm) return rcm;
  int rc = check_params(S->mesh, p);
  if (rc) return rc;
  S->prm = *p;
  return GD_OK;
}

// host (caller order) <-> device (device order) column transfers: one copy
// of the caller's columns, then a gather on the device (or the reverse)
static int put_cols(gd_state* S, double* dev, const double* host, int ncols, bool nodes) {
  const Mesh& M = *S->mesh;
  const int64_t n = nodes ? M.v.nn : M.v.nc;
  const size_t bytes = (size_t)n * ncols * sizeof(double);
  if (nodes ? M.nodes_identity : M.cells_identity) {
    GD_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, S->stream));
  } else {
    GD_HIP(hipMemcpyAsync(S->xfer, host, bytes, hipMemcpyHostToDevice, S->stream));
    GD_HIP(launch_permute(dev, S->xfer, nodes ? M.d_nperm : M.d_cperm, n, ncols, S->stream));
  }
  GD_HIP(hipStreamSynchronize(S->stream));
  return GD_OK;
}

static int get_cols(gd_state* S, const double* dev, double* host, int ncols, bool nodes) {
  const Mesh& M = *S->mesh;
  const int64_t n = nodes ? M.v.nn : M.v.nc;
  const size_t bytes = (size_t)n * ncols * sizeof(double);
  if (nodes ? M.nodes_identity : M.cells_identity) {
    GD_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, S->stream));
  } else {
    GD_HIP(launch_permute(S->xfer, dev, nodes ? M.d_niperm : M.d_ciperm, n, ncols, S->stream));
    GD_HIP(hipMemcpyAsync(host, S->xfer, bytes, hipMemcpyDeviceToHost, S->stream));
  }
  GD_HIP(hipStreamSynchronize(S->stream));
  return GD_OK;
}

// The bookkeeping a run leaves pending: the reference loop's DeltaX of the
// last step and UconsCellsNewX = UconsCellsOldX (evaluate2d.jl:100-104), and
// for the ns scheme the FP64 fields2d primitives, aSoundCells and VMAXCells
// of U (evaluate2d.jl:87-93).  Every entry point that reads or writes those
// arrays calls this first; the time loop itself never needs them.
static int materialize(gd_state* S) {
  S->ns_primed = false;  // every entry point but gd_run comes here first
  if (!S->pend_delta && !S->pend_derived) return GD_OK;
  const MeshView& v = S->mesh->v;
  if (S->pend_delta) {
    GD_HIP(launch_delta(v.nc, S->U[S->cur], S->U[1 - S->cur], S->Dl, S->stream));
    GD_HIP(launch_copy(S->U[1 - S->cur], S->U[S->cur], 4 * v.nc, S->stream));
    S->pend_delta = false;
  }
  if (S->pend_derived) {
    GD_HIP(launch_update_variables(v, S->prm.Gamma, S->U[S->cur], S->U[S->cur], S->Q[S->qc], S->aS, S->vmax, nullptr,
                                   S->stream));
    S->pend_derived = false;
  }
  return GD_OK;
}

#define GD_MATERIALIZE(S)                   \
  do {                                      \
    const int rc_m_ = materialize(S);       \
    if (rc_m_) return rc_m_;                \
  } while (0)

int gd_state_set_fields(gd_state* S, const double* rC, const double* uC, const double* vC, const double* pC,
                        const double* rN, const double* uN, const double* vN, const double* pN) {
  if (!S) return fail(GD_ERR_INVALID, "null state");
  GD_MATERIALIZE(S);
  const int64_t nc = S->mesh->v.nc, nn = S->mesh->v.nn;
  const double* cs[4] = {rC, uC, vC, pC};
  const double* ns[4] = {rN, uN, vN, pN};
  int rc = 0;
  for (int q = 0; q < 4; ++q)
    if (cs[q] && (rc = put_cols(S, S->Q[S->qc] + q * nc, cs[q], 1, false))) return rc;
  for (int q = 0; q < 4; ++q)
    if (ns[q] && (rc = put_cols(S, S->Qn + q * nn, ns[q], 1, true))) return rc;
  // aSoundCells / VMAXCells as createFields2d computes them (initfields_TGV2d.jl:47-48)
  GD_HIP(launch_fields_derived(nc, S->prm.Gamma, S->Q[S->qc], S->aS, S->vmax, S->stream));
  GD_HIP(hipStreamSynchronize(S->stream));
  return GD_OK;
}

int gd_state_get_fields(gd_state* S, double* rC, double* uC, double* vC, double* pC, double* aC, double* vmC,
                        double* rN, double* uN, double* vN, double* pN) {
  if (!S) return fail(GD_ERR_INVALID, "null state");
  GD_MATERIALIZE(S);
  const int64_t nc = S->mesh->v.nc, nn = S->mesh->v.nn;
  double* cs[4] = {rC, uC, vC, pC};
  double* ns[4] = {rN, uN, vN, pN};
  int rc = 0;
  for (int q = 0; q < 4; ++q)
    if (cs[q] && (rc = get_cols(S, S->Q[S->qc] + q * nc, cs[q], 1, false))) return rc;
  if (aC || vmC) {
    GD_HIP(launch_fields_derived(nc, S->prm.Gamma, S->Q[S->qc], S->aS, S->vmax, S->stream));
    if (aC && (rc = get_cols(S, S->aS, aC, 1, false))) return rc;
    if (vmC && (rc = get_cols(S, S->vmax, vmC, 1, false))) return rc;
  }
  for (int q = 0; q < 4; ++q)
    if (ns[q] && (rc = get_cols(S, S->Qn + q * nn, ns[q], 1, true))) return rc;
  return GD_OK;
}

static double* array_ptr(gd_state* S, int32_t which, bool* nodes) {
  *nodes = false;
  switch (which) {
    case GD_ARR_U_OLD: return S->U[S->cur];
    case GD_ARR_U_NEW: return S->U[1 - S->cur];
    case GD_ARR_U_DIFF: return S->Ud;
    case GD_ARR_FLUXES: return S->F;
    case GD_ARR_DELTA: return S->Dl;
    case GD_ARR_U_NODES: *nodes = true; return S->Un;
    default: return nullptr;
  }
}

int gd_state_set_array(gd_state* S, int32_t which, const double* host) {
  if (!S || !host) return fail(GD_ERR_INVALID, "null argument");
  GD_MATERIALIZE(S);
  bool nodes;
  double* p = array_ptr(S, which, &nodes);
  if (!p) return fail(GD_ERR_INVALID, "unknown array");
  return put_cols(S, p, host, 4, nodes);
}

int gd_state_get_array(gd_state* S, int32_t which, double* host) {
  if (!S || !host) return fail(GD_ERR_INVALID, "null argument");
  GD_MATERIALIZE(S);
  bool nodes;
  double* p = array_ptr(S, which, &nodes);
  if (!p) return fail(GD_ERR_INVALID, "unknown array");
  return get_cols(S, p, host, 4, nodes);
}

int gd_state_device_u(gd_state* S, double** dev) {
  if (!S || !dev) return fail(GD_ERR_INVALID, "null argument");
  S->u_exposed = true;  // the caller may write U: every later run starts from it afresh
  S->ns_primed = false;
  *dev = S->U[S->cur];
  return GD_OK;
}

int gd_state_device_adjoint(gd_state* S, double** dev) {
  if (!S || !dev) return fail(GD_ERR_INVALID, "null argument");
  if (!S->grad0) return fail(GD_ERR_INVALID, "no adjoint has run on this state");
  *dev = S->grad0;
  return GD_OK;
}

int gd_sync(gd_state* S) {
  if (!S) return fail(GD_ERR_INVALID, "null state");
  GD_HIP(hipStreamSynchronize(S->stream));
  return GD_OK;
}

static int need_ref(gd_state* S) {
  if (S->prm.scheme != GD_SCHEME_REF) return fail(GD_ERR_UNSUPPORTED, "operator belongs to the reference scheme");
  return GD_OK;
}

int gd_calc_one_stage(gd_state* S, int32_t variant, double betta, double dt, double flow_time) {
  if (!S) return fail(GD_ERR_INVALID, "null state");
  GD_MATERIALIZE(S);
  int rc = need_ref(S);
  if (rc) return rc;
  if (variant != GD_STAGE_MUSCL_FAST && variant != GD_STAGE_MUSCL) return fail(GD_ERR_INVALID, "unknown variant");
  RefStepArgs a{};
  a.m = S->mesh->v;
  a.prm = S->prm;
  a.Q = S->Q[S->qc];
  a.Qn = S->Qn;
  a.Uo = S->U[S->cur];
  a.Ud = variant == GD_STAGE_MUSCL_FAST ? S->Ud : nullptr;
  a.include_udiff = variant == GD_STAGE_MUSCL_FAST;
  a.Un = S->U[1 - S->cur];
  a.F = S->F;
  a.betta = betta;
  a.dt = dt;
  a.flow_time = flow_time;
  GD_HIP(launch_ref_step(a, S->stream));
  GD_HIP(hipStreamSynchronize(S->stream));
  return GD_OK;
}

int gd_calc_face_states(gd_state* S, double flow_time, double* ULeft, double* URight) {
  if (!S || !ULeft || !URight) return fail(GD_ERR_INVALID, "null argument");
  GD_MATERIALIZE(S);
  int rc = need_ref(S);
  if (rc) return rc;
  const Mesh& M = *S->mesh;
  const int64_t nc = M.v.nc;
  if (!S->fstate && (rc = dev_alloc(S->allocs, &S->fstate, 2 * 16 * (size_t)nc))) return rc;
  RefStepArgs a{};
  a.m = M.v;
  a.prm = S->prm;
  a.Q = S->Q[S->qc];
  a.Qn = S->Qn;
  a.flow_time = flow_time;
  double* dUL = S->fstate;
  double* dUR = S->fstate + 16 * nc;
  GD_HIP(launch_ref_face_states(a, dUL, dUR, S->stream));
  // [slot][4][nc] device order -> per slot a (4, n_cells) column-major array
  // in the caller's numbering; slots past a cell's faces keep their contents
  std::vector<double> h(16 * (size_t)nc);
  for (int side = 0; side < 2; ++side) {
    if ((rc = get_cols(S, side ? dUR : dUL, h.data(), 16, false))) return rc;
    double* out = side ? URight : ULeft;
    for (int k = 0; k < 4; ++k)
      for (int64_t i = 0; i < nc; ++i) {
        if (k >= M.nv_caller[i]) continue;
        for (int v = 0; v < 4; ++v) out[(int64_t)k * 4 * nc + 4 * i + v] = h[((int64_t)k * 4 + v) * nc + i];
      }
  }
  return GD_OK;
}

int gd_calc_one_stage_cuda(gd_state* S, int32_t precision, double betta, double dt, double flow_time,
                           const double* ULeft, const double* URight, double* iFlux) {
  if (!S) return fail(GD_ERR_INVALID, "null state");
  GD_MATERIALIZE(S);
  int rc = need_ref(S);
  if (rc) return rc;
  if (precision != GD_FP64 && precision != GD_FP32_MIXED) return fail(GD_ERR_INVALID, "unknown precision");
  if ((ULeft == nullptr) != (URight == nullptr)) return fail(GD_ERR_INVALID, "give both face-state arrays or neither");
  const Mesh& M = *S->mesh;
  const int64_t nc = M.v.nc;
  if (!S->fstate && (rc = dev_alloc(S->allocs, &S->fstate, 2 * 16 * (size_t)nc))) return rc;
  double* dUL = S->fstate;
  double* dUR = S->fstate + 16 * nc;
  if (ULeft) {
    // the caller's slots 1..3, each (4, n_cells) column-major, to [slot][4][nc] in device order
    std::vector<double> h(4 * (size_t)nc);
    for (int side = 0; side < 2; ++side) {
      const double* src = side ? URight : ULeft;
      for (int k = 0; k < 3; ++k) {
        for (int64_t i = 0; i < nc; ++i)
          for (int v = 0; v < 4; ++v) h[(size_t)v * nc + i] = src[(int64_t)k * 4 * nc + 4 * i + v];
        if ((rc = put_cols(S, (side ? dUR : dUL) + (int64_t)k * 4 * nc, h.data(), 4, false))) return rc;
      }
    }
  } else {  // phase one on the device from the state's cell and node primitives
    RefStepArgs a{};
    a.m = M.v;
    a.prm = S->prm;
    a.Q = S->Q[S->qc];
    a.Qn = S->Qn;
    a.flow_time = flow_time;
    GD_HIP(launch_ref_face_states(a, dUL, dUR, S->stream));
  }
  GD_HIP(launch_ref_cuda_stage(M.v, S->prm.Gamma, precision == GD_FP32_MIXED, dUL, dUR, S->U[S->cur], betta, dt, S->F,
                               S->U[1 - S->cur], S->stream));
  if (iFlux) {
    if ((rc = get_cols(S, S->F, iFlux, 4, false))) return rc;
  } else {
    GD_HIP(hipStreamSynchronize(S->stream));
  }
  return GD_OK;
}

int gd_calc_diff_term(gd_state* S) {
  if (!S) return fail(GD_ERR_INVALID, "null state");
  GD_MATERIALIZE(S);
  int rc = need_ref(S);
  if (rc) return rc;
  const MeshView& v = S->mesh->v;
  GD_HIP(launch_ref_phi(v, S->U[S->cur], S->phiN, S->stream));
  GD_HIP(launch_ref_gg1(v, S->phiN, S->g6, S->stream));
  GD_HIP(launch_ref_gg_nodes(v, S->g6, S->g6N, S->stream));
  GD_HIP(launch_ref_diff_term(v, S->prm, S->Q[S->qc], S->g6N, S->Ud, S->stream));
  GD_HIP(hipStreamSynchronize(S->stream));
  return GD_OK;
}

int gd_update_variables(gd_state* S) {
  if (!S) return fail(GD_ERR_INVALID, "null state");
  GD_MATERIALIZE(S);
  const MeshView& v = S->mesh->v;
  GD_HIP(launch_update_variables(v, S->prm.Gamma, S->U[1 - S->cur], S->U[S->cur], S->Q[S->qc], S->aS, S->vmax, S->Dl,
                                 S->stream));
  GD_HIP(hipStreamSynchronize(S->stream));
  return GD_OK;
}

int gd_cells2nodes(gd_state* S) {
  if (!S) return fail(GD_ERR_INVALID, "null state");
  GD_MATERIALIZE(S);
  const MeshView& v = S->mesh->v;
  GD_HIP(launch_ref_nodes(v, S->Q[S->qc], S->U[S->cur], S->Qn, S->Un, nullptr, S->stream));
  GD_HIP(hipStreamSynchronize(S->stream));
  return GD_OK;
}

// ---------------------------------------------------------------- time loop

namespace {
struct Timer {
  gd_state* S;
  bool on;
  gd_run_stats* st;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> recs;
  size_t next = 0;
  std::vector<std::string> names;
  std::vector<double> bytes;
  hipEvent_t get() {
    if (next < S->ev_pool.size()) return S->ev_pool[next++];
    hipEvent_t e;
    (void)hipEventCreate(&e);
    S->ev_pool.push_back(e);
    ++next;
    return e;
  }
  int kid(const char* name, double b) {
    for (size_t i = 0; i < names.size(); ++i)
      if (names[i] == name) return (int)i;
    names.push_back(name);
    bytes.push_back(b);
    return (int)names.size() - 1;
  }
  template <class F>
  hipError_t run(const char* name, double b, F&& f) {
    if (!on) return f();
    const int k = kid(name, b);
    hipEvent_t e0 = get(), e1 = get();
    (void)hipEventRecord(e0, S->stream);
    hipError_t e = f();
    (void)hipEventRecord(e1, S->stream);
    recs.push_back({k, {e0, e1}});
    return e;
  }
  void finish() {
    if (!on || !st) return;
    st->n_kernels = (int32_t)std::min<size_t>(names.size(), GD_MAX_KERNELS);
    for (int k = 0; k < st->n_kernels; ++k) {
      std::snprintf(st->kernel_names[k], 32, "%s", names[k].c_str());
      st->kernel_ms_total[k] = 0.0;
      st->kernel_launches[k] = 0;
      st->kernel_bytes[k] = bytes[k];
    }
    for (auto& r : recs) {
      if (r.first >= GD_MAX_KERNELS) continue;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, r.second.first, r.second.second);
      st->kernel_ms_total[r.first] += ms;
      st->kernel_launches[r.first] += 1;
    }
  }
};
}  // namespace

// algorithmic HBM bytes per launch (compulsory arrays once per launch at their
// stored width; gathered neighbour/node data counted once at unique size)
static double bytes_ref_step(const MeshView& v, bool visc) {
  const double nc = (double)v.nc, nn = (double)v.nn;
  double b = nc * (32 /*Q own*/ + 12 /*nb*/ + 72 /*nx,ny,len*/ + 24 /*dn1,up1*/ + 8 /*rarea*/ + 8 /*hmin*/ +
                   32 /*U old*/ + 32 /*U new*/ + 32 /*Q new*/) +
             nn * 32 /*node primitives*/;
  if (visc) b += nc * (24 /*ea,eb*/ + 8 /*area*/) + nn * 48 /*g6N*/;
  return b;
}
// the face pass: per mirrored face its owner, neighbour and far-node ids, the
// owner slot's geometry and the flux written; per one-sided entry its cell and
// slot besides; cell and node primitives gathered at unique size
static double bytes_ref_faces(const MeshView& v) {
  const double n1 = (double)(v.n_rfaces - v.n_rmirror);
  return (double)v.n_rmirror * (4 /*cell*/ + 12 /*nb, dn1, up1*/ + 24 /*nx, ny, len*/ + 32 /*flux*/) +
         n1 * (8 /*cell, slot*/ + 12 /*nb, dn1, up1*/ + 24 /*nx, ny, len*/ + 32 /*flux*/) + (double)v.nc * 32 +
         (double)v.nn * 32;
}
// the step with the face pass: every face slot's flux read (unique per entry)
static double bytes_ref_step_faces(const MeshView& v, bool visc) {
  const double nc = (double)v.nc, nn = (double)v.nn;
  double b = nc * (32 /*Q own*/ + 12 /*rcf*/ + 8 /*rarea*/ + 8 /*hmin*/ + 32 /*U old*/ + 32 /*U new*/ + 32 /*Q new*/) +
             (double)v.n_rfaces * 32 /*face fluxes*/;
  if (visc) b += nc * (24 /*ea,eb*/ + 72 /*nx,ny,len*/ + 8 /*area*/) + nn * 48 /*g6N*/;
  return b;
}
static double bytes_ref_nodes(const MeshView& v, bool visc) {
  const double nc = (double)v.nc, nn = (double)v.nn;
  const double inc = 3.0 * nc;  // cell-node incidences of a triangle mesh
  return inc * (4 + 8) /*ids + weights*/ + nc * 32 /*Q*/ + nn * 32 /*Qn*/ + (visc ? nc * 24 + nn * 24 : 0.0);
}
static double bytes_ref_gg1(const MeshView& v) {
  return (double)v.nc * (24 + 72 + 8 + 48) + (double)v.nn * 24;
}
static double bytes_ref_gg_nodes(const MeshView& v) {
  return 3.0 * v.nc * 4 + (double)v.nc * 48 + (double)v.nn * 48;
}
// ns: t = bytes of the compute type (geometry read at t except len, rarea,
// hmin); gathered neighbour data counted once at unique size
static double bytes_ns_grad(const MeshView& v, double t) {
  return (double)v.nc * (4 * t /*q*/ + 12 /*nb*/ + 6 * t /*lsq*/ + 6 * t /*rf*/ + t /*hmin*/ + 16 * t /*record*/);
}
static double bytes_ns_face(const MeshView& v, double t) {
  return (double)v.n_faces * (8 /*cells*/ + 7 * t /*nx,ny,len,r0,r1*/ + 4 * t /*flux write*/) +
         (double)v.nc * 16 * t /*records, gathered at unique size*/;
}
static double bytes_ns_update(const MeshView& v, double t) {
  return (double)v.nc * (4.0 * v.nf /*face ids*/ + 32 /*U old*/ + 32 /*U new*/ + 8 /*1/A*/ + t /*h*/ +
                         4 * t /*q new*/) +
         (double)v.n_faces * 4 * t /*face fluxes, unique*/;
}

int gd_run(gd_state* S, int32_t n_steps, gd_run_params* rp, gd_run_stats* stats) {
  if (!S || !rp) return fail(GD_ERR_INVALID, "null argument");
  if (n_steps < 0) return fail(GD_ERR_INVALID, "negative step count");
  if (rp->dt_policy < GD_DT_FIXED || rp->dt_policy > GD_DT_LOCAL) return fail(GD_ERR_INVALID, "unknown dt policy");
  const MeshView& v = S->mesh->v;
  const gd_params& P = S->prm;
  Timer T{S, rp->time_kernels != 0, stats};
  const bool wall = rp->wall_code != 0 && v.n_wall > 0;
  const bool ns = P.scheme == GD_SCHEME_NS;
  const bool mixed = P.precision == GD_FP32_MIXED;
  const int64_t nparts = (v.nc + kThreads - 1) / kThreads;
  // a run that steps supersedes the previous run's pending Delta / U_NEW
  if (n_steps > 0) S->pend_delta = false;
  // a run that follows an ns run on the same state, with nothing between
  // them, continues from where that one ended (see State::ns_primed)
  const bool warm = ns && S->ns_primed && !S->u_exposed;
  S->ns_primed = false;
  int tq = warm ? S->ns_tq : 0;
  GD_HIP(launch_run_begin(S->acc, rp->flow_time, wall ? S->wall_f : nullptr, v.n_wall, S->drag, S->stream));
  if (ns) {
    // the ns scheme advances U: its compute-type primitives and the first dt
    // inputs come from U in one pass (folded by the first gradient launch),
    // unless the last update left both
    if (!warm)
      GD_HIP(launch_ns_prime(v, P, S->U[S->cur], S->Qt[0], mixed, S->acc, S->part, rp->flow_time, P.viscous,
                             S->stream, false));
    if (n_steps == 0) GD_HIP(launch_acc_fold(S->part, nparts, &S->acc[0], S->stream));
  } else {
    GD_HIP(launch_fields_derived(v.nc, P.Gamma, S->Q[S->qc], S->aS, S->vmax, S->stream));
    GD_HIP(launch_prime_acc(v, P, S->Q[S->qc], S->vmax, S->acc, S->part, rp->flow_time, rp->dt_policy, 0,
                            S->stream));
  }
  if (!ns) {
    const bool visc = P.viscous != 0;
    // the face pass is exact where the two sides' evaluations are exact
    // negatives: AUSM+ with an odd, symmetric limiter (GODUNOV_REF_PER_SIDE=1
    // evaluates every face from each of its cells, for comparison)
    const char* per_side = std::getenv("GODUNOV_REF_PER_SIDE");
    const bool face_pass = v.n_rfaces > 0 && P.flux == GD_FLUX_AUSM_PLUS &&
                           (P.limiter == GD_LIM_REF_HARMONIC || P.limiter == GD_LIM_VAN_ALBADA_A ||
                            P.limiter == GD_LIM_MINMOD) &&
                           !(per_side && per_side[0] == '1');
    double* rflux = static_cast<double*>(S->fflux);  // sized for n_rfaces entries
    if (visc) GD_HIP(T.run("ref_phi", 0.0, [&] { return launch_ref_phi(v, S->U[S->cur], S->phiN, S->stream); }));
    for (int32_t s = 0; s < n_steps; ++s) {
      if (visc) {
        GD_HIP(T.run("ref_gg1", bytes_ref_gg1(v), [&] { return launch_ref_gg1(v, S->phiN, S->g6, S->stream); }));
        GD_HIP(T.run("ref_gg_nodes", bytes_ref_gg_nodes(v),
                     [&] { return launch_ref_gg_nodes(v, S->g6, S->g6N, S->stream); }));
      }
      RefStepArgs a{};
      a.m = v;
      a.prm = P;
      a.Q = S->Q[S->qc];
      a.Qn = S->Qn;
      a.Uo = S->U[S->cur];
      a.g6N = visc ? S->g6N : nullptr;
      a.Un = S->U[1 - S->cur];
      a.Qw = S->Q[1 - S->qc];
      a.acc = S->acc;
      a.part = S->part;
      a.step = s;
      a.dt_policy = rp->dt_policy;
      a.dt = rp->dt;
      a.cfl = rp->cfl;
      a.betta = rp->betta;
      a.include_udiff = 1;
      a.wall_f = wall ? S->wall_f : nullptr;
      a.wall_code = rp->wall_code;
      if (face_pass) {
        GD_HIP(T.run("ref_faces", bytes_ref_faces(v), [&] { return launch_ref_faces(a, rflux, S->stream); }));
        a.rflux = rflux;
      }
      GD_HIP(T.run("ref_step", face_pass ? bytes_ref_step_faces(v, visc) : bytes_ref_step(v, visc),
                   [&] { return launch_ref_step(a, S->stream); }));
      S->cur ^= 1;
      S->qc ^= 1;
      // the node pass also closes the step: dt partials into the next slot,
      // the wall faces' forces into drag
      FoldArgs fa;
      fa.part = S->part;
      fa.nparts = nparts;
      fa.slot = &S->acc[(s + 1) % 3];
      fa.wall_f = wall ? S->wall_f : nullptr;
      fa.n_wall = v.n_wall;
      fa.drag = S->drag;
      GD_HIP(T.run("ref_nodes", bytes_ref_nodes(v, visc), [&] {
        return launch_ref_nodes(v, S->Q[S->qc], S->U[S->cur], S->Qn, nullptr, visc ? S->phiN : nullptr, S->stream,
                                fa);
      }));
    }
  } else {
    const double t = mixed ? 4.0 : 8.0;
    auto grad = mixed ? launch_ns_grad_f32 : launch_ns_grad_f64;
    auto face = mixed ? launch_ns_face_f32 : launch_ns_face_f64;
    auto update = mixed ? launch_ns_update_f32 : launch_ns_update_f64;
    for (int32_t s = 0; s < n_steps; ++s) {
      NsStepArgs a{};
      a.m = v;
      a.prm = P;
      a.Uo = S->U[S->cur];
      a.Un = S->U[1 - S->cur];
      a.Q = S->Qt[tq];
      a.Qw = S->Qt[1 - tq];
      a.rec = S->rec;
      a.fflux = S->fflux;
      a.acc = S->acc;
      a.part = S->part;
      a.step = s;
      a.dt_policy = rp->dt_policy;
      a.dt = rp->dt;
      a.cfl = rp->cfl;
      a.wall_f = wall ? S->wall_f : nullptr;
      a.wall_code = rp->wall_code;
      a.fold_prev = 1;  // step 0 folds the prime pass's partials (wall_f is still zero then)
      a.nparts = nparts;
      a.drag = S->drag;
      GD_HIP(T.run(mixed ? "ns_grad_fp32" : "ns_grad_fp64", bytes_ns_grad(v, t), [&] { return grad(a, S->stream); }));
      GD_HIP(T.run(mixed ? "ns_face_fp32" : "ns_face_fp64", bytes_ns_face(v, t), [&] { return face(a, S->stream); }));
      GD_HIP(T.run(mixed ? "ns_update_fp32" : "ns_update_fp64", bytes_ns_update(v, t),
                   [&] { return update(a, S->stream); }));
      S->cur ^= 1;
      tq ^= 1;
    }
    if (n_steps > 0)  // the last step's reductions (earlier ones ride on the next gradient launch)
      GD_HIP(launch_ns_close(S->part, nparts, &S->acc[n_steps % 3], wall ? S->wall_f : nullptr, v.n_wall, S->drag,
                             S->stream));
    S->pend_derived = true;  // fields2d's FP64 primitives of U, when someone asks
  }
  // the reference's bookkeeping after the step (Delta of the last step,
  // U_NEW == U_OLD) is left pending for the caller that reads it
  if (n_steps > 0) {
    if (rp->residuals)
      GD_HIP(launch_residuals(v, S->U[S->cur], S->U[1 - S->cur], S->red, S->hdr->red, S->stream));
    S->pend_delta = true;
  }
  // one copy brings back the accumulators, the wall force and the residuals
  GD_HIP(hipMemcpyAsync(S->hdr_host, S->hdr, sizeof(RunHdr), hipMemcpyDeviceToHost, S->stream));
  GD_HIP(hipStreamSynchronize(S->stream));
  if (ns) {
    S->ns_primed = true;
    S->ns_tq = tq;
  }
  const RunHdr& H = *S->hdr_host;
  double red[4] = {0, 0, 0, 0}, drag[4] = {0, 0, 0, 0};
  if (rp->residuals && n_steps > 0)
    for (int q = 0; q < 4; ++q) red[q] = H.red[q];
  if (wall)
    for (int q = 0; q < 4; ++q) drag[q] = H.drag[q];
  const Acc& last = H.acc[n_steps % 3];
  rp->flow_time = last.time;
  if (stats) {
    stats->flow_time = last.time;
    stats->last_dt = last.dt;
    for (int q = 0; q < 4; ++q) stats->residuals[q] = red[q];
    stats->wall_force[0] = drag[0];
    stats->wall_force[1] = drag[1];
    stats->wall_force_sum[0] = drag[2];
    stats->wall_force_sum[1] = drag[3];
    stats->nonphysical = (int32_t)last.nonphysical;
    if (!T.on) stats->n_kernels = 0;
  }
  T.finish();
  if (last.nonphysical) {
    char buf[128];
    std::snprintf(buf, sizeof(buf), "%u cells with non-positive density/pressure or NaN", last.nonphysical);
    g_err = buf;
    return GD_ERR_NONPHYSICAL;
  }
  return GD_OK;
}


// ---- discrete adjoint of the reference loop --------------------------------
// The reference-scheme adjoint's incidence lists (RefAdjMaps), built once per
// mesh from its device arrays; flags meshes outside what the adjoint covers.
template <class T>
static int download(const T* dev, int64_t n, std::vector<T>& host) {
  host.resize((size_t)std::max<int64_t>(n, 0));
  if (n > 0) GD_HIP(hipMemcpy(host.data(), dev, (size_t)n * sizeof(T), hipMemcpyDeviceToHost));
  return GD_OK;
}
template <class T>
static void csr(int64_t rows, const std::vector<std::pair<int64_t, T>>& items, std::vector<int64_t>& off,
                std::vector<T>& ref) {
  off.assign((size_t)rows + 1, 0);
  for (const auto& it : items) ++off[(size_t)it.first + 1];
  for (int64_t r = 0; r < rows; ++r) off[(size_t)r + 1] += off[(size_t)r];
  ref.resize(items.size());
  std::vector<int64_t> pos(off.begin(), off.end() - 1);
  for (const auto& it : items) ref[(size_t)pos[(size_t)it.first]++] = it.second;  // stable: items' order per row
}

static int ref_adj_maps(Mesh* M) {
  RefAdjMaps& a = M->radj;
  if (a.built) return GD_OK;
  const MeshView& v = M->v;
  const int64_t nc = v.nc, nn = v.nn, nf = v.nf, nr = v.n_rfaces, nm = v.n_rmirror;
  std::vector<int32_t> nb, dn1, up1, dn2, up2, ea, eb, rfc, rfs, rfnb, rfdn, rfup, cl, nv;
  std::vector<double> w, fnx, fny, flen, area;
  int rc = 0;
  if ((rc = download(v.nb, nf * nc, nb)) || (rc = download(v.dn1, nf * nc, dn1)) || (rc = download(v.up1, nf * nc, up1)) ||
      (rc = download(v.ea, nf * nc, ea)) || (rc = download(v.eb, nf * nc, eb)) || (rc = download(v.rf_cell, nr, rfc)) ||
      (rc = download(v.rf_slot, nr, rfs)) || (rc = download(v.rf_nb, nm, rfnb)) || (rc = download(v.rf_dn1, nm, rfdn)) ||
      (rc = download(v.rf_up1, nm, rfup)) || (rc = download(v.cl, (int64_t)v.ncl * nn, cl)) ||
      (rc = download(v.w, (int64_t)v.ncl * nn, w)) || (rc = download(v.fnx, nf * nc, fnx)) ||
      (rc = download(v.fny, nf * nc, fny)) || (rc = download(v.flen, nf * nc, flen)) ||
      (rc = download(v.area, nc, area)))
    return rc;
  if (v.has_quads) {
    if ((rc = download(v.nv, nc, nv)) || (rc = download(v.dn2, nf * nc, dn2)) || (rc = download(v.up2, nf * nc, up2)))
      return rc;
  } else {
    nv.assign((size_t)nc, 3);
  }
  // every list index below is checked here, so the lists address only rows
  // that exist
  auto in = [](int64_t x, int64_t n) { return x >= 0 && x < n; };
  a.supported = 2 * nr + 1 <= INT32_MAX;  // entry codes 2 f + role are int32
  for (int64_t i = 0; i < nc && a.supported; ++i) {
    a.supported = nv[(size_t)i] == 3 || (nv[(size_t)i] == 4 && v.has_quads && nf == 4);
    for (int64_t k = 0; k < nv[(size_t)i] && a.supported; ++k) {
      const size_t sl = (size_t)k * nc + i;
      a.supported = nb[sl] < nc && in(ea[sl], nn) && in(eb[sl], nn);
      if (a.supported && nb[sl] >= 0) {
        a.supported = in(dn1[sl], nn) && in(up1[sl], nn);
        if (a.supported && nv[(size_t)i] == 4) a.supported = in(dn2[sl], nn) && in(up2[sl], nn);
      }
    }
  }
  for (int64_t f = 0; f < nm && a.supported; ++f)
    a.supported = in(rfnb[(size_t)f], nc) && in(rfdn[(size_t)f], nn) && in(rfup[(size_t)f], nn);
  for (int64_t f = 0; f < nr && a.supported; ++f)
    a.supported = in(rfc[(size_t)f], nc) && rfs[(size_t)f] >= 0 && rfs[(size_t)f] < nv[(size_t)rfc[(size_t)f]];
  for (size_t q = 0; q < cl.size() && a.supported; ++q) a.supported = cl[q] >= -1 && cl[q] < nc;
  // the Green-Gauss coefficient pair of node J in cell i, in the cluster slot
  // of J that holds i: sum over the cell's edges ending at J of
  // 0.5 (-ny, -nx) len, over A_i; a vertex missing from its cluster is outside
  // what the adjoint covers
  std::vector<double> ccx((size_t)v.ncl * nn, 0.0), ccy((size_t)v.ncl * nn, 0.0);
  for (int64_t i = 0; i < nc && a.supported; ++i)
    for (int64_t k = 0; k < nv[(size_t)i] && a.supported; ++k) {
      const size_t sl = (size_t)k * nc + i;
      const double fx = 0.5 * -fny[sl] * flen[sl], fy = 0.5 * -fnx[sl] * flen[sl];
      for (const int64_t J : {(int64_t)ea[sl], (int64_t)eb[sl]}) {
        int q = 0;
        while (q < v.ncl && cl[(size_t)q * nn + J] != i) ++q;
        if (q == v.ncl) { a.supported = false; break; }
        ccx[(size_t)q * nn + J] += fx;
        ccy[(size_t)q * nn + J] += fy;
      }
    }
  a.built = true;
  if (!a.supported) return GD_OK;
  for (size_t q = 0; q < cl.size(); ++q)
    if (cl[q] >= 0) { ccx[q] /= area[(size_t)cl[q]]; ccy[q] /= area[(size_t)cl[q]]; }
  // the entries' far nodes (a quad cell's one-sided entry: both far-side
  // nodes, whose halves the face kernel writes) and cells
  std::vector<std::pair<int64_t, int32_t>> nq, ce;
  nq.reserve((size_t)2 * nr); ce.reserve((size_t)2 * nr);
  M->radj_code.assign((size_t)(nr - nm), 0);
  for (int64_t f = 0; f < nr; ++f) {
    const int64_t i = rfc[(size_t)f];
    ce.push_back({i, (int32_t)(2 * f)});
    if (f < nm) {
      nq.push_back({rfdn[(size_t)f], (int32_t)(2 * f)});
      nq.push_back({rfup[(size_t)f], (int32_t)(2 * f + 1)});
      ce.push_back({rfnb[(size_t)f], (int32_t)(2 * f + 1)});
      continue;
    }
    const size_t sl = (size_t)rfs[(size_t)f] * nc + i;
    M->radj_code[(size_t)(f - nm)] = nb[sl] < 0 ? nb[sl] + 1 : 0;
    if (nb[sl] < 0) continue;  // a boundary face reads only its cell
    nq.push_back({dn1[sl], (int32_t)(2 * f)});
    if (nv[(size_t)i] == 4) nq.push_back({dn2[sl], (int32_t)(2 * f)});
    nq.push_back({up1[sl], (int32_t)(2 * f + 1)});
    if (nv[(size_t)i] == 4) nq.push_back({up2[sl], (int32_t)(2 * f + 1)});
    ce.push_back({nb[sl], (int32_t)(2 * f + 1)});
  }
  struct Memb { int32_t node; double w; };
  std::vector<std::pair<int64_t, Memb>> cm;
  cm.reserve((size_t)v.ncl * nn);
  for (int64_t J = 0; J < nn; ++J) {  // the forward node means' weights, summed in slot order
    double det = 0.0;
    for (int q = 0; q < v.ncl; ++q)
      if (cl[(size_t)q * nn + J] >= 0) det = det + w[(size_t)q * nn + J];
    for (int q = 0; q < v.ncl; ++q) {
      const int32_t c = cl[(size_t)q * nn + J];
      if (c < 0) continue;
      const double wi = w[(size_t)q * nn + J];
      cm.push_back({c, Memb{(int32_t)J, det != 0 ? wi / det : wi}});
    }
  }
  std::vector<int64_t> nq_off, ce_off, cm_off;
  std::vector<int32_t> nq_ref, ce_ref;
  std::vector<Memb> cmr;
  csr(nn, nq, nq_off, nq_ref);
  csr(nc, ce, ce_off, ce_ref);
  csr(nc, cm, cm_off, cmr);
  std::vector<int32_t> cm_node(cmr.size());
  std::vector<double> cm_w(cmr.size());
  for (size_t q = 0; q < cmr.size(); ++q) { cm_node[q] = cmr[q].node; cm_w[q] = cmr[q].w; }
  int64_t* B = &M->device_bytes;
  auto& al = M->allocs;
  if ((rc = upload(al, &a.nq_off, nq_off, B)) || (rc = upload(al, &a.nq_ref, nq_ref, B)) ||
      (rc = upload(al, &a.ccx, ccx, B)) || (rc = upload(al, &a.ccy, ccy, B)) ||
      (rc = upload(al, &a.ce_off, ce_off, B)) || (rc = upload(al, &a.ce_ref, ce_ref, B)) ||
      (rc = upload(al, &a.cm_off, cm_off, B)) || (rc = upload(al, &a.cm_node, cm_node, B)) ||
      (rc = upload(al, &a.cm_w, cm_w, B))) {
    a.built = false;
    return rc;
  }
  return GD_OK;
}

// gd_adjoint_run for the reference scheme: J = sum_n w . F_n + <dJ/dU_N, U_N>
// over N steps of the reference loop at a fixed dt (betta 1), F_n the step's
// force through the faces of wall_code, from the state's U with the loop's
// primitives and node primitives recomputed from it; dJ/dU0 and dJ/dbc of the
// fixed boundary states.
static int ref_adjoint_run(gd_state* S, const gd_adjoint_params* ap, const double* dJ_dUN, double* dJ_dU0,
                           double* dJ_dbc, gd_adjoint_stats* stats) {
  const gd_params& P = S->prm;
  if (P.flux != GD_FLUX_AUSM_PLUS) return fail(GD_ERR_UNSUPPORTED, "the reference-scheme adjoint covers AUSM+");
  if (P.limiter != GD_LIM_REF_HARMONIC && P.limiter != GD_LIM_VAN_ALBADA_A && P.limiter != GD_LIM_MINMOD)
    return fail(GD_ERR_UNSUPPORTED, "the reference-scheme adjoint covers the odd, symmetric limiters");
  if (P.viscous && P.energy_law != GD_ENERGY_CONST)
    return fail(GD_ERR_UNSUPPORTED, "the reference-scheme adjoint covers constant mu and k");
  if (ap->n_steps < 1 || ap->checkpoint_every < 1) return fail(GD_ERR_INVALID, "n_steps and checkpoint_every >= 1");
  if (!(ap->dt > 0.0)) return fail(GD_ERR_INVALID, "the adjoint needs a fixed dt > 0");
  GD_MATERIALIZE(S);
  int rc = 0;
  if ((rc = ref_adj_maps(S->mesh))) return rc;
  const RefAdjMaps& maps = S->mesh->radj;
  if (!maps.supported)
    return fail(GD_ERR_UNSUPPORTED,
                "the reference-scheme adjoint covers cells of 3 or 4 vertices, each in its vertices' clusters");
  const MeshView& v = S->mesh->v;
  const int64_t nc = v.nc, nn = v.nn;
  const bool visc = P.viscous != 0;
  const bool wall = ap->wall_code != 0 && v.n_wall > 0;
  const int64_t nparts = (nc + kThreads - 1) / kThreads;
  const int32_t N = ap->n_steps;
  const int32_t c = std::min(ap->checkpoint_every, N);
  const int32_t nseg = (N + c - 1) / c;
  const double bdt = 1.0 * ap->dt;  // betta 1, the explicit Euler step of the reference drivers
  const double mu[3] = {P.mu_const, P.mu_const, P.mu_const * P.k_const};

  // What the transposed step s reads of the forward loop is Q_s and its node
  // means Qn_s.  They are kept from the forward sweep for every step of the K
  // trailing segments that record_budget holds (steps sK .. N-1); the segments
  // before them keep U at their first step and are recomputed on the way
  // back, writing their steps' Q and Qn into one segment's store.
  const size_t U4 = 4 * (size_t)nc * sizeof(double), N4 = 4 * (size_t)nn * sizeof(double);
  const int32_t K = ap->record_budget > 0
                        ? (int32_t)std::min<int64_t>(nseg, ap->record_budget / (int64_t)((size_t)c * (U4 + N4)))
                        : 0;
  const int32_t sK = std::min(N, (nseg - K) * c);
  const int32_t n_ck = nseg - K;
  const size_t nr4 = 4 * (size_t)std::max<int64_t>(v.n_rfaces, 1) * sizeof(double);
  // workspace: checkpoints, one recomputed segment's Q and Qn, the kept Q and
  // Qn, a scratch U, two adjoint states ([cell][4] records), the entries'
  // adjoints, node- and cell-level adjoints
  const size_t nseg_st = (size_t)(n_ck > 0 ? c : 0), nkeep = (size_t)(N - sK);
  const size_t n1 = (size_t)std::max<int64_t>(v.n_rfaces - v.n_rmirror, 1);  // one-sided entries
  const size_t parts[] = {(size_t)n_ck * U4, nseg_st * U4, nseg_st * N4, nkeep * U4, nkeep * N4, U4,
                          U4, U4, nr4, nr4, nr4, nr4, N4, 6 * (size_t)nn * sizeof(double),
                          6 * (size_t)nc * sizeof(double), N4, 4 * n1 * sizeof(double)};
  constexpr int kParts = (int)(sizeof(parts) / sizeof(parts[0]));
  size_t offs[kParts], need = 0;
  for (int q = 0; q < kParts; ++q) {
    offs[q] = need;
    need += (parts[q] + 255) & ~size_t(255);
  }
  if (need > S->adj_ws_bytes) {
    if (S->adj_ws) (void)hipFree(S->adj_ws);
    S->adj_ws = nullptr;
    S->adj_ws_bytes = 0;
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess) return fail(GD_ERR_NOMEM, std::string("hipMalloc (adjoint workspace): ") + hipGetErrorString(e));
    S->adj_ws = p;
    S->adj_ws_bytes = need;
  }
  S->adj_geo_precision = -1;  // the ns adjoint's slot geometry at the front of the workspace is overwritten
  if (!S->grad0 && (rc = dev_alloc(S->allocs, &S->grad0, 4 * nc))) return rc;
  char* W = static_cast<char*>(S->adj_ws);
  auto at = [&](int q) { return reinterpret_cast<double*>(W + offs[q]); };
  double* ck = at(0);
  double *segQ = at(1), *segQn = at(2), *keepQ = at(3), *keepQn = at(4);
  double* Uscr = at(5);
  double* lam[2] = {at(6), at(7)};
  double *lqo = at(8), *lqn = at(9), *ldn = at(10), *lup = at(11);
  double* lamQn = at(12);
  double* lamG6N = at(13);
  double* tG = at(14);
  double* lamPhiN = at(15);
  double* lbc = at(16);  // fixed states' adjoints per one-sided entry, summed over the steps
  double* rflux = static_cast<double*>(S->fflux);
  const int64_t bytes = (int64_t)(parts[0] + parts[1] + parts[2] + parts[3] + parts[4]);
  const size_t q4 = 4 * (size_t)nc, n4 = 4 * (size_t)nn;  // doubles per Q, per Qn

  // the loop's start from U: primitives, node primitives, cluster means
  auto start = [&](const double* U, double* Q, double* Qn) -> int {
    GD_HIP(launch_update_variables(v, P.Gamma, U, const_cast<double*>(U), Q, S->aS, S->vmax, nullptr, S->stream));
    GD_HIP(launch_ref_nodes(v, Q, U, Qn, nullptr, visc ? S->phiN : nullptr, S->stream));
    return GD_OK;
  };
  // one forward step as gd_run's reference loop takes it, at the fixed dt:
  // (U, Q, Qn) of step s -> (Un, Qw, Qnw) of step s + 1; the forward sweep's
  // steps (fwd) also sum the wall force of the objective's faces
  auto fstep = [&](int32_t s, const double* Uo, const double* Qo, const double* Qno, double* Un, double* Qw,
                   double* Qnw, bool fwd) -> int {
    if (visc) {
      GD_HIP(launch_ref_gg1(v, S->phiN, S->g6, S->stream));
      GD_HIP(launch_ref_gg_nodes(v, S->g6, S->g6N, S->stream));
    }
    RefStepArgs a{};
    a.m = v;
    a.prm = P;
    a.Q = Qo;
    a.Qn = Qno;
    a.Uo = Uo;
    a.g6N = visc ? S->g6N : nullptr;
    a.Un = Un;
    a.Qw = Qw;
    a.acc = S->acc;
    a.part = S->part;
    a.step = s;
    a.dt_policy = GD_DT_FIXED;
    a.dt = ap->dt;
    a.cfl = 0.0;
    a.betta = 1.0;
    a.include_udiff = 1;
    a.wall_f = fwd && wall ? S->wall_f : nullptr;
    a.wall_code = fwd && wall ? ap->wall_code : 0;
    GD_HIP(launch_ref_faces(a, rflux, S->stream));
    a.rflux = rflux;
    GD_HIP(launch_ref_step(a, S->stream));
    FoldArgs fa;  // the step's wall forces summed into drag by the node pass
    if (fwd && wall) {
      fa.part = S->part;
      fa.nparts = nparts;
      fa.slot = &S->acc[(s + 1) % 3];
      fa.wall_f = S->wall_f;
      fa.n_wall = v.n_wall;
      fa.drag = S->drag;
    }
    GD_HIP(launch_ref_nodes(v, Qw, Un, Qnw, nullptr, visc ? S->phiN : nullptr, S->stream, fa));
    return GD_OK;
  };
  // the flow time of each step, as the forward accumulates it
  std::vector<double> ts((size_t)N + 1);
  ts[0] = ap->flow_time;
  for (int32_t s = 0; s < N; ++s) ts[(size_t)s + 1] = ts[(size_t)s] + ap->dt;

  hipEvent_t ev[3];
  for (auto& e : ev) GD_HIP(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int q = 0; q < 3; ++q) (void)hipEventDestroy(e[q]);
    }
  } evg{ev};

  // ---- forward sweep from the state's U: U in the state's two buffers, Q
  // and Qn in the state's buffers before sK and where they are kept from sK
  // on; U_N, Q_N and Qn_N end in the state's buffers
  GD_HIP(hipEventRecord(ev[0], S->stream));
  if (wall) {
    GD_HIP(launch_fill(S->wall_f, 2 * (int64_t)v.n_wall, 0.0, S->stream));
    GD_HIP(launch_fill(S->drag, 4, 0.0, S->stream));
  }
  double* Ucur = S->U[S->cur];
  double* Qcur = sK == 0 ? keepQ : S->Q[S->qc];
  double* Qncur = sK == 0 ? keepQn : S->Qn;
  if ((rc = start(Ucur, Qcur, Qncur))) return rc;
  GD_HIP(launch_fields_derived(nc, P.Gamma, Qcur, S->aS, S->vmax, S->stream));
  GD_HIP(launch_prime_acc(v, P, Qcur, S->vmax, S->acc, S->part, ap->flow_time, GD_DT_FIXED, 0, S->stream));
  for (int32_t s = 0; s < N; ++s) {
    if (s < sK && s % c == 0) GD_HIP(launch_copy(ck + (size_t)(s / c) * q4, Ucur, 4 * nc, S->stream));
    double* Unext = Ucur == S->U[0] ? S->U[1] : S->U[0];
    double *Qnext, *Qnnext;
    if (s + 1 < N && s + 1 >= sK) {
      Qnext = keepQ + (size_t)(s + 1 - sK) * q4;
      Qnnext = keepQn + (size_t)(s + 1 - sK) * n4;
    } else {
      Qnext = Qcur == S->Q[0] ? S->Q[1] : S->Q[0];
      Qnnext = S->Qn;  // read by the step, then rewritten by its node pass
    }
    if ((rc = fstep(s, Ucur, Qcur, Qncur, Unext, Qnext, Qnnext, true))) return rc;
    Ucur = Unext;
    Qcur = Qnext;
    Qncur = Qnnext;
  }
  S->cur = Ucur == S->U[0] ? 0 : 1;
  S->qc = Qcur == S->Q[0] ? 0 : 1;
  GD_HIP(hipEventRecord(ev[1], S->stream));
  double drag[4] = {0, 0, 0, 0};
  if (wall) {
    GD_HIP(hipMemcpyAsync(drag, S->drag, sizeof(drag), hipMemcpyDeviceToHost, S->stream));
    GD_HIP(hipStreamSynchronize(S->stream));
  }
  double J = ap->force_weight[0] * drag[2] + ap->force_weight[1] * drag[3];
  if (dJ_dUN) {
    std::vector<double> UN((size_t)4 * nc);
    if ((rc = get_cols(S, S->U[S->cur], UN.data(), 4, false))) return rc;
    for (size_t q = 0; q < UN.size(); ++q) J += dJ_dUN[q] * UN[q];
  }

  // ---- reverse sweep: each segment's Q and Qn (kept, or recomputed from its
  // checkpoint), then its steps transposed from the last
  int li = 0;
  if (dJ_dUN) {
    if ((rc = put_cols(S, S->grad0, dJ_dUN, 4, false))) return rc;
    GD_HIP(launch_radj_to_rec(nc, S->grad0, lam[0], S->stream));
  } else {
    GD_HIP(launch_fill(lam[0], 4 * nc, 0.0, S->stream));
  }
  GD_HIP(launch_fill(lbc, 4 * (int64_t)n1, 0.0, S->stream));
  int32_t recomputed = 0;
  for (int32_t g = nseg - 1; g >= 0; --g) {
    const int32_t s0 = g * c, s1 = std::min(N, s0 + c);
    const bool kept = s0 >= sK;
    const double* Qb = kept ? keepQ + (size_t)(s0 - sK) * q4 : segQ;
    const double* Qnb = kept ? keepQn + (size_t)(s0 - sK) * n4 : segQn;
    if (!kept) {
      // U ping-pongs between the scratch and the state's free buffer (the
      // state's other buffer holds U_N)
      double* Ua = Uscr;
      double* Ub = S->U[1 - S->cur];
      GD_HIP(launch_copy(Ua, ck + (size_t)g * q4, 4 * nc, S->stream));
      if ((rc = start(Ua, segQ, segQn))) return rc;
      // the step's flow time, which the boundary ghosts read, as the forward had it
      GD_HIP(launch_fill(&S->acc[s0 % 3].time, 1, ts[(size_t)s0], S->stream));
      for (int32_t s = s0; s < s1 - 1; ++s) {
        const size_t k = (size_t)(s - s0);
        if ((rc = fstep(s, Ua, segQ + k * q4, segQn + k * n4, Ub, segQ + (k + 1) * q4, segQn + (k + 1) * n4, false)))
          return rc;
        std::swap(Ua, Ub);
        ++recomputed;
      }
    }
    for (int32_t s = s1 - 1; s >= s0; --s) {
      const double* Qs = Qb + (size_t)(s - s0) * q4;
      const double* Qns = Qnb + (size_t)(s - s0) * n4;
      GD_HIP(launch_radj_face(v, P, Qs, Qns, lam[li], bdt, ts[(size_t)s], wall ? ap->wall_code : 0, ap->force_weight,
                              lqo, lqn, ldn, lup, lbc, S->stream));
      GD_HIP(launch_radj_nodes(v, maps, ldn, lup, lam[li], bdt, mu, visc, lamQn, lamG6N, S->stream));
      if (visc) {
        GD_HIP(launch_radj_g6(v, maps, lamG6N, tG, S->stream));
        GD_HIP(launch_radj_phi(v, maps, tG, lamPhiN, S->stream));
      }
      GD_HIP(launch_radj_cell(v, maps, P.Gamma, lqo, lqn, lamQn, visc ? lamPhiN : nullptr, Qs, lam[li], lam[1 - li],
                              S->stream));
      li ^= 1;
    }
  }
  GD_HIP(hipEventRecord(ev[2], S->stream));
  GD_HIP(launch_radj_from_rec(nc, lam[li], S->grad0, S->stream));
  if (dJ_dU0 && (rc = get_cols(S, S->grad0, dJ_dU0, 4, false))) return rc;
  if (dJ_dbc) {  // the fixed states' adjoints, summed per code in entry order
    for (int q = 0; q < 4 * (GD_MAX_BC + 1); ++q) dJ_dbc[q] = 0.0;
    const std::vector<int32_t>& code = S->mesh->radj_code;
    bool any = false;
    for (const int32_t c0 : code)
      if (c0 < 0 && -c0 <= GD_MAX_BC && P.bc[-c0].kind == GD_BC_FIXED) any = true;
    if (any) {
      std::vector<double> h(4 * code.size());
      GD_HIP(hipMemcpyAsync(h.data(), lbc, h.size() * sizeof(double), hipMemcpyDeviceToHost, S->stream));
      GD_HIP(hipStreamSynchronize(S->stream));
      for (size_t e = 0; e < code.size(); ++e) {
        const int32_t c0 = code[e];
        if (c0 < 0 && -c0 <= GD_MAX_BC && P.bc[-c0].kind == GD_BC_FIXED)
          for (int q = 0; q < 4; ++q) dJ_dbc[4 * (-c0) + q] += h[4 * e + q];
      }
    }
  }
  // the state holds U_N, Q_N and Qn_N as after gd_run (the cluster means,
  // which recomputed segments overwrite, are gd_run's first pass), the
  // bookkeeping pending
  S->pend_delta = true;
  S->pend_derived = true;
  Acc acc[3];
  GD_HIP(hipMemcpyAsync(acc, S->acc, sizeof(acc), hipMemcpyDeviceToHost, S->stream));
  GD_HIP(hipStreamSynchronize(S->stream));
  if (stats) {
    float f_ms = 0.f, r_ms = 0.f;
    (void)hipEventElapsedTime(&f_ms, ev[0], ev[1]);
    (void)hipEventElapsedTime(&r_ms, ev[1], ev[2]);
    stats->J = J;
    stats->force_sum[0] = drag[2];
    stats->force_sum[1] = drag[3];
    stats->recomputed_steps = recomputed;
    stats->n_checkpoints = nseg;
    stats->storage_bytes = bytes;
    stats->ms_forward = f_ms;
    stats->ms_reverse = r_ms;
    stats->flow_time = ap->flow_time + N * ap->dt;
  }
  return GD_OK;
}

// ---- discrete adjoint of the ns loop ---------------------------------------
// Forward sweep from the state's U keeps U every checkpoint_every steps and the
// cell records of the current segment; the reverse sweep walks the segments
// backwards, recomputing each (except the last) from its checkpoint into the
// segment's record slots, then applies the transposed steps.
int gd_adjoint_run(gd_state* S, const gd_adjoint_params* ap, const double* dJ_dUN, double* dJ_dU0, double* dJ_dbc,
                   gd_adjoint_stats* stats) {
  if (!S || !ap) return fail(GD_ERR_INVALID, "null argument");
  const gd_params& P = S->prm;
  if (P.scheme == GD_SCHEME_REF) return ref_adjoint_run(S, ap, dJ_dUN, dJ_dU0, dJ_dbc, stats);
  if (P.scheme != GD_SCHEME_NS) return fail(GD_ERR_UNSUPPORTED, "unknown scheme");
  if (P.flux != GD_FLUX_AUSM_PLUS && P.flux != GD_FLUX_ROE)
    return fail(GD_ERR_UNSUPPORTED, "the adjoint covers the AUSM+ and Roe fluxes");
  if (ap->n_steps < 1 || ap->checkpoint_every < 1) return fail(GD_ERR_INVALID, "n_steps and checkpoint_every >= 1");
  if (!(ap->dt > 0.0)) return fail(GD_ERR_INVALID, "the adjoint needs a fixed dt > 0");
  GD_MATERIALIZE(S);
  const MeshView& v = S->mesh->v;
  const int64_t nc = v.nc;
  const bool mixed = P.precision == GD_FP32_MIXED;
  const size_t t = mixed ? 4 : 8;
  const int32_t N = ap->n_steps;
  const int32_t c = std::min(ap->checkpoint_every, N);
  const int32_t nseg = (N + c - 1) / c;
  const int64_t nparts = (nc + kThreads - 1) / kThreads;
  const bool wall = ap->wall_code != 0 && v.n_wall > 0;
  const int64_t nw = std::max<int32_t>(v.n_wall, 1);

  if (ap->record_budget < 0) return fail(GD_ERR_INVALID, "negative record budget");
  // records kept for the last `keep` segments; earlier segments are recomputed
  const size_t seg_bytes = (size_t)c * 16 * nc * t;
  const int32_t keep = (int32_t)std::max<int64_t>(1, std::min<int64_t>(nseg, ap->record_budget / (int64_t)seg_bytes));
  // one workspace per state, kept between calls and grown when a call needs
  // more (an optimisation loop calls the adjoint again and again): the slot
  // geometry first (valid across calls of one precision), then the rest
  const size_t geo_bytes = kAdjSlotGeo * (size_t)v.nf * nc * t;
  const size_t ck_bytes = (size_t)nseg * 4 * nc * sizeof(double);
  const size_t rec_bytes = (size_t)keep * seg_bytes;
  const size_t parts[] = {geo_bytes, ck_bytes, rec_bytes, 4 * nc * sizeof(double), 4 * nc * sizeof(double),
                          4 * nc * sizeof(double), 4 * nc * sizeof(double), 4 * nw * sizeof(double),
                          4 * nc * t, 4 * nc * t, 4 * (size_t)v.nf * nc * t};
  constexpr int kParts = (int)(sizeof(parts) / sizeof(parts[0]));
  size_t offs[kParts], need = 0;
  for (int q = 0; q < kParts; ++q) {
    offs[q] = need;
    need += (parts[q] + 255) & ~size_t(255);
  }
  if (need > S->adj_ws_bytes) {
    if (S->adj_ws) (void)hipFree(S->adj_ws);
    S->adj_ws = nullptr;
    S->adj_ws_bytes = 0;
    S->adj_geo_precision = -1;
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess) return fail(GD_ERR_NOMEM, std::string("hipMalloc (adjoint workspace): ") + hipGetErrorString(e));
    S->adj_ws = p;
    S->adj_ws_bytes = need;
  }
  char* W = static_cast<char*>(S->adj_ws);
  char* sgeo = W + offs[0];
  double* ck = reinterpret_cast<double*>(W + offs[1]);
  char* recs = W + offs[2];
  double* Uw0 = reinterpret_cast<double*>(W + offs[3]);
  double* Uw1 = reinterpret_cast<double*>(W + offs[4]);
  double* lU0 = reinterpret_cast<double*>(W + offs[5]);
  double* lU1 = reinterpret_cast<double*>(W + offs[6]);
  double* bcsum = reinterpret_cast<double*>(W + offs[7]);
  char* lamF = W + offs[8];
  char* selfq = W + offs[9];
  char* nrev = W + offs[10];
  const int64_t bytes = (int64_t)(ck_bytes + rec_bytes);  // checkpoints + kept records
  int rc = 0;
  if (S->adj_geo_precision != P.precision) {
    GD_HIP((mixed ? launch_adj_slot_geo_f32 : launch_adj_slot_geo_f64)(v, sgeo, S->stream));
    S->adj_geo_precision = P.precision;
  }
  // record slot of step s: kept segments have their own region; the others
  // (recomputed later) share the first kept region, which only later steps
  // of the forward sweep overwrite and which the reverse sweep no longer needs
  // when it reaches them
  // in the forward sweep, a step of a segment that the reverse sweep
  // recomputes writes its record into the state's own record buffer, reused
  // step after step as gd_run does (one buffer rather than one per step of a
  // 2.7 GB region: the sweep at the 16 GiB budget 30.5 -> 30.0 ms, with no
  // budget 30.5 -> 29.2 ms)
  bool fwd_sweep = true;
  auto rec_slot = [&](int32_t s) {
    const int32_t kidx = std::max<int32_t>(0, s / c - (nseg - keep));
    return (void*)(recs + (size_t)kidx * seg_bytes + (size_t)(s % c) * 16 * nc * t);
  };

  auto prims = mixed ? launch_ns_prims_f32 : launch_ns_prims_f64;
  auto grad = mixed ? launch_ns_grad_f32 : launch_ns_grad_f64;
  auto face = mixed ? launch_ns_face_f32 : launch_ns_face_f64;
  auto update = mixed ? launch_ns_update_f32 : launch_ns_update_f64;
  auto acell = mixed ? launch_adj_cell_f32 : launch_adj_cell_f64;
  auto astate = mixed ? launch_adj_state_f32 : launch_adj_state_f64;

  // one forward step s from Uo into Un, the record into its segment slot;
  // records_only: the record alone (the last step of a recomputed segment,
  // whose U_{s+1} is the next checkpoint)
  auto fstep = [&](int32_t s, const double* Uo, double* Un, int tq, bool fold, bool with_wall,
                   bool records_only = false) -> int {
    NsStepArgs a{};
    a.m = v;
    a.prm = P;
    a.Uo = Uo;
    a.Un = Un;
    a.Q = S->Qt[tq];
    a.Qw = S->Qt[1 - tq];
    a.rec = (fwd_sweep && s / c < nseg - keep) ? S->rec : rec_slot(s);
    a.fflux = S->fflux;
    a.acc = S->acc;
    a.part = S->part;
    a.step = s;
    a.dt_policy = GD_DT_FIXED;
    a.dt = ap->dt;
    a.cfl = 0.0;
    a.wall_f = with_wall ? S->wall_f : nullptr;
    a.wall_code = ap->wall_code;
    a.fold_prev = fold ? 1 : 0;
    a.nparts = nparts;
    a.drag = S->drag;
    GD_HIP(grad(a, S->stream));
    if (records_only) return GD_OK;
    GD_HIP(face(a, S->stream));
    GD_HIP(update(a, S->stream));
    return GD_OK;
  };

  hipEvent_t ev[3];
  for (auto& e : ev) GD_HIP(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int q = 0; q < 3; ++q) (void)hipEventDestroy(e[q]);
    }
  } evg{ev};

  // ---- forward sweep (as gd_run: the prime pass, each gradient launch
  // folding the step before it)
  GD_HIP(hipEventRecord(ev[0], S->stream));
  GD_HIP(launch_fill(S->wall_f, 2 * nw, 0.0, S->stream));
  GD_HIP(launch_fill(S->drag, 4, 0.0, S->stream));
  int tq = 0;
  GD_HIP(launch_ns_prime(v, P, S->U[S->cur], S->Qt[tq], mixed, S->acc, S->part, ap->flow_time, P.viscous,
                         S->stream));
  for (int32_t s = 0; s < N; ++s) {
    if (s % c == 0) GD_HIP(launch_copy(ck + (size_t)(s / c) * 4 * nc, S->U[S->cur], 4 * nc, S->stream));
    if ((rc = fstep(s, S->U[S->cur], S->U[1 - S->cur], tq, true, wall))) return rc;
    S->cur ^= 1;
    tq ^= 1;
  }
  fwd_sweep = false;
  GD_HIP(launch_acc_fold(S->part, nparts, &S->acc[N % 3], S->stream));
  if (wall) GD_HIP(launch_wall_sum(S->wall_f, v.n_wall, S->drag, S->stream));
  GD_HIP(hipEventRecord(ev[1], S->stream));
  Acc acc[3];
  double drag[4] = {0, 0, 0, 0};
  GD_HIP(hipMemcpyAsync(acc, S->acc, sizeof(acc), hipMemcpyDeviceToHost, S->stream));
  GD_HIP(hipMemcpyAsync(drag, S->drag, sizeof(drag), hipMemcpyDeviceToHost, S->stream));
  GD_HIP(hipStreamSynchronize(S->stream));
  if (acc[N % 3].nonphysical) {
    char buf[128];
    std::snprintf(buf, sizeof(buf), "%u cells with non-positive density/pressure or NaN", acc[N % 3].nonphysical);
    return fail(GD_ERR_NONPHYSICAL, buf);
  }
  double J = ap->force_weight[0] * drag[2] + ap->force_weight[1] * drag[3];
  if (dJ_dUN) {  // + <dJ/dU_N, U_N>
    std::vector<double> UN((size_t)4 * nc);
    if ((rc = get_cols(S, S->U[S->cur], UN.data(), 4, false))) return rc;
    for (size_t q = 0; q < UN.size(); ++q) J += dJ_dUN[q] * UN[q];
  }

  // ---- reverse sweep
  double* lU[2] = {lU0, lU1};
  int li = 0;
  if (dJ_dUN) {
    if ((rc = put_cols(S, lU[0], dJ_dUN, 4, false))) return rc;
  } else {
    GD_HIP(launch_fill(lU[0], 4 * nc, 0.0, S->stream));
  }
  GD_HIP(launch_fill(bcsum, 4 * nw, 0.0, S->stream));
  AdjStepArgs r{};
  r.m = v;
  r.prm = P;
  r.lamF = lamF;
  r.selfq = selfq;
  r.nrev = nrev;
  r.sgeo = sgeo;
  r.bcsum = bcsum;
  r.dt = ap->dt;
  r.wall_code = wall ? ap->wall_code : 0;
  r.fw[0] = ap->force_weight[0];
  r.fw[1] = ap->force_weight[1];
  r.rec = rec_slot(N - 1);
  r.lamU_in = lU[li];
  GD_HIP(astate(r, false, S->stream));  // lamF of the last step from dJ/dU_N
  int32_t recomputed = 0;
  for (int32_t seg = nseg - 1; seg >= 0; --seg) {
    const int32_t s0 = seg * c, s1 = std::min(N, s0 + c);
    if (seg < nseg - keep) {
      GD_HIP(launch_copy(Uw0, ck + (size_t)seg * 4 * nc, 4 * nc, S->stream));
      double* Uw[2] = {Uw0, Uw1};
      int wc = 0, rq = 0;
      // the primitives of the checkpoint exactly as the forward sweep had them:
      // from the prime pass at step 0, from the update pass after
      if (s0 == 0)
        GD_HIP(launch_ns_prime(v, P, Uw[wc], S->Qt[rq], mixed, S->acc, S->part, ap->flow_time, P.viscous, S->stream));
      else
        GD_HIP(prims(nc, P.Gamma, Uw[wc], S->Qt[rq], S->stream));
      for (int32_t s = s0; s < s1; ++s) {
        if ((rc = fstep(s, Uw[wc], Uw[1 - wc], rq, false, false, s == s1 - 1))) return rc;
        wc ^= 1;
        rq ^= 1;
        ++recomputed;
      }
    }
    for (int32_t s = s1 - 1; s >= s0; --s) {
      r.rec = rec_slot(s);
      r.lamU_in = lU[li];
      r.lamU_out = lU[1 - li];
      GD_HIP(acell(r, S->stream));
      GD_HIP(astate(r, true, S->stream));
      li ^= 1;
    }
  }
  GD_HIP(hipEventRecord(ev[2], S->stream));
  if (!S->grad0 && (rc = dev_alloc(S->allocs, &S->grad0, 4 * nc))) return rc;
  GD_HIP(launch_copy(S->grad0, lU[li], 4 * nc, S->stream));
  if (dJ_dU0 && (rc = get_cols(S, lU[li], dJ_dU0, 4, false))) return rc;
  if (dJ_dbc) {
    std::vector<double> bs((size_t)4 * nw);
    GD_HIP(hipMemcpyAsync(bs.data(), bcsum, bs.size() * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    GD_HIP(hipStreamSynchronize(S->stream));
    for (int q = 0; q < 4 * (GD_MAX_BC + 1); ++q) dJ_dbc[q] = 0.0;
    const std::vector<int32_t>& wc = S->mesh->wall_codes;
    for (int32_t w = 0; w < v.n_wall; ++w) {
      const int idx = -wc[w];
      if (idx < 0 || idx > GD_MAX_BC || P.bc[idx].kind != GD_BC_FIXED) continue;
      for (int q = 0; q < 4; ++q) dJ_dbc[4 * idx + q] += bs[4 * (size_t)w + q];
    }
  }
  // the state holds U_N, as after gd_run (its bookkeeping pending likewise)
  S->pend_delta = N > 0;
  S->pend_derived = true;
  GD_HIP(hipStreamSynchronize(S->stream));
  if (stats) {
    float f_ms = 0.f, r_ms = 0.f;
    (void)hipEventElapsedTime(&f_ms, ev[0], ev[1]);
    (void)hipEventElapsedTime(&r_ms, ev[1], ev[2]);
    stats->J = J;
    stats->force_sum[0] = drag[2];
    stats->force_sum[1] = drag[3];
    stats->recomputed_steps = recomputed;
    stats->n_checkpoints = nseg;
    stats->storage_bytes = bytes;
    stats->ms_forward = f_ms;
    stats->ms_reverse = r_ms;
    stats->flow_time = acc[N % 3].time;
  }
  return GD_OK;
}
