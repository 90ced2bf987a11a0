// Internal device data model of libgodunov (not part of the C ABI).
//
// HBM layout: structure-of-arrays, 0-based, in device cell order.  Per-face
// arrays are stored face-slot-major, [slot][cell], so that thread i of a
// wavefront reads element i of each slot: every load of the cell-centric
// kernels is a unit-stride 64-lane access except the neighbour gathers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/godunov.h"

namespace gd {

// MI355X dispatches workgroup b to XCD (b mod 8).  Remapping the launch order
// so that each XCD receives one contiguous range of logical blocks keeps the
// neighbours of a cell range (which, in a locality-preserving order such as
// Hilbert or RCM, sit in nearby blocks) in the same 4 MB L2 instead of
// spreading the halo over eight L2s.
constexpr unsigned kNumXcd = 8;
// threads per block of every cell / face / node kernel (the per-block partial
// reductions are sized by it)
constexpr int kThreads = 256;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned x = b % kNumXcd, k = b / kNumXcd;
  const unsigned q = nb / kNumXcd, r = nb % kNumXcd;
  return x * q + (x < r ? x : r) + k;
}

typedef unsigned int stage_u32x4 __attribute__((ext_vector_type(4)));
// One output record of K 16 B pieces per lane, stored as the wave's contiguous
// range through a per-wave LDS slab, so that each store instruction writes
// 1 KB of whole lines.  item is the lane's record index, and a wave's lanes
// hold consecutive items (block base + threadIdx.x).  Every lane of the wave
// takes part; records at or past n (lanes standing in past the end) are not
// stored.  A lane's own 32 or 48 B record stored directly splits lines between
// store instructions; staged, the reference loop's 2M TGV step's face,
// first-gradient and node-gradient passes go 72.2 -> 69.2, 68.5 -> 58.6 and
// 42.1 -> 38.7 us (step 0.3501 -> 0.3350 ms), and its adjoint's reverse sweep
// of 100 steps 65.7 -> 64.0 ms, same bits.
template <int K>
struct WaveStage {
  stage_u32x4 slab[kThreads / 64][64 * K];
  __device__ __forceinline__ void store(double* base, int64_t item, int64_t n, const double* o) {
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    const int lane = threadIdx.x & 63;
    stage_u32x4* S = slab[threadIdx.x >> 6];
#pragma unroll
    for (int q = 0; q < K; ++q) S[lane * K + q] = __builtin_bit_cast(stage_u32x4, f64x2{o[2 * q], o[2 * q + 1]});
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int64_t w0 = item - lane;
    stage_u32x4* d = reinterpret_cast<stage_u32x4*>(base) + w0 * K;
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int idx = q * 64 + lane;
      if (w0 + idx / K < n) d[idx] = S[idx];
    }
  }
};

// Everything a kernel needs about the mesh, passed by value as a kernel argument.
struct MeshView {
  int64_t nc, nn;
  int32_t nf;          // face slots stored: 3 (triangles only) or 4
  int32_t ncl;         // cluster width stored
  int32_t has_quads;
  int32_t n_wall;      // faces in the wall list
  // per cell (nc) unless noted
  const int32_t* nv;   // vertices per cell (3/4)
  const int32_t* nb;   // [nf][nc] neighbour (>= 0) or (code - 1) (< 0) for a boundary face
  const double* fnx;   // [nf][nc] inward unit normal
  const double* fny;
  const double* flen;  // [nf][nc] edge length
  const int32_t* dn1;  // [nf][nc] own far-side node(s) of the face (node2cellsL2down)
  const int32_t* dn2;
  const int32_t* up1;  // [nf][nc] neighbour far-side node(s) (node2cellsL2up)
  const int32_t* up2;
  const int32_t* ea;   // [nf][nc] edge end nodes (cells2nodes)
  const int32_t* eb;
  const double* area;
  const double* rarea; // 1/A
  const double* hx;    // longest edge (HX)
  const double* hmin;  // 2A / perimeter
  const float* hmin32;
  const double* xc;    // centroid (cell_mid_points)
  const double* yc;
  // ns geometry
  const double* lsq;   // [nf][2][nc] least-squares weights: grad = sum_k w_k (q_nb - q_i)
  const double* rf;    // [nf][2][nc] face midpoint - own centroid
  const int32_t* wall_id; // [nf][nc] index into the wall list or -1
  const int32_t* kn;      // [nf][nc] the neighbour's slot of the shared face, -1 at a boundary
  // float copies for the FP32-mixed kernels of values that those kernels round
  // to float at first use (bit-identical results, half the bytes)
  const float* fnx32;  // [nf][nc]
  const float* fny32;
  const float* lsq32;  // [nf][2][nc]
  const float* rf32;   // [nf][2][nc]
  // face list of the ns flux pass: each face once, owned by the cell with the
  // lower caller id; the owner's inward normal; [2][n_faces] pairs are x, y
  int64_t n_faces;
  const int32_t* fc0;   // owner cell
  const int32_t* fc1;   // neighbour cell, or (code - 1) < 0 at a boundary
  const int32_t* fwall; // wall-list index of a boundary face, or -1
  const double* f_nx;
  const double* f_ny;
  const double* f_len;
  const double* f_r0;   // [2][n_faces] face midpoint - owner centroid
  const double* f_r1;   // [2][n_faces] face midpoint - neighbour centroid (its own vertices)
  const float* f_nx32;
  const float* f_ny32;
  const float* f_r032;
  const float* f_r132;
  const float* f_len32;
  const int32_t* cface; // [nf][nc]: 2 f + (1 if the cell is the face's neighbour), -1 for no face
  // reference-scheme face list, the entries of the loop's face pass.  First
  // the n_rmirror interior faces between two triangles whose two sides
  // mirror each other exactly (inward normals exact negatives, equal lengths,
  // each side's far vertex the other's), evaluated once from the owner (lower
  // caller id) and numbered in owner order; then one entry per remaining
  // (cell, slot) — boundary faces, faces of quads, faces whose two sides do
  // not mirror bit for bit — evaluated from that cell's side.
  int64_t n_rfaces;
  int64_t n_rmirror;
  const int32_t* rf_cell;  // owner cell / the cell of a one-sided entry
  const int32_t* rf_slot;  // its slot of the face
  const int32_t* rcf;      // [nf][nc]: 2 f + (1 if the cell is a mirrored face's non-owner), -1 for no face
  // the mirrored faces' owner-slot data in face order: neighbour, far nodes
  // of the owner and of the neighbour side, owner's inward normal and length
  const int32_t* rf_nb;
  const int32_t* rf_dn1;
  const int32_t* rf_up1;
  const double* rf_nx;
  const double* rf_ny;
  const double* rf_len;
  // per node
  const int32_t* cl;   // [ncl][nn] cluster cell ids, -1 padded
  const double* w;     // [ncl][nn] interpolation weights
  const double* xn;
  const double* yn;
  double max_hx;
};

// Incidence lists of the reference-scheme adjoint, built from the mesh on its
// first use (compressed rows: row r's items are off[r] .. off[r+1]-1, in a
// fixed order, so every transposed sum is deterministic).
struct RefAdjMaps {
  bool built = false;
  bool supported = false;      // every index in range
  const int64_t* nq_off;       // per node: the face-list entries that read it as a far node
  const int32_t* nq_ref;       //   2 f + (1 for the neighbour side's far node "up", 0 for "down")
  const double* ccx;           // per cluster slot [ncl][nn] (cell cl[q nn + J]): the Green-Gauss
  const double* ccy;           //   coefficient pair of the node in that cell (0 if not a vertex)
  const int64_t* ce_off;       // per cell: the face-list entries that read its primitives
  const int32_t* ce_ref;       //   2 f + (0 as the entry's own cell, 1 as its neighbour)
  const int64_t* cm_off;       // per cell: the node clusters it belongs to
  const int32_t* cm_node;
  const double* cm_w;          //   weight / the node's weight sum (node primitives)
};

struct Mesh {
  MeshView v{};
  RefAdjMaps radj{};
  std::vector<int32_t> radj_code;  // per one-sided face-list entry: its boundary code (< 0), else 0
  std::vector<void*> allocs;
  std::vector<int64_t> cperm, ciperm;  // cells new->old, old->new
  std::vector<int64_t> nperm, niperm;  // nodes new->old, old->new
  int64_t n_boundary_faces = 0;
  int64_t device_bytes = 0;
  int64_t bandwidth = 0;
  int32_t max_cluster = 0;
  int device = 0;
  std::vector<int32_t> wall_codes;  // boundary code of each wall-list entry
  std::vector<int32_t> nv_caller;   // vertices per cell in the caller's numbering
  // the same permutations on the device, for column transfers; the identity
  // is flagged and skipped
  const int32_t *d_cperm = nullptr, *d_ciperm = nullptr, *d_nperm = nullptr, *d_niperm = nullptr;
  bool cells_identity = false, nodes_identity = false;
};

// Run-time accumulator slot, rotated over three buffers (see gd_run).
struct Acc {
  unsigned long long vmax_bits;   // max over cells of |u|+a (positive doubles compare as integers)
  unsigned long long hdt_bits;    // min over cells of local dt / CFL
  double time;
  double dt;
  unsigned int nonphysical;
  unsigned int pad;
  double force[2];
};

// What a run reads back at its end, in one device block so that one copy into
// pinned host memory brings it all: the accumulator slots, the wall force
// (last step x, y; running sums x, y) and the residual norms.
struct RunHdr {
  Acc acc[3];
  double drag[4];
  double red[4];
};

struct State {
  Mesh* mesh = nullptr;
  gd_params prm{};
  hipStream_t stream = nullptr;
  double* U[2] = {nullptr, nullptr};   // [4][nc]
  double* Q[2] = {nullptr, nullptr};   // [4][nc] primitives rho,u,v,p
  int cur = 0;                          // U[cur] is U_OLD, U[1-cur] is U_NEW
  int qc = 0;                           // Q[qc] are the current cell primitives
  double* aS = nullptr;                 // [nc]
  double* vmax = nullptr;               // [nc]
  double* Qn = nullptr;                 // [4][nn]
  double* Un = nullptr;                 // [4][nn]
  double* Ud = nullptr;                 // [4][nc]
  double* F = nullptr;                  // [4][nc]
  double* Dl = nullptr;                 // [4][nc]
  double* phiN = nullptr;               // [nn][4]: rho*u, rho*v, E cluster means, 0
  double* g6 = nullptr;                 // [nc][6]: Green-Gauss level-1 gradients per cell
  double* g6N = nullptr;                // [nn][6]: their cluster means per node
  void* rec = nullptr;                  // ns: per-cell record [nc][16] (q, grad q, limiter), compute type
  void* fflux = nullptr;                // ns: [n_faces][4] face fluxes, compute type
  void* Qt[2] = {nullptr, nullptr};     // ns: primitives [nc][4], compute type
  double* wall_f = nullptr;             // [n_wall][2]
  RunHdr* hdr = nullptr;                // device: acc, drag, residual norms
  RunHdr* hdr_host = nullptr;           // pinned host copy of hdr
  double* drag = nullptr;               // hdr->drag: last step force x,y; running sums
  Acc* acc = nullptr;                   // hdr->acc [3]
  double* part = nullptr;               // [3 x cell blocks] partial reductions
  double* red = nullptr;                // reduction scratch (per-block residual partials)
  double* xfer = nullptr;               // [4][max(nc, nn)] staging of host transfers
  double* grad0 = nullptr;              // [4][nc] dJ/dU0 of the last adjoint run (allocated by the first)
  double* fstate = nullptr;             // [2][4 slots][4][nc] two-phase face states (allocated by the first request)
  void* adj_ws = nullptr;                // adjoint workspace, kept between calls (grown when needed)
  size_t adj_ws_bytes = 0;
  int32_t adj_geo_precision = -1;       // precision of the slot geometry held in it, -1: none
  // bookkeeping a run leaves to whoever reads it (done by the next call that does):
  bool pend_delta = false;    // Dl = U[cur] - U[1-cur] of the last step, then U[1-cur] = U[cur]
  bool pend_derived = false;  // Q[qc], aS, vmax from U[cur] (updateVariablesSA)
  // The ns loop may start where the last one ended: Qt[ns_tq] holds the
  // compute-type primitives of U[cur] and part the dt partials of U[cur], as
  // the last k_ns_update wrote them.  Set by an ns gd_run, cleared by every
  // other entry point on the state (materialize) and never trusted once
  // gd_state_device_u has handed U to the caller, who may write it.
  bool ns_primed = false;
  int ns_tq = 0;
  bool u_exposed = false;
  std::vector<void*> allocs;
};

// ---- shared by the C ABI's translation units (gd_api.hip) ----
int set_error(int code, const std::string& msg);  // records gd_last_error's message, returns code
int check_params(const Mesh* m, const gd_params* p);

// ---- kernel launchers (gd_ref_kernels.hip) ----
struct RefStepArgs {
  MeshView m;
  gd_params prm;
  const double* Q;      // cell primitives (read)
  const double* Qn;     // node primitives
  const double* Uo;     // U_OLD
  const double* Ud;     // U_DIFF (stage variant) or nullptr
  const double* g6N;    // node means of gradients (fused viscous) or nullptr
  double* Un;           // U_NEW
  double* F;            // FLUXES or nullptr
  double* Qw;           // fused update: new primitives, or nullptr
  double* aS;
  double* vmax;
  Acc* acc;             // fused run: accumulators (3 slots) or nullptr
  double* part;         // fused run: per-block partials of the dt inputs [3 x blocks]
  int32_t step;         // rotates the accumulator slots
  int32_t dt_policy;
  double dt, cfl, betta, flow_time;
  int32_t include_udiff;
  double* wall_f;
  int32_t wall_code;
  // face pass: the fluxes of the listed faces (MeshView::rcf), or nullptr:
  // every face evaluated from each of its cells
  const double* rflux;
};

// what a launch closes of the step before it (see k_ref_nodes); slot null: nothing
struct FoldArgs {
  const double* part = nullptr;   // per-block dt partials
  int64_t nparts = 0;
  Acc* slot = nullptr;            // the accumulator slot they fold into
  const double* wall_f = nullptr; // wall-face forces of the step (nullptr: no wall)
  int32_t n_wall = 0;
  double* drag = nullptr;         // {fx, fy, running sums}
};

hipError_t launch_ref_step(const RefStepArgs& a, hipStream_t s);
hipError_t launch_ref_face_states(const RefStepArgs& a, double* UL, double* UR, hipStream_t s);
// the face pass: the flux of every listed face from its owner, into rflux [n_rfaces][4]
hipError_t launch_ref_faces(const RefStepArgs& a, double* rflux, hipStream_t s);
// calcOneStageCUDA (x32 false) / calcOneStageCUDAx32 from face states UL/UR
// [slot][4][nc]: F = iFlux [4][nc], Un = Uo - F betta dt / A
hipError_t launch_ref_cuda_stage(const MeshView& m, double gamma, bool x32, const double* UL, const double* UR,
                                 const double* Uo, double betta, double dt, double* F, double* Un, hipStream_t s);
hipError_t launch_ref_nodes(const MeshView& m, const double* Q, const double* U, double* Qn, double* Un,
                            double* phiN, hipStream_t s, const FoldArgs& fold = FoldArgs{});
// start of the ns loop: resets the accumulators to flow_time (unless
// reset_acc is false), writes the compute-type primitives of U into Qt and the
// first step's dt partials into part (folded by the first gradient launch)
hipError_t launch_ns_prime(const MeshView& m, const gd_params& prm, const double* U, void* Qt, bool f32, Acc* acc,
                           double* part, double flow_time, int32_t viscous_ns, hipStream_t s, bool reset_acc = true);
// start of a run, one block: the accumulator slots reset to flow_time, the
// wall-face forces (2 n_wall values, when wall_f is given) and drag zeroed
hipError_t launch_run_begin(Acc* acc, double flow_time, double* wall_f, int32_t n_wall, double* drag, hipStream_t s);
// end of the ns loop, one launch: the last update's dt partials into slot and,
// when wall_f is given, the last face pass's wall forces into drag
hipError_t launch_ns_close(const double* part, int64_t nparts, Acc* slot, const double* wall_f, int32_t n_wall,
                           double* drag, hipStream_t s);
hipError_t launch_ref_phi(const MeshView& m, const double* U, double* phiN, hipStream_t s);
hipError_t launch_ref_gg1(const MeshView& m, const double* phiN, double* g6, hipStream_t s);
hipError_t launch_ref_gg_nodes(const MeshView& m, const double* g6, double* g6N, hipStream_t s);
hipError_t launch_ref_diff_term(const MeshView& m, const gd_params& prm, const double* Q, const double* g6N,
                                double* Ud, hipStream_t s);
hipError_t launch_update_variables(const MeshView& m, double gamma, const double* Unew, double* Uold, double* Q,
                                   double* aS, double* vmax, double* Delta, hipStream_t s);
hipError_t launch_prime_acc(const MeshView& m, const gd_params& prm, const double* Q, const double* vmax, Acc* acc,
                            double* part, double flow_time, int32_t dt_policy, int32_t viscous_ns, hipStream_t s);
hipError_t launch_acc_fold(const double* part, int64_t nparts, Acc* slot, hipStream_t s);
// sums of (Unew - Uold)^2 per variable into out[0..3], through per-block partials in scratch
hipError_t launch_residuals(const MeshView& m, const double* Unew, const double* Uold, double* scratch, double* out,
                            hipStream_t s);
hipError_t launch_wall_sum(const double* wall_f, int32_t n, double* out4, hipStream_t s);
hipError_t launch_fields_derived(int64_t nc, double gamma, const double* Q, double* aS, double* vmax, hipStream_t s);
hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s);
hipError_t launch_copy(double* dst, const double* src, int64_t n, hipStream_t s);
hipError_t launch_permute(double* dst, const double* src, const int32_t* idx, int64_t n, int32_t ncols,
                          hipStream_t s);
hipError_t launch_delta(int64_t nc, const double* Un, const double* Uo, double* D, hipStream_t s);

// ---- ns scheme (gd_ns_fp64.hip, gd_ns_fp32.hip) ----
struct NsStepArgs {
  MeshView m;
  gd_params prm;
  const double* Uo;
  double* Un;
  const void* Q;        // current primitives [nc][4], compute type
  void* Qw;             // next primitives
  void* rec;            // [nc][16]: q, (d/dx, d/dy) of rho,u,v,p, limiter of rho,u,v,p
  void* fflux;          // [n_faces][4] flux through each face, leaving its owner
  Acc* acc;
  double* part;         // per-block partials of the dt inputs [3 x blocks]
  int32_t step;
  int32_t dt_policy;
  double dt, cfl;
  double* wall_f;
  int32_t wall_code;
  // fold_prev != 0: two extra blocks of the gradient launch fold the previous
  // update's partials into acc[step % 3] and sum wall_f (the previous face
  // pass) into drag
  int32_t fold_prev;
  int64_t nparts;
  double* drag;
};
hipError_t launch_ns_prims_f64(int64_t nc, double gamma, const double* U, void* Q, hipStream_t s);
hipError_t launch_ns_grad_f64(const NsStepArgs& a, hipStream_t s);
hipError_t launch_ns_prims_f32(int64_t nc, double gamma, const double* U, void* Q, hipStream_t s);
hipError_t launch_ns_grad_f32(const NsStepArgs& a, hipStream_t s);
// ---- discrete adjoint of the ns step (gd_adj_fp64.hip / gd_adj_fp32.hip) ----
// values per cell slot in the adjoint's slot geometry: the owner's normal
// (2), the length, the other cell's offset to the face midpoint (2)
constexpr int kAdjSlotGeo = 5;

struct AdjStepArgs {
  MeshView m;
  gd_params prm;
  const void* rec;      // [nc][16] forward record of the step, compute type
  void* lamF;           // [nc][4] -dt/A lamU^{n+1}, compute type
  void* selfq;          // [nc][4] primitive adjoint, own part
  void* nrev;           // [nc][nf][4] primitive adjoint sent to each neighbour
  const void* sgeo;     // [kAdjSlotGeo][nf][nc] geometry of each cell slot's face (see launch_adj_slot_geo)
  double* bcsum;        // [n_wall][4]
  const double* lamU_in;
  double* lamU_out;
  double dt;
  int32_t wall_code;
  double fw[2];
};
// per cell slot, the face-list geometry of the slot's face as the flux pass
// reads it, component-major [kAdjSlotGeo][nf][nc]
hipError_t launch_adj_slot_geo_f64(const MeshView& m, void* sgeo, hipStream_t s);
hipError_t launch_adj_slot_geo_f32(const MeshView& m, void* sgeo, hipStream_t s);
hipError_t launch_adj_cell_f64(const AdjStepArgs& a, hipStream_t s);
hipError_t launch_adj_state_f64(const AdjStepArgs& a, bool gather, hipStream_t s);
hipError_t launch_adj_cell_f32(const AdjStepArgs& a, hipStream_t s);
hipError_t launch_adj_state_f32(const AdjStepArgs& a, bool gather, hipStream_t s);

hipError_t launch_ns_face_f64(const NsStepArgs& a, hipStream_t s);
hipError_t launch_ns_update_f64(const NsStepArgs& a, hipStream_t s);
hipError_t launch_ns_face_f32(const NsStepArgs& a, hipStream_t s);
hipError_t launch_ns_update_f32(const NsStepArgs& a, hipStream_t s);

// the reference-scheme adjoint (gd_ref_adj.hip), one reversed loop step from
// the adjoint lamU of U_{s+1} and the primitives Q / node primitives Qn of
// step s (records [n][4] for the node-level adjoints, [n][6] for gradients):
//   face:  per face-list entry, the adjoint of its flux pushed through AUSM+
//          and the limiters to its two cells and two far nodes -> lamE [n_rfaces][16]
//   nodes: per node, those far-node parts -> lamQn; with the viscous term the
//          second Green-Gauss level transposed -> lamG6N
//   g6:    per cell, the node means of gradients and the first level's 1/A
//          transposed -> tG
//   phi:   per node, the first Green-Gauss level transposed -> lamPhiN
//   cell:  per cell, the entries' cell parts, the node means and cluster means
//          transposed, the primitives' Jacobian at U_s and the identity -> lamU of U_s
hipError_t launch_radj_face(const MeshView& m, const gd_params& p, const double* Q, const double* Qn,
                            const double* lam, double bdt, double t, int32_t wall_code, const double fw[2],
                            double* lqo, double* lqn, double* ldn, double* lup, double* lbc, hipStream_t s);
hipError_t launch_radj_nodes(const MeshView& m, const RefAdjMaps& a, const double* ldn, const double* lup,
                             const double* lam, double bdt, const double mu[3], bool visc, double* lamQn,
                             double* lamG6N, hipStream_t s);
hipError_t launch_radj_g6(const MeshView& m, const RefAdjMaps& a, const double* lamG6N, double* tG, hipStream_t s);
hipError_t launch_radj_phi(const MeshView& m, const RefAdjMaps& a, const double* tG, double* lamPhiN, hipStream_t s);
hipError_t launch_radj_cell(const MeshView& m, const RefAdjMaps& a, double gamma, const double* lqo,
                            const double* lqn, const double* lamQn, const double* lamPhiN, const double* Q,
                            const double* lam, double* lamOut, hipStream_t s);
// [4][nc] <-> [nc][4]
hipError_t launch_radj_to_rec(int64_t nc, const double* soa, double* rec, hipStream_t s);
hipError_t launch_radj_from_rec(int64_t nc, const double* rec, double* soa, hipStream_t s);

}  // namespace gd
