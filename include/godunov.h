/*
 * libgodunov: explicit Godunov step of the 2-D compressible Navier-Stokes
 * finite-volume solver, as hand-written HIP kernels for MI355X (gfx950).
 *
 * Plain C ABI: opaque handles, raw pointers and sizes, int status (0 = ok,
 * < 0 = error, message via gd_last_error).  Host arrays use the reference's
 * layout: Julia column-major (nCells, k) arrays, i.e. k contiguous column
 * vectors, and 1-based ids, so a Julia caller passes its arrays as they are.
 * The library keeps its own 0-based structure-of-arrays copy in HBM, in a cell
 * order of its choosing (as given, RCM, a Hilbert curve, or a caller
 * permutation); every host transfer is in the caller's numbering.
 *
 * Reference interfaces replaced (dimaZloy/2D-DNS-godunov-solver-):
 *   mesh2d_Int32 / readMesh2dHDF5   primeObjects.jl:109-135, utilsIO.jl:107-316 -> gd_mesh_create
 *   fields2d                        primeObjects.jl:85-106                      -> gd_state_*_fields
 *   calcOneStageMUSCLcpuFast        calcOneStageMUSCLcpuFast.jl:3-107           -> gd_calc_one_stage(GD_STAGE_MUSCL_FAST)
 *   SecondOrderUpwindM2             SOUscheme.jl:3-118                          -> same kernel
 *   calcOneStageMUSCL               calcOneStageMUSCL.jl:4-106                  -> gd_calc_one_stage(GD_STAGE_MUSCL)
 *   calcOneStageCUDA / ...x32       calcOneStageCUDA.jl:4-135, calcOneStageCUDAx32.jl:5-155
 *                                                                               -> gd_calc_one_stage_cuda
 *   calcDiffTerm                    calcDiffTerm.jl:31-121                      -> gd_calc_diff_term
 *   updateVariablesSA               evaluate2d.jl:52-110                        -> gd_update_variables
 *   cells2nodesSolutionReconstructionWithStencils utilsFVM2dp.jl:299-365       -> gd_cells2nodes
 *   the time loop body              testViscousFluxesTGV2d.jl:414-481           -> gd_run
 *
 * Threading: a gd_state owns one HIP stream and is not re-entrant; distinct
 * states are independent.  Every call returns after its work is complete
 * unless noted (gd_run_async).
 */
#ifndef GODUNOV_H
#define GODUNOV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define GD_OK 0
#define GD_ERR_INVALID (-1)      /* bad argument, inconsistent mesh */
#define GD_ERR_HIP (-2)          /* HIP runtime error */
#define GD_ERR_NOMEM (-3)        /* device allocation failed */
#define GD_ERR_UNSUPPORTED (-4)  /* option combination not implemented */
#define GD_ERR_NONPHYSICAL (-5)  /* gd_run found non-positive density/pressure or NaN */
#define GD_ERR_NO_DEVICE (-6)    /* no HIP device */

typedef struct gd_mesh gd_mesh;
typedef struct gd_state gd_state;

/* mesh2d_Int32 (primeObjects.jl:109-135): pointers to the Julia arrays as-is. */
typedef struct {
  int64_t n_cells, n_nodes;
  int32_t n_neib_cells;                  /* columns of cell_clusters / node_stencils (8) */
  int32_t reserved;
  const double *xNodes, *yNodes;         /* n_nodes */
  const int32_t *mesh_connectivity;      /* n_cells x 7: id, type(3 tri, 2 quad), nVertices, v1..v4 */
  const double *cell_nodes_X, *cell_nodes_Y; /* n_cells x 4, vertex coordinates seen from the cell */
  const double *cell_mid_points;         /* n_cells x 2 */
  const double *cell_areas, *HX;         /* n_cells */
  const double *cell_edges_Nx, *cell_edges_Ny, *cell_edges_length; /* n_cells x 4, inward unit normals */
  const int32_t *cell_stiffness;         /* n_cells x 4: neighbour id (>0) or boundary code (<=0) */
  const int32_t *cell_clusters;          /* n_nodes x n_neib_cells, 0 padded */
  const double *node_stencils;           /* n_nodes x n_neib_cells */
  const int32_t *node2cellsL2up, *node2cellsL2down, *cells2nodes; /* n_cells x 8 */
} gd_mesh_desc;

/* device cell order */
#define GD_ORDER_AS_GIVEN 0
#define GD_ORDER_RCM 1            /* reverse Cuthill-McKee of the face graph */
#define GD_ORDER_PERMUTATION 2    /* caller's new->old map (0-based) */
#define GD_ORDER_HILBERT 3        /* Hilbert curve through the cell centroids */

typedef struct {
  int64_t n_cells, n_nodes, n_boundary_faces;
  int32_t has_quads, max_cluster;
  int64_t device_bytes;           /* HBM held by the mesh */
  int64_t bandwidth;              /* max |i - j| over neighbouring cells in device order */
} gd_mesh_info;

int gd_mesh_create(const gd_mesh_desc* desc, int32_t order, const int64_t* perm_new_to_old, gd_mesh** out);
int gd_mesh_info_get(const gd_mesh* mesh, gd_mesh_info* info);
/* The device cell order, for callers that work on the device arrays
   (gd_state_device_u, gd_state_device_adjoint): new_to_old[d] = the caller's
   0-based id of device cell d (n_cells entries).  No counterpart in the
   reference, which has one numbering. */
int gd_mesh_cell_order(const gd_mesh* mesh, int64_t* new_to_old);
int gd_mesh_destroy(gd_mesh* mesh);

/* ---- numerics ---- */
#define GD_SCHEME_REF 0   /* the reference scheme: vertex MUSCL (computeslope2d.jl) + calcDiffTerm Laplacian */
#define GD_SCHEME_NS 1    /* cell gradients (LSQ) + Venkatakrishnan + Sutherland stress/heat fluxes */

#define GD_FP64 0
#define GD_FP32_MIXED 1   /* slopes/limiters/reconstruction/fluxes in FP32; geometry, accumulation, update, p/T in FP64 */

#define GD_FLUX_AUSM_PLUS 0        /* AUSMplusFlux2dFast, AUSMflux2dFast.jl:87-175 */
#define GD_FLUX_ROE 1              /* RoeFlux2dFast, RoeFlux2dFast.jl:1-198 */
#define GD_FLUX_AUSM_PLUS_M2020 2  /* AUSMplusM2020Flux2dFast, AUSMflux2dFast.jl:234-284 (ref scheme only) */

#define GD_LIM_REF_HARMONIC 0      /* Minmod_Limiter, limiters.jl:33-35 */
#define GD_LIM_VAN_LEER_A 1        /* vanLeer_LimiterA, limiters.jl:41-43 */
#define GD_LIM_VAN_ALBADA_A 2      /* vanAlbada_LimiterA, limiters.jl:54-56 */
#define GD_LIM_MINMOD 3            /* 0.5 (sign x + sign y) min(|x|,|y|), limiters.jl:45-47 */
#define GD_LIM_VENKATAKRISHNAN 4   /* ns scheme */

#define GD_ENERGY_AIR_POLY 0       /* calcAirConductivityFromT, thermo.jl:57-60 */
#define GD_ENERGY_CONST 1          /* k(T) = k_const */

#define GD_BC_NONE 0
#define GD_BC_FIXED 1              /* data = rho, u, v, p          (boundaryConditions2d.jl:34-39) */
#define GD_BC_EXTRAPOLATE 2        /* zero-gradient copy           (boundaryConditions2d.jl:47-52) */
#define GD_BC_SLIP_WALL 3          /* velocity reflection          (bcInviscidWall.jl:109-126) */
#define GD_BC_TGV_MIRROR 4         /* data = P, U, L, mu: analytic TGV at the mirrored centroid
                                      (boundaryConditions_TGV2d.jl:16-30) */
#define GD_BC_ISOTHERMAL_WALL 5    /* data[0] = wall temperature; no-slip */
#define GD_MAX_BC 16

typedef struct { int32_t kind; int32_t reserved; double data[4]; } gd_bc;

typedef struct {
  int32_t scheme, precision, flux, limiter;
  double limiter_eps;      /* 1e-12 (computeslope2d.jl:116) */
  double venkat_K;         /* ns: eps^2 = (K h)^3 s_v^2, h = 2A/perimeter */
  double venkat_scale[4];  /* ns: s_v for rho, u, v, p */
  int32_t viscous;         /* ref: calcDiffTerm Laplacian on/off; ns: Sutherland stresses on/off */
  int32_t energy_law;      /* ref: GD_ENERGY_* */
  double mu_const;         /* ref: 1e-2 (calcDiffTerm.jl:110) */
  double k_const;          /* ref: k for GD_ENERGY_CONST */
  /* THERMOPHYSICS (thermo.jl:4-10) */
  double RGAS, Gamma;
  /* ns: mu(T) = mu_ref (T/T_ref)^1.5 (T_ref + S)/(T + S); Prandtl number */
  double mu_ref, T_ref, S_suth, Pr;
  gd_bc bc[GD_MAX_BC + 1]; /* indexed by -code (code 0 -> bc[0]) */
} gd_params;

void gd_params_default(gd_params* p);

int gd_state_create(gd_mesh* mesh, const gd_params* params, gd_state** out);
int gd_state_destroy(gd_state* st);
int gd_state_set_params(gd_state* st, const gd_params* params);

/* fields2d: any pointer may be NULL (left unchanged / not read back) */
int gd_state_set_fields(gd_state* st, const double* densityCells, const double* UxCells, const double* UyCells,
                        const double* pressureCells, const double* densityNodes, const double* UxNodes,
                        const double* UyNodes, const double* pressureNodes);
int gd_state_get_fields(gd_state* st, double* densityCells, double* UxCells, double* UyCells, double* pressureCells,
                        double* aSoundCells, double* VMAXCells, double* densityNodes, double* UxNodes,
                        double* UyNodes, double* pressureNodes);

/* (n_cells, 4) column-major conservative arrays of the reference time loop */
#define GD_ARR_U_OLD 0    /* UconsCellsOldX */
#define GD_ARR_U_NEW 1    /* UconsCellsNewX */
#define GD_ARR_U_DIFF 2   /* UConsDiffCellsX */
#define GD_ARR_FLUXES 3   /* iFLUXX / FLUXES */
#define GD_ARR_DELTA 4    /* DeltaX */
#define GD_ARR_U_NODES 5  /* UconsNodesOldX, (n_nodes, 4) */
int gd_state_set_array(gd_state* st, int32_t which, const double* host);
int gd_state_get_array(gd_state* st, int32_t which, double* host);

/* ---- reference operators (GD_SCHEME_REF, FP64) ---- */
#define GD_STAGE_MUSCL_FAST 0  /* calcOneStageMUSCLcpuFast / SecondOrderUpwindM2: adds betta*dt*Udiff */
#define GD_STAGE_MUSCL 1       /* calcOneStageMUSCL / calcOneStageCUDA: no Udiff */
int gd_calc_one_stage(gd_state* st, int32_t variant, double betta, double dt, double flow_time);
/* phase one of calcOneStageMUSCL (calcOneStageMUSCL.jl:4-106 with
   computeMUSCLStencilsCUDA, computeMUSCLStencilsCUDA.jl:3-28): the face
   states ULeft1..4 / URight1..4 from the state's cell and node primitives at
   flow_time; ULeft, URight: four (4, n_cells) column-major arrays each, slot
   after slot, caller numbering; slots past a cell's faces are left as given */
int gd_calc_face_states(gd_state* st, double flow_time, double* ULeft, double* URight);
/* calcOneStageCUDA (precision GD_FP64, calcOneStageCUDA.jl:4-135) and
   calcOneStageCUDAx32 (GD_FP32_MIXED, calcOneStageCUDAx32.jl:5-155): AUSM+ as
   the reference's GPU kernel writes it (AUSMflux2dCUDA*.jl:56-98) on face
   slots 1..3 of the face states ULeft1..3 / URight1..3 (each (4, n_cells)
   column-major, slot after slot as for gd_calc_face_states; slot 4 is not
   read: the source leaves the 4th face's launch commented out), the per-cell
   sums accumulated in Float64 or Float32, then U_NEW = U_OLD - iFlux betta dt
   / A in Float64 without Udiff.  ULeft = URight = NULL: the face states of
   the state's primitives at flow_time (phase one on the device).  iFlux
   (n_cells, 4) or NULL: the accumulated sums (FLUXES holds them too). */
int gd_calc_one_stage_cuda(gd_state* st, int32_t precision, double betta, double dt, double flow_time,
                           const double* ULeft, const double* URight, double* iFlux);
int gd_calc_diff_term(gd_state* st);
/* calcDiffTerm on the caller's device arrays, allocating nothing.  No
   counterpart in the reference: for callers that cannot allocate device
   memory (work recorded into HIP graphs, tight time loops).  d_U (conservative
   state) and d_Q (cell primitives rho, u, v, p) are 4 x n_cells SoA device
   arrays in the mesh's device order (gd_mesh_cell_order); columns 1..3 of
   d_Udiff (same layout) are written, column 0 is left as it is.  Enqueued on
   `stream` (a hipStream_t, NULL for the null stream) without waiting; the
   current device must be the mesh's.  The intermediates live in one static
   device array of 256 MiB per device (held until the process ends), so a mesh
   may need at most that: gd_diff_term_scratch_bytes(n_cells, n_nodes) <= 256 MiB
   (about 3M triangles), else GD_ERR_INVALID.  Calls that overlap on one
   device share that array and race. */
size_t gd_diff_term_scratch_bytes(int64_t n_cells, int64_t n_nodes);
int gd_calc_diff_term_noalloc(const gd_mesh* mesh, const gd_params* params, const double* d_U, const double* d_Q,
                              double* d_Udiff, void* stream);
int gd_update_variables(gd_state* st);
int gd_cells2nodes(gd_state* st);

/* ---- device-resident time loop ---- */
#define GD_DT_FIXED 0     /* dt as given */
#define GD_DT_REF_TAU 1   /* CFL * max(HX) / max(VMAX) (testViscousFluxesTGV2d.jl:420-423) */
#define GD_DT_LOCAL 2     /* CFL * min_i h_i / (|u_i| + a_i [+ viscous]) with h = 2A/perimeter */
typedef struct {
  int32_t dt_policy;
  int32_t time_kernels;   /* record HIP events around every launch and report per-kernel times */
  double dt, cfl, betta;
  double flow_time;       /* start time; advanced by the loop */
  int32_t residuals;      /* 1: accumulate sum(Delta^2) per variable over the run's last step */
  int32_t wall_code;      /* boundary code whose faces integrate the wall force (0: none) */
} gd_run_params;

#define GD_MAX_KERNELS 8
typedef struct {
  double flow_time;           /* after the run */
  double last_dt;
  double residuals[4];        /* sum over cells of Delta^2, last step */
  double wall_force[2];       /* last step: integral over wall faces of p n - tau.n (x, y) */
  double wall_force_sum[2];   /* sum over the run's steps */
  int32_t n_kernels;
  int32_t nonphysical;        /* count of cells with rho<=0, p<=0 or NaN after the last step */
  char kernel_names[GD_MAX_KERNELS][32];
  double kernel_ms_total[GD_MAX_KERNELS];   /* HIP-event time summed over launches */
  int64_t kernel_launches[GD_MAX_KERNELS];
  double kernel_bytes[GD_MAX_KERNELS];       /* algorithmic bytes per launch */
} gd_run_stats;

int gd_run(gd_state* st, int32_t n_steps, gd_run_params* rp, gd_run_stats* stats);

/* ---- discrete adjoint of the ns loop (AUSM+ or Roe, FP64 or FP32-mixed) and
   of the reference loop (FP64) ----
   Not in the reference, which has no adjoint; the objective is
   J = sum_{n<N} w . F_n + <dJ_dUN, U_N>, F_n the wall force of step n on the
   faces of code wall_code, at a fixed dt.  With the reference scheme the step
   is gd_run's reference loop at betta 1 (calcDiffTerm, the fast MUSCL stage,
   updateVariablesSA, the node interpolation), started from the primitives of
   the state's U and their node means, F_n the AUSM+ momentum flux through the
   faces of wall_code; it covers AUSM+ with the harmonic, van Albada or minmod
   limiter, constant mu and k (or inviscid), triangles and quads, and every
   boundary code; anything else returns GD_ERR_UNSUPPORTED.  Runs N steps forward from the
   state's U (which ends at U_N, as after gd_run), keeping U every
   checkpoint_every steps and recomputing the steps between checkpoints on the
   way back (unless record_budget holds their records).  dJ_dU0 (NULL: kept
   on the device only, see gd_state_device_adjoint) / dJ_dUN: (n_cells, 4) column-major, caller numbering.
   dJ_dbc: [(GD_MAX_BC + 1)][4], derivative with respect to the fixed states
   bc[-code].data[0..3] (rho, u, v, p) of GD_BC_FIXED codes, zero elsewhere.
   The device workspace (checkpoints, kept records, adjoint fields) stays with
   the state for the next call, grown when a call needs more, and is freed by
   gd_state_destroy. */
typedef struct {
  int32_t n_steps;
  int32_t checkpoint_every;
  double dt;
  double flow_time;
  int32_t wall_code;
  int32_t reserved;
  double force_weight[2];
  int64_t record_budget;     /* bytes of HBM for cell records: the records of as many trailing
                                segments as fit are kept from the forward sweep instead of
                                being recomputed (0: one segment) */
} gd_adjoint_params;

typedef struct {
  double J;
  double force_sum[2];       /* sum over the N steps of the wall force */
  double flow_time;
  int32_t recomputed_steps;
  int32_t n_checkpoints;
  int64_t storage_bytes;     /* checkpoints + the cell records of the kept segments */
  double ms_forward;         /* forward sweep, HIP events */
  double ms_reverse;         /* recomputation + transposed steps */
} gd_adjoint_stats;

int gd_adjoint_run(gd_state* st, const gd_adjoint_params* ap, const double* dJ_dUN, double* dJ_dU0, double* dJ_dbc,
                   gd_adjoint_stats* stats);
/* dJ/dU0 of the last gd_adjoint_run, on the device: 4 x n_cells SoA in device
   order; valid until the next gd_adjoint_run on the state (dJ_dU0 may be NULL
   for callers that keep the gradient on the device) */
int gd_state_device_adjoint(gd_state* st, double** dev_grad_soa);
int gd_sync(gd_state* st);

/* device pointers for zero-copy callers; valid until the next gd_run/step.
   The caller may read and write U through it: once it has been handed out,
   every later gd_run on the state starts from U as it then is. */
int gd_state_device_u(gd_state* st, double** dev_u_soa /* 4 x n_cells, device order */);

/* Measurement only, no counterpart in the reference: the current device's
   HBM streaming rates, for reading the step's bandwidth against what a kernel
   that only streams reaches.  A plain copy of `bytes` (a multiple of 16, at
   least 16 KiB) between two buffers this call allocates and frees, and a plain
   read of one; best of cached and non-temporal 16 B vector accesses, `iters`
   timed launches each after one warm launch.  copy_gbs counts bytes read plus
   bytes written. */
int gd_hbm_stream_rates(int64_t bytes, int32_t iters, double* copy_gbs, double* read_gbs);

int gd_device_count(int32_t* n);
int gd_last_error(char* buf, size_t n);
const char* gd_version(void);

#ifdef __cplusplus
}
#endif
#endif
