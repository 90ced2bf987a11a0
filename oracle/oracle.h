/*
 * CPU oracle for the explicit Godunov step: TEST INFRASTRUCTURE ONLY.
 *
 * A scalar C++ restatement of the reference's Julia step, used by the tests,
 * by the smoke check and by the bench's CPU-baseline leg as the checker.  The
 * product (libgodunov, HIP) never links or calls it.
 *
 * All arrays use the reference layout: Julia column-major (nCells, k) arrays
 * (contiguous column vectors), 1-based ids, boundary codes <= 0 in
 * cell_stiffness.  Every function cites the reference lines it restates.
 *
 * Pinning (tests/test_oracle.py, test_mesh.py): every function is checked
 * against hand-computed values, closed forms, invariants or an independent
 * library computation.  orc_calc_one_stage_muscl_cpu_fast /
 * orc_calc_one_stage_muscl are pinned value by value by the closed form on a
 * linear field (face states and flux sums), by the order of accuracy of their
 * residual on a smooth field, and by whole-step properties (free stream,
 * conservation, numbering independence, mirror symmetry).  Parity unpinned,
 * since the Julia reference cannot run here: the rounding of Julia's x^4 and
 * pow in orc_air_conductivity, and of its @fastmath node interpolation
 * (utilsFVM2dp.jl:305), which may contract or reorder; orc_calc_one_stage_cuda
 * follows the reference's GPU kernels (AUSMflux2dCUDA*.jl:56-98) at the level
 * of their source expressions and Julia's type promotion, not of the
 * CUDA.jl/LLVM code generation, which may contract multiply-adds.
 */
#ifndef GODUNOV_ORACLE_H
#define GODUNOV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int64_t nCells, nNodes;
  int32_t nNeibCells;
  const double *xNodes, *yNodes;
  const int32_t *mesh_connectivity;       /* nCells x 7 */
  const double *cell_nodes_X, *cell_nodes_Y; /* nCells x 4 */
  const double *cell_mid_points;          /* nCells x 2 */
  const double *cell_areas, *HX;          /* nCells */
  const double *cell_edges_Nx, *cell_edges_Ny, *cell_edges_length; /* nCells x 4 */
  const int32_t *cell_stiffness;          /* nCells x 4 */
  const int32_t *cell_clusters;           /* nNodes x 8 */
  const double *node_stencils;            /* nNodes x 8 */
  const int32_t *node2cellsL2up, *node2cellsL2down, *cells2nodes; /* nCells x 8 */
} orc_mesh;

typedef struct {
  double *densityCells, *UxCells, *UyCells, *pressureCells, *aSoundCells, *VMAXCells;
  double *densityNodes, *UxNodes, *UyNodes, *pressureNodes;
} orc_fields;

typedef struct { double RGAS, Gamma, Cp, mu, kGas; } orc_thermo;

enum {
  ORC_BC_NONE = 0,
  ORC_BC_FIXED = 1,          /* boundaryConditions2d.jl:34-39 (data = rho,u,v,p) */
  ORC_BC_EXTRAPOLATE = 2,    /* boundaryConditions2d.jl:47-52 */
  ORC_BC_SLIP_WALL = 3,      /* bcInviscidWall.jl:21-38 */
  ORC_BC_TGV_MIRROR = 4,     /* boundaryConditions_TGV2d.jl:16-30 (data = P,U,L,mu) */
  ORC_BC_ISOTHERMAL_WALL = 5 /* no-slip wall at temperature data[0] */
};
typedef struct { int32_t kind; int32_t pad; double data[4]; } orc_bc;

enum { ORC_FLUX_AUSM_PLUS = 0, ORC_FLUX_ROE = 1, ORC_FLUX_AUSM_PLUS_M2020 = 2 };
enum { ORC_LIM_REF_HARMONIC = 0, ORC_LIM_VAN_LEER_A = 1, ORC_LIM_VAN_ALBADA_A = 2, ORC_LIM_MINMOD = 3,
       ORC_LIM_VENKATAKRISHNAN = 4 };
enum { ORC_ENERGY_AIR_POLY = 0, ORC_ENERGY_CONST = 1 };

#define ORC_MAX_BC 16
typedef struct {
  int32_t flux;          /* ORC_FLUX_* */
  int32_t limiter;       /* ORC_LIM_* (ref mode: 0..3) */
  double limiter_eps;    /* 1e-12 on the step path, computeslope2d.jl:116 */
  int32_t include_udiff; /* 1: calcOneStageMUSCLcpuFast; 0: calcOneStageMUSCL / CUDA */
  int32_t energy_law;    /* ORC_ENERGY_* for calcDiffTerm's energy row */
  double mu_const;       /* 1e-2 in calcDiffTerm.jl:110 */
  double k_const;        /* conductivity factor for ORC_ENERGY_CONST */
  orc_bc bc[ORC_MAX_BC + 1]; /* indexed by -code */
} orc_params;

/* ---- pointwise kernels ---- */
double orc_limiter(int32_t kind, double x, double y, double c);
void orc_ausm_plus_flux(const double* q8, double nx, double ny, double side, double gamma, double* flux4);
void orc_ausm_plus_m2020_flux(const double* q8, double nx, double ny, double side, double gamma, double* flux4);
void orc_roe_flux(const double* q8, double nx, double ny, double side, double gamma, double* flux4);
void orc_ns_ausm_flux(const double* q8, double nx, double ny, double gamma, double* flux4);
void orc_ns_roe_flux(const double* q8, double nx, double ny, double gamma, double* flux4);
double orc_air_conductivity(double T);
double orc_sutherland_viscosity(double T);

/* ---- reference operators (ref mode) ---- */
void orc_ref_bc_ghost(const orc_mesh* m, const orc_thermo* th, const orc_params* prm, int64_t i, int32_t k,
                      const double* cur, double t, double* out4);
void orc_distribute_cells_in_threads(int32_t nThreads, int64_t nCells, int32_t* table /* nThreads x 2 */);
void orc_phs2dcns(const orc_fields* f, int64_t nCells, double gamma, double* U);
int orc_cells2nodes(const int32_t* nodesThreads, int32_t nThreads, const orc_mesh* m, const orc_fields* f,
                    const double* Ucells, double* Unodes /* may be NULL */);
int orc_calc_one_stage_muscl_cpu_fast(const int32_t* cellsThreads, int32_t nThreads, double betta, double dt,
                                      double flowTime, const orc_mesh* m, const orc_fields* f, const orc_thermo* th,
                                      const orc_params* prm, const double* UOld, double* FLUXES, const double* Udiff,
                                      double* UNew);
int orc_calc_one_stage_muscl(const int32_t* cellsThreads, int32_t nThreads, double betta, double dt, double flowTime,
                             const orc_mesh* m, const orc_fields* f, const orc_thermo* th, const orc_params* prm,
                             const double* UOld, double* UNew, double* FLUXES,
                             double* ULeft /* optional: ULeft1..4, four (4, nCells) arrays */,
                             double* URight /* optional: URight1..4 */);
int orc_calc_diff_term(const int32_t* cellsThreads, int32_t nThreads, const orc_mesh* m, const orc_fields* f,
                       const orc_thermo* th, const orc_params* prm, const double* UC, double* Udiff,
                       double* lap3 /* optional nCells x 3: laplasUCu, laplasUCv, laplasUCe */);
int orc_green_gauss_gradient(const int32_t* cellsThreads, int32_t nThreads, const orc_mesh* m, const double* phi,
                             double* gradX, double* gradY);
int orc_update_variables(const int32_t* cellsThreads, int32_t nThreads, int64_t nCells, double gamma,
                         const double* UNew, double* UOld, double* Delta, const orc_fields* f);
double orc_ref_tau(const orc_mesh* m, const orc_fields* f, double cfl);
/* calcOneStageCUDA (x32 = 0) / calcOneStageCUDAx32 (x32 = 1): AUSM+ of the
   caller's face states on face slots 1..3 with device-side accumulation in
   Float64 / Float32, then the Float64 update; iFlux: (nCells, 4) or NULL */
int orc_calc_one_stage_cuda(int32_t x32, const orc_mesh* m, const orc_thermo* th, double betta, double dt,
                            const double* UOld, const double* ULeft, const double* URight, double* UNew,
                            double* iFlux);
double orc_cuda_split(int32_t x32, int32_t which, double M);
double orc_local_dt(const orc_mesh* m, const orc_fields* f, double cfl);

/* ---- Navier-Stokes (ns) scheme: cell gradients, Venkatakrishnan, Sutherland ---- */
typedef struct {
  double gamma, R, Pr;
  double mu_ref, T_ref, S;  /* Sutherland: mu = mu_ref (T/T_ref)^1.5 (T_ref+S)/(T+S) */
  double venkat_K;
  double venkat_scale[4];   /* per-variable scales s_v in eps^2 = (K h)^3 s_v^2 */
  int32_t flux;             /* ORC_FLUX_AUSM_PLUS or ORC_FLUX_ROE */
  int32_t viscous;          /* 0 inviscid, 1 Sutherland stresses + heat flux */
  int32_t mixed;            /* 1: slopes/limiter/reconstruction/fluxes in float */
  int32_t pad;
  orc_bc bc[ORC_MAX_BC + 1];
} orc_ns_params;

int orc_ns_step(const orc_mesh* m, const orc_ns_params* prm, double dt, const double* UOld, double* UNew,
                double* grad8 /* optional nCells x 8 */, double* phi4 /* optional nCells x 4 */,
                double* wall_force2 /* optional: pressure+viscous force on code-wall faces */);
double orc_ns_dt(const orc_mesh* m, const orc_ns_params* prm, const double* U, double cfl);
/* tangent-linear FP64 ns step (forward-mode dual numbers); dbc: [ORC_MAX_BC+1][4] or NULL */
int orc_ns_step_tl(const orc_mesh* m, const orc_ns_params* prm, double dt, const double* UOld, const double* dUOld,
                   const double* dbc, double* UNew, double* dUNew, double* force2, double* dforce2);
void orc_set_threads(int32_t n); /* OpenMP threads of the ns oracle (default 1) */

#ifdef __cplusplus
}
#endif
#endif
