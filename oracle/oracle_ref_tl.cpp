// Tangent-linear restatement of the reference scheme's time loop: TEST
// INFRASTRUCTURE ONLY (the checker of the reference-scheme adjoint).
//
// One loop step is what the reference's viscous TGV driver does per time step
// (testViscousFluxesTGV2d.jl:429-481) and what the oracle's pointwise
// functions in oracle.cpp restate:
//   calcDiffTerm (calcDiffTerm.jl:31-121, constant mu and k),
//   calcOneStageMUSCLcpuFast (calcOneStageMUSCLcpuFast.jl:3-107 with
//     computeInterfaceSlope, computeslope2d.jl:3-136, and AUSMplusFlux2dFast,
//     AUSMflux2dFast.jl:87-175),
//   updateVariablesSA (evaluate2d.jl:83-93) and the node interpolation
//   (utilsFVM2dp.jl:299-365),
// at a fixed dt, the flow time advancing by dt per step (the ghost states of
// the TGV mirror depend on it).  The state is U alone: the loop starts from the
// primitives of U0 and their node means.  Every function is templated on the
// scalar: with double it is the oracle's loop operation for operation; with
// the dual number of dual.hpp it carries one directional derivative dU through
// the same arithmetic.  Triangles and quads, interior faces and the boundary
// ghosts of oracle.cpp's bc_ghost (fixed state, zero gradient, slip wall, TGV
// mirror, isothermal wall); the summed wall force of the faces of one code
// (each step's AUSM+ momentum flux through them) is carried along.
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "dual.hpp"
#include "oracle.h"

namespace {

using orc_ad::Dual;

inline int32_t Ix(const int32_t* a, int64_t n, int64_t i, int k) { return a[(int64_t)(k - 1) * n + (i - 1)]; }
inline double Dx(const double* a, int64_t n, int64_t i, int k) { return a[(int64_t)(k - 1) * n + (i - 1)]; }

template <class T> inline T tabs(T x) { using std::fabs; using orc_ad::fabs; return fabs(x); }
template <class T> inline T tsqrt(T x) { using std::sqrt; using orc_ad::sqrt; return sqrt(x); }
template <class T> inline T tfma(T a, T b, T c) { using std::fma; using orc_ad::fma; return fma(a, b, c); }
// Julia's sign (derivative 0) and NaN-propagating min, as in oracle.cpp
template <class T> inline T jsignT(T x) { return T(orc_ad::val(x) > 0.0 ? 1.0 : (orc_ad::val(x) < 0.0 ? -1.0 : orc_ad::val(x))); }
template <class T> inline T jminT(T x, T y) {
  const double a = orc_ad::val(x), b = orc_ad::val(y);
  return (a != a || b != b) ? x + y : (b < a ? y : x);
}

// limiters.jl:33-56 (oracle.cpp: limiter)
template <class T>
T limiterT(int kind, T x, T y, double c) {
  switch (kind) {
    case ORC_LIM_VAN_LEER_A: return (x * y + tabs(x * y)) / (x + y + T(c));
    case ORC_LIM_VAN_ALBADA_A: return (x * (y * y + T(c)) + y * (x * x + T(c))) / (x * x + y * y + T(2.0 * c));
    case ORC_LIM_MINMOD: return T(0.5) * (jsignT(x) + jsignT(y)) * jminT(tabs(x), tabs(y));
    default: return (jsignT(x) + jsignT(y)) * tabs(x * y) / (tabs(x) + tabs(y) + T(c));
  }
}

// AUSMflux2dFast.jl:23-53 (oracle.cpp: P_m .. M_m)
template <class T> T Palfa_m(T M, double A) { return T(0.25) * (M - T(1.0)) * (M - T(1.0)) * (T(2.0) + M) - T(A) * M * (M * M - T(1.0)) * (M * M - T(1.0)); }
template <class T> T Palfa_p(T M, double A) { return T(0.25) * (M + T(1.0)) * (M + T(1.0)) * (T(2.0) - M) + T(A) * M * (M * M - T(1.0)) * (M * M - T(1.0)); }
template <class T> T Mbetta_p(T M, double B) { return T(0.25) * (M + T(1.0)) * (M + T(1.0)) + T(B) * (M * M - T(1.0)) * (M * M - T(1.0)); }
template <class T> T Mbetta_m(T M, double B) { return T(-0.25) * (M - T(1.0)) * (M - T(1.0)) - T(B) * (M * M - T(1.0)) * (M * M - T(1.0)); }
template <class T> T P_mT(T M, double A) { return (orc_ad::val(tabs(M)) >= 1.0) ? T(0.5) * (T(1.0) - jsignT(M)) : Palfa_m(M, A); }
template <class T> T P_pT(T M, double A) { return (orc_ad::val(tabs(M)) >= 1.0) ? T(0.5) * (T(1.0) + jsignT(M)) : Palfa_p(M, A); }
template <class T> T M_pT(T M, double B) { return (orc_ad::val(tabs(M)) >= 1.0) ? T(0.5) * (M + tabs(M)) : Mbetta_p(M, B); }
template <class T> T M_mT(T M, double B) { return (orc_ad::val(tabs(M)) >= 1.0) ? T(0.5) * (M - tabs(M)) : Mbetta_m(M, B); }

// AUSMplusFlux2dFast, AUSMflux2dFast.jl:87-175 (oracle.cpp: ausm_plus)
template <class T>
void ausm_plusT(T rhoL, T UL, T VL, T PL, T rhoR, T UR, T VR, T PR, double nx, double ny, double side, double gamma,
                T* flux) {
  const T g1(gamma - 1.0), nxT(nx), nyT(ny), sideT(side);
  const T htL = tfma(T(0.5), (UL * UL + VL * VL), PL / rhoL + PL / rhoL / g1);
  const T htR = tfma(T(0.5), (UR * UR + VR * VR), PR / rhoR + PR / rhoR / g1);
  const T cst(2.0 * (gamma - 1.0) / (gamma + 1.0));
  const T aL_tilda = tsqrt(cst * htL) * jminT(T(1.0), (tsqrt(cst * htL)) / tabs(UL * nxT + VL * nyT));
  const T aR_tilda = tsqrt(cst * htR) * jminT(T(1.0), (tsqrt(cst * htR)) / tabs(UR * nxT + VR * nyT));
  const T a12 = jminT(aL_tilda, aR_tilda);
  const T m = M_pT((UL * nxT + VL * nyT) / a12, 1.0 / 8.0) + M_mT((UR * nxT + VR * nyT) / a12, 1.0 / 8.0);
  const T p12 = P_pT((UL * nxT + VL * nyT) / a12, 3.0 / 16.0) * PL + P_mT((UR * nxT + VR * nyT) / a12, 3.0 / 16.0) * PR;
  const T mp = T(0.5) * (m + tabs(m)), mm = T(0.5) * (m - tabs(m));
  flux[0] = -(a12 * (mp * rhoL + mm * rhoR) + T(0.0)) * sideT;
  flux[1] = -(a12 * (mp * rhoL * UL + mm * rhoR * UR) + p12 * nxT) * sideT;
  flux[2] = -(a12 * (mp * rhoL * VL + mm * rhoR * VR) + p12 * nyT) * sideT;
  flux[3] = -(a12 * (mp * rhoL * htL + mm * rhoR * htR) + T(0.0)) * sideT;
}

template <class T>
struct RefLoop {
  const orc_mesh* m;
  const orc_thermo* th;
  const orc_params* prm;
  int viscous;
  double betta, dt;
  int64_t nc, nn;
  std::vector<T> U, Q, Qn, Ud, Un;
  double t = 0.0;        // flow time of the step
  int32_t wall_code = 0; // faces whose force is summed (0: none)
  const double* dbc = nullptr;  // [ORC_MAX_BC + 1][4]: tangent of the fixed states, or NULL
  T force[2] = {T(0.0), T(0.0)};

  // updateVariablesSA's primitives, evaluate2d.jl:87-90 (oracle.cpp: orc_update_variables)
  void prims() {
    const T G1(th->Gamma - 1.0);
#pragma omp parallel for
    for (int64_t i = 0; i < nc; ++i) {
      Q[i] = U[i];
      Q[nc + i] = U[nc + i] / U[i];
      Q[2 * nc + i] = U[2 * nc + i] / U[i];
      Q[3 * nc + i] = G1 * (U[3 * nc + i] - T(0.5) * (U[nc + i] * U[nc + i] + U[2 * nc + i] * U[2 * nc + i]) / U[i]);
    }
  }
  // cells2nodesSolutionReconstructionWithStencils, utilsFVM2dp.jl:299-365
  void node_means() {
#pragma omp parallel for
    for (int64_t J = 1; J <= nn; ++J) {
      T det(0.0), r[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
      for (int j = 1; j <= 8; ++j) {
        const int32_t c = Ix(m->cell_clusters, nn, J, j);
        if (c != 0) {
          const T wi(Dx(m->node_stencils, nn, J, j));
          for (int v = 0; v < 4; ++v) r[v] += Q[v * nc + c - 1] * wi;
          det += wi;
        }
      }
      if (orc_ad::val(det) != 0)
        for (int v = 0; v < 4; ++v) r[v] = r[v] / det;
      for (int v = 0; v < 4; ++v) Qn[v * nn + J - 1] = r[v];
    }
  }
  // calcPhiInCellCluster, calcGradients.jl:3-18
  T phi_in_cluster(int32_t node, const T* field) const {
    T phi(0.0);
    int64_t counter = 0;
    for (int s = 1; s <= m->nNeibCells; ++s) {
      const int32_t cell = Ix(m->cell_clusters, nn, node, s);
      if (cell > 0 && cell <= nc) { phi += field[cell - 1]; counter += 1; }
    }
    return counter != 0 ? phi / T((double)counter) : T(0.0);
  }
  // calcScalarFieldGreenGaussNodesBasedGradient, calcGradients.jl:36-139 (normals swapped)
  void green_gauss(const T* phi, T* gx, T* gy) const {
#pragma omp parallel for
    for (int64_t C = 1; C <= nc; ++C) {
      const int nf = Ix(m->mesh_connectivity, nc, C, 3) == 4 ? 4 : 3;
      T fx[4], fy[4];
      for (int k = 1; k <= nf; ++k) {
        const double side = Dx(m->cell_edges_length, nc, C, k);
        const double nx = Dx(m->cell_edges_Nx, nc, C, k), ny = Dx(m->cell_edges_Ny, nc, C, k);
        const T pl = phi_in_cluster(Ix(m->cells2nodes, nc, C, 2 * k - 1), phi);
        const T pr = phi_in_cluster(Ix(m->cells2nodes, nc, C, 2 * k), phi);
        fx[k - 1] = T(0.5) * (pl + pr) * T(-ny) * T(side);
        fy[k - 1] = T(0.5) * (pl + pr) * T(-nx) * T(side);
      }
      // the 4th slot of a triangle: 0.5 (0 + 0) (-0) 0, as the oracle adds it
      if (nf == 3) {
        fx[3] = T(0.5) * (T(0.0) + T(0.0)) * T(-0.0) * T(0.0);
        fy[3] = fx[3];
      }
      const T A(m->cell_areas[C - 1]);
      gx[C - 1] = (fx[0] + fx[1] + fx[2] + fx[3]) / A;
      gy[C - 1] = (fy[0] + fy[1] + fy[2] + fy[3]) / A;
    }
  }
  // calcDiffTerm, calcDiffTerm.jl:31-121 with constant mu and k
  void diff_term() {
    std::vector<T> g[3][2], dd[3][2], dummy(nc);
    for (int q = 0; q < 3; ++q)
      for (int c = 0; c < 2; ++c) { g[q][c].assign(nc, T(0.0)); dd[q][c].assign(nc, T(0.0)); }
    for (int q = 0; q < 3; ++q) green_gauss(U.data() + (q + 1) * nc, g[q][0].data(), g[q][1].data());
    for (int q = 0; q < 3; ++q) {
      green_gauss(g[q][0].data(), dd[q][0].data(), dummy.data());
      green_gauss(g[q][1].data(), dummy.data(), dd[q][1].data());
    }
    const T mu(prm->mu_const), k(prm->k_const);
    for (int64_t i = 0; i < nc; ++i) {
      Ud[i] = T(0.0);
      Ud[nc + i] = mu * (dd[0][0][i] + dd[0][1][i]);
      Ud[2 * nc + i] = mu * (dd[1][0][i] + dd[1][1][i]);
      Ud[3 * nc + i] = mu * k * (dd[2][0][i] + dd[2][1][i]);
    }
  }
  // the ghost state of boundary face k of cell i (oracle.cpp: bc_ghost):
  // ComputeUPhysFromBoundaries, boundaryConditions2d.jl:29-57, the wall
  // reflection bcInviscidWall.jl:109-126, the TGV mirror
  // boundaryConditions_TGV2d.jl:16-30 (a function of the geometry and the flow
  // time only, taken from the oracle's own ghost), the isothermal no-slip wall
  void ghost(int64_t i, int k, const T* cur, T* bnd) const {
    const int32_t code = Ix(m->cell_stiffness, nc, i, k);
    const int32_t idx = -code;
    const orc_bc* b = (idx >= 0 && idx <= ORC_MAX_BC) ? &prm->bc[idx] : nullptr;
    const int kind = b ? b->kind : ORC_BC_EXTRAPOLATE;
    const double nx = Dx(m->cell_edges_Nx, nc, i, k), ny = Dx(m->cell_edges_Ny, nc, i, k);
    switch (kind) {
      case ORC_BC_FIXED:
        for (int v = 0; v < 4; ++v) {
          bnd[v] = T(b->data[v]);
          if constexpr (std::is_same_v<T, Dual>)
            if (dbc) bnd[v].d = dbc[4 * idx + v];
        }
        break;
      case ORC_BC_SLIP_WALL:
        bnd[0] = cur[0];
        bnd[1] = cur[1] * T(ny * ny - nx * nx) - T(2.0 * nx * ny) * cur[2];
        bnd[2] = cur[2] * T(nx * nx - ny * ny) - T(2.0 * nx * ny) * cur[1];
        bnd[3] = cur[3];
        break;
      case ORC_BC_TGV_MIRROR: {
        double c[4], g[4];
        for (int v = 0; v < 4; ++v) c[v] = orc_ad::val(cur[v]);
        orc_ref_bc_ghost(m, th, prm, i, k, c, t, g);
        for (int v = 0; v < 4; ++v) bnd[v] = T(g[v]);
        break;
      }
      case ORC_BC_ISOTHERMAL_WALL: {
        const T R(th->RGAS);
        const T Tc = cur[3] / (cur[0] * R);
        T Tg = T(2.0 * b->data[0]) - Tc;
        if (orc_ad::val(Tg) < 0.5 * b->data[0]) Tg = T(0.5 * b->data[0]);
        bnd[0] = cur[3] / (R * Tg); bnd[1] = -cur[1]; bnd[2] = -cur[2]; bnd[3] = cur[3];
        break;
      }
      default:
        for (int v = 0; v < 4; ++v) bnd[v] = cur[v];
    }
  }
  // computeInterfaceSlope (computeslope2d.jl:44-126): a triangle's far nodes,
  // a quad's two far-side nodes averaged, or at a boundary the ghost with
  // uDown = uL and uUp = uR (a first-order face)
  void face(int64_t i, int k, T* e) const {
    const int32_t ek = Ix(m->cell_stiffness, nc, i, k);
    const int32_t type = Ix(m->mesh_connectivity, nc, i, 2);
    const int index = 2 * k - 1;
    T qL[4], qR[4], up[4], dn[4], UpLeft[4], UpRight[4];
    for (int v = 0; v < 4; ++v) { qL[v] = Q[v * nc + i - 1]; up[v] = T(0.0); dn[v] = T(0.0); }
    if (ek >= 1 && ek <= nc) {
      if (type == 3) {
        const int32_t pDown1 = Ix(m->node2cellsL2down, nc, i, index), pUp1 = Ix(m->node2cellsL2up, nc, i, index);
        for (int v = 0; v < 4; ++v) { up[v] = Qn[v * nn + pUp1 - 1]; dn[v] = Qn[v * nn + pDown1 - 1]; }
      } else if (type == 2) {
        const int32_t pDown1 = Ix(m->node2cellsL2down, nc, i, index), pDown2 = Ix(m->node2cellsL2down, nc, i, index + 1);
        const int32_t pUp1 = Ix(m->node2cellsL2up, nc, i, index), pUp2 = Ix(m->node2cellsL2up, nc, i, index + 1);
        for (int v = 0; v < 4; ++v) {
          up[v] = T(0.5) * (Qn[v * nn + pUp1 - 1] + Qn[v * nn + pUp2 - 1]);
          dn[v] = T(0.5) * (Qn[v * nn + pDown1 - 1] + Qn[v * nn + pDown2 - 1]);
        }
      }
      for (int v = 0; v < 4; ++v) qR[v] = Q[v * nc + ek - 1];
    } else {
      ghost(i, k, qL, qR);
      for (int v = 0; v < 4; ++v) { dn[v] = qL[v]; up[v] = qR[v]; }
    }
    const double ksi = prm->limiter_eps;
    for (int v = 0; v < 4; ++v)
      UpLeft[v] = qL[v] + T(0.5) * limiterT(prm->limiter, qL[v] - dn[v], qR[v] - qL[v], ksi);
    for (int v = 0; v < 4; ++v)
      UpRight[v] = qR[v] - T(0.5) * limiterT(prm->limiter, qR[v] - qL[v], up[v] - qR[v], ksi);
    ausm_plusT(UpRight[0], UpRight[1], UpRight[2], UpRight[3], UpLeft[0], UpLeft[1], UpLeft[2], UpLeft[3],
               Dx(m->cell_edges_Nx, nc, i, k), Dx(m->cell_edges_Ny, nc, i, k), Dx(m->cell_edges_length, nc, i, k),
               th->Gamma, e);
  }
  // calcOneStageMUSCLcpuFast, calcOneStageMUSCLcpuFast.jl:3-107 / SOUscheme.jl:56-111
  void stage() {
    const T b(betta), h(dt);
    std::vector<T> wf(2 * nc, T(0.0));  // each cell's wall faces, summed in cell order below
#pragma omp parallel for
    for (int64_t i = 1; i <= nc; ++i) {
      const int nf = Ix(m->mesh_connectivity, nc, i, 3) == 4 ? 4 : 3;
      T e[4][4], F[4];
      for (int k = 1; k <= nf; ++k) {
        face(i, k, e[k - 1]);
        if (wall_code != 0 && Ix(m->cell_stiffness, nc, i, k) == wall_code) {
          wf[2 * (i - 1)] += e[k - 1][1];
          wf[2 * (i - 1) + 1] += e[k - 1][2];
        }
      }
      for (int v = 0; v < 4; ++v)
        F[v] = nf == 3 ? e[0][v] + e[1][v] + e[2][v] : e[0][v] + e[1][v] + e[2][v] + e[3][v];
      const T Rarea(1.0 / m->cell_areas[i - 1]);
      for (int v = 0; v < 4; ++v) {
        const int64_t s = v * nc + i - 1;
        Un[s] = viscous ? (U[s] - F[v] * b * h * Rarea + b * h * Ud[s]) : (U[s] - F[v] * b * h * Rarea);
      }
    }
    for (int64_t i = 0; i < nc; ++i) {
      force[0] += wf[2 * i];
      force[1] += wf[2 * i + 1];
    }
  }
  void step() {
    if (viscous) diff_term();
    stage();
    U.swap(Un);
    prims();
    node_means();
    t = t + dt;
  }
};

bool supported(const orc_mesh* m, const orc_params* prm, int32_t viscous) {
  if (prm->flux != ORC_FLUX_AUSM_PLUS) return false;
  if (viscous && prm->energy_law != ORC_ENERGY_CONST) return false;
  for (int64_t i = 1; i <= m->nCells; ++i) {
    const int32_t nv = Ix(m->mesh_connectivity, m->nCells, i, 3), type = Ix(m->mesh_connectivity, m->nCells, i, 2);
    if (!((nv == 3 && type == 3) || (nv == 4 && type == 2))) return false;
    for (int k = 1; k <= nv; ++k)
      if (Ix(m->cell_stiffness, m->nCells, i, k) > m->nCells) return false;
  }
  return true;
}

}  // namespace

extern "C" {

// N steps of the reference loop from U0 (nCells x 4, column-major) at a fixed
// dt from flow time t0, with the directional derivative along dU0 and, for
// the fixed-state codes, along dbc ([ORC_MAX_BC + 1][4] by -code, or NULL)
// carried along: UN = U_N and dUN its derivative; force / dforce (or NULL):
// the sum over the N steps of the wall force (x, y) of the faces of code
// wall_code (0: none) and its derivative.  Returns -1 where the mesh or the
// parameters are outside what it covers (another flux, k(T), cells other
// than triangles and quads).
int orc_ref_loop_tl(const orc_mesh* m, const orc_thermo* th, const orc_params* prm, int32_t viscous, double betta,
                    double dt, int32_t n_steps, double t0, int32_t wall_code, const double* U0, const double* dU0,
                    const double* dbc, double* UN, double* dUN, double* force, double* dforce) {
  if (!supported(m, prm, viscous)) return -1;
  RefLoop<Dual> L{m, th, prm, viscous, betta, dt, m->nCells, m->nNodes, {}, {}, {}, {}, {}};
  L.t = t0;
  L.wall_code = wall_code;
  L.dbc = dbc;
  const int64_t nc = m->nCells;
  L.U.resize(4 * nc);
  L.Q.resize(4 * nc);
  L.Un.resize(4 * nc);
  L.Ud.assign(4 * nc, Dual(0.0));
  L.Qn.resize(4 * m->nNodes);
  for (int64_t s = 0; s < 4 * nc; ++s) L.U[s] = Dual(U0[s], dU0 ? dU0[s] : 0.0);
  L.prims();
  L.node_means();
  for (int32_t n = 0; n < n_steps; ++n) L.step();
  for (int64_t s = 0; s < 4 * nc; ++s) {
    UN[s] = L.U[s].v;
    if (dUN) dUN[s] = L.U[s].d;
  }
  for (int q = 0; q < 2; ++q) {
    if (force) force[q] = L.force[q].v;
    if (dforce) dforce[q] = L.force[q].d;
  }
  return 0;
}

}  // extern "C"
