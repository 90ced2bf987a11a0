// Discrete adjoint of the reference scheme's time loop (not in the reference,
// which has no adjoint): the transpose of one loop step
//   U' = U - b dt (sum of the cell's face fluxes) / A + b dt Udiff,
// with the face fluxes from vertex MUSCL + limiter + AUSM+ of the primitives Q
// = P(U) and their node means Qn = N(Q), and Udiff = mu Laplacian of the
// cluster means of U through two Green-Gauss levels (calcDiffTerm with
// constant mu and k), on triangle meshes whose faces are all interior.
//
// The nonlinear part is per face-list entry (the face pass's list: a mirrored
// face once from its owner, every other face slot from its cell): the entry's
// reconstruction and AUSM+ are recomputed from Q and Qn and differentiated
// with dual numbers, seeding each input in turn (the templated reference
// arithmetic of gd_ref_math.hpp, so the values are the forward kernels' own).
// Everything else is linear and transposed as gathers over the incidence lists
// of RefAdjMaps, each summed in a fixed order.  Built without contraction, like
// the forward reference kernels.
#include "gd_ad.hpp"
#include "gd_fold.hpp"
#include "gd_internal.hpp"
#include "gd_ref_math.hpp"

namespace gd {
namespace {

constexpr int kB = kThreads;

template <int K>
__device__ __forceinline__ void ld2(const double* p, double* o) {
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const double2 v = *reinterpret_cast<const double2*>(p + 2 * q);
    o[2 * q] = v.x; o[2 * q + 1] = v.y;
  }
}
template <int K>
__device__ __forceinline__ void st2(double* p, const double* o) {
#pragma unroll
  for (int q = 0; q < K; ++q) *reinterpret_cast<double2*>(p + 2 * q) = make_double2(o[2 * q], o[2 * q + 1]);
}

// One face-list entry f: the flux e = AUSM+(fR, fL) enters U' of its cell i
// with -b dt / A_i and, for a mirrored face, of the neighbour j with
// +b dt / A_j; its inputs are qL = Q_i, qR = Q_j and the far nodes' dn, up:
//   fL = qL + 0.5 lim(qL - dn, qR - qL),  fR = qR - 0.5 lim(qR - qL, up - qR).
// Writes (lam qL, lam qR, lam dn, lam up), 16 values.
__global__ __launch_bounds__(kB) void k_radj_face(MeshView m, gd_params p, const double* Q, const double* Qn,
                                                   const double* lamU, double bdt, double* lamE) {
  const int64_t f = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (f >= m.n_rfaces) return;
  const int64_t nc = m.nc, nn = m.nn;
  const int64_t i = m.rf_cell[f];
  const bool mir = f < m.n_rmirror;
  int64_t j;
  int32_t d1, u1;
  double nx, ny, len;
  if (mir) {
    j = m.rf_nb[f]; d1 = m.rf_dn1[f]; u1 = m.rf_up1[f];
    nx = m.rf_nx[f]; ny = m.rf_ny[f]; len = m.rf_len[f];
  } else {
    const int64_t s = (int64_t)m.rf_slot[f] * nc + i;
    j = m.nb[s]; d1 = m.dn1[s]; u1 = m.up1[s];
    nx = m.fnx[s]; ny = m.fny[s]; len = m.flen[s];
  }
  double lam[4];
  const double ci = -bdt * m.rarea[i];
#pragma unroll
  for (int v = 0; v < 4; ++v) lam[v] = ci * lamU[v * nc + i];
  if (mir) {
    const double cj = bdt * m.rarea[j];
#pragma unroll
    for (int v = 0; v < 4; ++v) lam[v] = lam[v] + cj * lamU[v * nc + j];
  }
  double qL[4], qR[4], dn[4], up[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    qL[v] = Q[v * nc + i]; qR[v] = Q[v * nc + j];
    dn[v] = Qn[v * nn + d1]; up[v] = Qn[v * nn + u1];
  }
  const double eps = p.limiter_eps;
  double fL[4], fR[4], dLqL[4], dLqR[4], dLdn[4], dRqR[4], dRqL[4], dRup[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double a = qL[v] - dn[v], b = qR[v] - qL[v], c = up[v] - qR[v];
    const AD l1x = refm::limiter(p.limiter, AD(a, 1.0), AD(b, 0.0), eps);
    const AD l1y = refm::limiter(p.limiter, AD(a, 0.0), AD(b, 1.0), eps);
    const AD l2x = refm::limiter(p.limiter, AD(b, 1.0), AD(c, 0.0), eps);
    const AD l2y = refm::limiter(p.limiter, AD(b, 0.0), AD(c, 1.0), eps);
    fL[v] = qL[v] + 0.5 * l1x.v;
    fR[v] = qR[v] - 0.5 * l2x.v;
    dLqL[v] = 1.0 + 0.5 * (l1x.d - l1y.d);
    dLdn[v] = -0.5 * l1x.d;
    dLqR[v] = 0.5 * l1y.d;
    dRqR[v] = 1.0 - 0.5 * (l2x.d - l2y.d);
    dRqL[v] = 0.5 * l2x.d;
    dRup[v] = -0.5 * l2y.d;
  }
  // AUSM+ with the neighbour side fR as "L": the adjoint of each of its 8
  // inputs from one dual evaluation seeded on that input
  double lfR[4], lfL[4];
  for (int k = 0; k < 8; ++k) {
    AD L[4], R[4], e[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      L[v] = AD(fR[v], k == v ? 1.0 : 0.0);
      R[v] = AD(fL[v], k == v + 4 ? 1.0 : 0.0);
    }
    refm::ausm_plus(L, R, nx, ny, len, p.Gamma, e);
    const double g = lam[0] * e[0].d + lam[1] * e[1].d + lam[2] * e[2].d + lam[3] * e[3].d;
    if (k < 4) lfR[k] = g;
    else lfL[k - 4] = g;
  }
  double o[16];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    o[v] = lfL[v] * dLqL[v] + lfR[v] * dRqL[v];
    o[4 + v] = lfL[v] * dLqR[v] + lfR[v] * dRqR[v];
    o[8 + v] = lfL[v] * dLdn[v];
    o[12 + v] = lfR[v] * dRup[v];
  }
  st2<8>(lamE + 16 * f, o);
}

// Per node: the far-node parts of the entries that read it -> lamQn [nn][4];
// with the viscous term, the second Green-Gauss level transposed: a cell i's
// Laplacian of field c is (sum over its edges of 0.5 (g_a + g_b) (-n) len) / A_i
// with the x/y swap of calcGradients.jl, and enters U'[c+1] with b dt mu_c.
__global__ __launch_bounds__(kB) void k_radj_nodes(MeshView m, RefAdjMaps a, const double* lamE, const double* lamU,
                                                    double bdt, double mu0, double mu1, double mu2, int visc,
                                                    double* lamQn, double* lamG6N) {
  const int64_t J = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (J >= m.nn) return;
  double q[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t r = a.nq_off[J]; r < a.nq_off[J + 1]; ++r) {
    const int32_t code = a.nq_ref[r];
    double x[4];
    ld2<2>(lamE + 16 * (int64_t)(code >> 1) + 8 + 4 * (code & 1), x);
#pragma unroll
    for (int v = 0; v < 4; ++v) q[v] = q[v] + x[v];
  }
  st2<2>(lamQn + 4 * J, q);
  if (!visc) return;
  const int64_t nc = m.nc;
  const double mu[3] = {mu0, mu1, mu2};
  double g[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t r = a.ne_off[J]; r < a.ne_off[J + 1]; ++r) {
    const int64_t s = a.ne_ref[r];
    const int64_t i = s % nc;
    const double fx = 0.5 * -m.fny[s] * m.flen[s], fy = 0.5 * -m.fnx[s] * m.flen[s];
    const double sc = bdt / m.area[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double l = mu[c] * lamU[(c + 1) * nc + i] * sc;
      g[2 * c] = g[2 * c] + fx * l;
      g[2 * c + 1] = g[2 * c + 1] + fy * l;
    }
  }
  st2<3>(lamG6N + 6 * J, g);
}

// Per cell: the node means of the gradients transposed (1 / cluster size),
// then the first Green-Gauss level's 1/A -> tG [nc][6]
__global__ __launch_bounds__(kB) void k_radj_g6(MeshView m, RefAdjMaps a, const double* lamG6N, double* tG) {
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (i >= m.nc) return;
  double g[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t r = a.cm_off[i]; r < a.cm_off[i + 1]; ++r) {
    double x[6];
    ld2<3>(lamG6N + 6 * (int64_t)a.cm_node[r], x);
    const double c = a.cm_c[r];
#pragma unroll
    for (int t = 0; t < 6; ++t) g[t] = g[t] + c * x[t];
  }
  const double A = m.area[i];
#pragma unroll
  for (int t = 0; t < 6; ++t) g[t] = g[t] / A;
  st2<3>(tG + 6 * i, g);
}

// Per node: the first Green-Gauss level transposed -> lamPhiN [nn][4]
__global__ __launch_bounds__(kB) void k_radj_phi(MeshView m, RefAdjMaps a, const double* tG, double* lamPhiN) {
  const int64_t J = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (J >= m.nn) return;
  const int64_t nc = m.nc;
  double ph[4] = {0, 0, 0, 0};
  for (int64_t r = a.ne_off[J]; r < a.ne_off[J + 1]; ++r) {
    const int64_t s = a.ne_ref[r];
    const int64_t i = s % nc;
    const double fx = 0.5 * -m.fny[s] * m.flen[s], fy = 0.5 * -m.fnx[s] * m.flen[s];
    double t[6];
    ld2<3>(tG + 6 * i, t);
#pragma unroll
    for (int c = 0; c < 3; ++c) ph[c] = ph[c] + (fx * t[2 * c] + fy * t[2 * c + 1]);
  }
  st2<2>(lamPhiN + 4 * J, ph);
}

// Per cell: lamQ from the entries that read Q_i and from the node primitives
// (weight / weight sum), then lamU of U_s = lamU' + P'(U_s)^T lamQ, plus the
// cluster means of rho*u, rho*v, E transposed (1 / cluster size).
__global__ __launch_bounds__(kB) void k_radj_cell(MeshView m, RefAdjMaps a, double gamma, const double* lamE,
                                                   const double* lamQn, const double* lamPhiN, const double* U,
                                                   const double* lamU, double* lamOut) {
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (i >= m.nc) return;
  const int64_t nc = m.nc;
  double lq[4] = {0, 0, 0, 0};
  for (int64_t r = a.ce_off[i]; r < a.ce_off[i + 1]; ++r) {
    const int32_t code = a.ce_ref[r];
    double x[4];
    ld2<2>(lamE + 16 * (int64_t)(code >> 1) + 4 * (code & 1), x);
#pragma unroll
    for (int v = 0; v < 4; ++v) lq[v] = lq[v] + x[v];
  }
  double lph[3] = {0, 0, 0};
  for (int64_t r = a.cm_off[i]; r < a.cm_off[i + 1]; ++r) {
    const int64_t n = a.cm_node[r];
    double x[4];
    ld2<2>(lamQn + 4 * n, x);
    const double w = a.cm_w[r];
#pragma unroll
    for (int v = 0; v < 4; ++v) lq[v] = lq[v] + w * x[v];
    if (lamPhiN) {
      double y[4];
      ld2<2>(lamPhiN + 4 * n, y);
      const double c = a.cm_c[r];
#pragma unroll
      for (int t = 0; t < 3; ++t) lph[t] = lph[t] + c * y[t];
    }
  }
  const double u0 = U[i], u1 = U[nc + i], u2 = U[2 * nc + i];
  const double g1 = gamma - 1.0;
  const double r0 = 1.0 / u0;
  // Q = (u0, u1/u0, u2/u0, (gamma-1)(u3 - (u1^2 + u2^2)/(2 u0)))
  double o[4];
  o[0] = lamU[i] + lq[0] - lq[1] * u1 * r0 * r0 - lq[2] * u2 * r0 * r0 +
         lq[3] * g1 * 0.5 * (u1 * u1 + u2 * u2) * r0 * r0;
  o[1] = lamU[nc + i] + lq[1] * r0 - lq[3] * g1 * u1 * r0 + lph[0];
  o[2] = lamU[2 * nc + i] + lq[2] * r0 - lq[3] * g1 * u2 * r0 + lph[1];
  o[3] = lamU[3 * nc + i] + lq[3] * g1 + lph[2];
#pragma unroll
  for (int v = 0; v < 4; ++v) lamOut[v * nc + i] = o[v];
}

unsigned nb(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

}  // namespace

hipError_t launch_radj_face(const MeshView& m, const gd_params& p, const double* Q, const double* Qn,
                            const double* lamU, double bdt, double* lamE, hipStream_t s) {
  if (m.n_rfaces == 0) return hipSuccess;
  hipLaunchKernelGGL(k_radj_face, dim3(nb(m.n_rfaces)), dim3(kB), 0, s, m, p, Q, Qn, lamU, bdt, lamE);
  return hipGetLastError();
}
hipError_t launch_radj_nodes(const MeshView& m, const RefAdjMaps& a, const double* lamE, const double* lamU, double bdt,
                             const double mu[3], bool visc, double* lamQn, double* lamG6N, hipStream_t s) {
  hipLaunchKernelGGL(k_radj_nodes, dim3(nb(m.nn)), dim3(kB), 0, s, m, a, lamE, lamU, bdt, mu[0], mu[1], mu[2],
                     visc ? 1 : 0, lamQn, lamG6N);
  return hipGetLastError();
}
hipError_t launch_radj_g6(const MeshView& m, const RefAdjMaps& a, const double* lamG6N, double* tG, hipStream_t s) {
  hipLaunchKernelGGL(k_radj_g6, dim3(nb(m.nc)), dim3(kB), 0, s, m, a, lamG6N, tG);
  return hipGetLastError();
}
hipError_t launch_radj_phi(const MeshView& m, const RefAdjMaps& a, const double* tG, double* lamPhiN, hipStream_t s) {
  hipLaunchKernelGGL(k_radj_phi, dim3(nb(m.nn)), dim3(kB), 0, s, m, a, tG, lamPhiN);
  return hipGetLastError();
}
hipError_t launch_radj_cell(const MeshView& m, const RefAdjMaps& a, double gamma, const double* lamE,
                            const double* lamQn, const double* lamPhiN, const double* U, const double* lamU,
                            double* lamOut, hipStream_t s) {
  hipLaunchKernelGGL(k_radj_cell, dim3(nb(m.nc)), dim3(kB), 0, s, m, a, gamma, lamE, lamQn, lamPhiN, U, lamU,
                     lamOut);
  return hipGetLastError();
}

}  // namespace gd
