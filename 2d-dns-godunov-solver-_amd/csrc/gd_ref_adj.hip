// Discrete adjoint of the reference scheme's time loop (not in the reference,
// which has no adjoint): the transpose of one loop step
//   U' = U - b dt (sum of the cell's face fluxes) / A + b dt Udiff,
// with the face fluxes from vertex MUSCL + limiter + AUSM+ of the primitives Q
// = P(U) and their node means Qn = N(Q), and Udiff = mu Laplacian of the
// cluster means of U through two Green-Gauss levels (calcDiffTerm with
// constant mu and k), on triangles and quads with the boundary ghosts, and
// the wall force of one boundary code summed over the steps.
//
// The nonlinear part is per face-list entry (the face pass's list: a mirrored
// face once from its owner, every other face slot from its cell): the entry's
// reconstruction and AUSM+ are recomputed from Q and Qn with the forward
// kernels' arithmetic (so every branch is taken the same way), the limiter's
// two partials carried by a two-direction dual number, and AUSM+ transposed in
// one reverse sweep.  Everything else is linear and transposed as gathers,
// each summed in a fixed order: face-list entries per node and per cell
// (RefAdjMaps' CSR lists), and the node clusters slot by slot.  The adjoint
// state is kept as [cell][4] records, the form every gather reads.  Built
// without contraction, like the forward reference kernels.
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

// Adds, in row order, the 4-double records rec(ref[r]) of the CSR row
// [r0, r1) to acc; refs and records are loaded four at a time before they are
// added, so the gathers are in flight together.
template <class F>
__device__ __forceinline__ void row_sum4(const int32_t* ref, int64_t r0, int64_t r1, F rec, double* acc) {
  for (int64_t r = r0; r < r1; r += 4) {
    int32_t c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = r + u < r1 ? ref[r + u] : -1;
    double x[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (c[u] >= 0) ld2<2>(rec(c[u]), x[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (c[u] >= 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = acc[v] + x[u][v];
      }
  }
}

// derivatives of the AUSM+ splittings of gd_ref_math.hpp (beta 1/8, alpha
// 3/16); at |M| >= 1 the branch's own derivative, with fabs'(M) = +1 at M = 0
__device__ __forceinline__ double dmach_plus(double M) {
  const double s = M < 0.0 ? -1.0 : 1.0;
  return fabs(M) >= 1.0 ? 0.5 * (1.0 + s) : 0.5 * (M + 1.0) + 0.5 * M * (M * M - 1.0);
}
__device__ __forceinline__ double dmach_minus(double M) {
  const double s = M < 0.0 ? -1.0 : 1.0;
  return fabs(M) >= 1.0 ? 0.5 * (1.0 - s) : -0.5 * (M - 1.0) - 0.5 * M * (M * M - 1.0);
}
__device__ __forceinline__ double dpres_plus(double M) {
  return fabs(M) >= 1.0 ? 0.0 : 0.75 * (1.0 - M * M) + (3.0 / 16.0) * (M * M - 1.0) * (5.0 * M * M - 1.0);
}
__device__ __forceinline__ double dpres_minus(double M) {
  return fabs(M) >= 1.0 ? 0.0 : 0.75 * (M * M - 1.0) - (3.0 / 16.0) * (M * M - 1.0) * (5.0 * M * M - 1.0);
}

// The reverse sweep of refm::ausm_plus (AUSMflux2dFast.jl:87-175): with the
// outputs' adjoints w, the gradients of w . out with respect to the states L
// and R.  The forward values are recomputed with the same expressions as
// refm::ausm_plus, so the branches (min(1, a*/|un|), min(aL, aR), |M| >= 1,
// the sign of m) are the forward kernels'.
__device__ __forceinline__ void ausm_plus_adj(const double* L, const double* R, double nx, double ny, double len,
                                              double g, const double* w, double* gL, double* gR) {
  const double rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const double rR = R[0], uR = R[1], vR = R[2], pR = R[3];
  const double htL = fma(0.5, (uL * uL + vL * vL), pL / rL + pL / rL / (g - 1.0));
  const double htR = fma(0.5, (uR * uR + vR * vR), pR / rR + pR / rR / (g - 1.0));
  const double cst = 2.0 * (g - 1.0) / (g + 1.0);
  const double unL = uL * nx + vL * ny;
  const double unR = uR * nx + vR * ny;
  const double asL = sqrt(cst * htL);
  const double asR = sqrt(cst * htR);
  const double tL = asL / fabs(unL), tR = asR / fabs(unR);
  const bool cL = tL < 1.0, cR = tR < 1.0;  // refm::jmin(1, t) keeps t
  const double aL = asL * (cL ? tL : 1.0);
  const double aR = asR * (cR ? tR : 1.0);
  const bool takeR = aR < aL;  // refm::jmin(aL, aR)
  const double a12 = takeR ? aR : aL;
  const double ML = unL / a12, MR = unR / a12;
  const double m = refm::mach_plus(ML) + refm::mach_minus(MR);
  const double sm = m < 0.0 ? -1.0 : 1.0;
  const double mp = 0.5 * (m + fabs(m)), mm = 0.5 * (m - fabs(m));
  // out_k = -(a12 (mp rL phiL_k + mm rR phiR_k) + p12 n_k) len, phi = (1, u, v, ht)
  const double W0 = -w[0] * len, W1 = -w[1] * len, W2 = -w[2] * len, W3 = -w[3] * len;
  const double sL = W0 + W1 * uL + W2 * vL + W3 * htL;
  const double sR = W0 + W1 * uR + W2 * vR + W3 * htR;
  double a12b = mp * rL * sL + mm * rR * sR;
  const double mb = (a12 * rL * sL) * 0.5 * (1.0 + sm) + (a12 * rR * sR) * 0.5 * (1.0 - sm);
  double rLb = a12 * mp * sL, rRb = a12 * mm * sR;
  const double fL = a12 * mp * rL, fR = a12 * mm * rR;
  double uLb = fL * W1, vLb = fL * W2, htLb = fL * W3;
  double uRb = fR * W1, vRb = fR * W2, htRb = fR * W3;
  const double p12b = W1 * nx + W2 * ny;
  double pLb = p12b * refm::pres_plus(ML), pRb = p12b * refm::pres_minus(MR);
  const double MLb = p12b * pL * dpres_plus(ML) + mb * dmach_plus(ML);
  const double MRb = p12b * pR * dpres_minus(MR) + mb * dmach_minus(MR);
  const double ia = 1.0 / a12;
  double unLb = MLb * ia, unRb = MRb * ia;
  a12b = a12b - (MLb * ML + MRb * MR) * ia;
  const double aLb = takeR ? 0.0 : a12b, aRb = takeR ? a12b : 0.0;
  // a = as * min(1, as / |un|): with t = as / |un| < 1, da/das = 2 t and
  // da/d|un| = -t^2
  double asLb = aLb, asRb = aRb;
  if (cL) {
    asLb = aLb * 2.0 * tL;
    unLb = unLb - aLb * tL * tL * (unL < 0.0 ? -1.0 : 1.0);
  }
  if (cR) {
    asRb = aRb * 2.0 * tR;
    unRb = unRb - aRb * tR * tR * (unR < 0.0 ? -1.0 : 1.0);
  }
  // as = sqrt(cst ht)
  htLb = htLb + asLb * (0.5 * cst) / asL;
  htRb = htRb + asRb * (0.5 * cst) / asR;
  uLb = uLb + unLb * nx; vLb = vLb + unLb * ny;
  uRb = uRb + unRb * nx; vRb = vRb + unRb * ny;
  // ht = (u^2 + v^2) / 2 + (p / r) (1 + 1 / (g - 1))
  const double k = 1.0 + 1.0 / (g - 1.0);
  const double kL = htLb * k / rL, kR = htRb * k / rR;
  uLb = uLb + htLb * uL; vLb = vLb + htLb * vL;
  pLb = pLb + kL; rLb = rLb - kL * (pL / rL);
  uRb = uRb + htRb * uR; vRb = vRb + htRb * vR;
  pRb = pRb + kR; rRb = rRb - kR * (pR / rR);
  gL[0] = rLb; gL[1] = uLb; gL[2] = vLb; gL[3] = pLb;
  gR[0] = rRb; gR[1] = uRb; gR[2] = vRb; gR[3] = pRb;
}

// The Jacobian dg/dq (row-major 4 x 4) of refm::bc_ghost for a boundary
// code: zero for a fixed state and the TGV mirror (functions of the geometry
// and the time only), the identity for zero gradient, the reflection of the
// slip wall, and the isothermal wall's (T_g = 2 T_w - T_c unless floored at
// T_w / 2, rho_g = p / (R T_g)).
__device__ __forceinline__ void ghost_jac(const gd_bc* table, int code, const double* q, double nx, double ny,
                                          double R, double* G) {
  const int idx = -code;
  const gd_bc b = (idx >= 0 && idx <= GD_MAX_BC) ? table[idx] : gd_bc{GD_BC_EXTRAPOLATE, 0, {0, 0, 0, 0}};
#pragma unroll
  for (int t = 0; t < 16; ++t) G[t] = 0.0;
  if (b.kind == GD_BC_FIXED || b.kind == GD_BC_TGV_MIRROR) return;
  if (b.kind == GD_BC_SLIP_WALL) {
    // g1 = q1 (ny^2 - nx^2) - 2 nx ny q2,  g2 = q2 (nx^2 - ny^2) - 2 nx ny q1
    const double a = ny * ny - nx * nx, c = 2.0 * nx * ny;
    G[0] = 1.0; G[5] = a; G[6] = -c; G[9] = -c; G[10] = nx * nx - ny * ny; G[15] = 1.0;
    return;
  }
  if (b.kind == GD_BC_ISOTHERMAL_WALL) {
    const double Tw = b.data[0];
    const double Tc = q[3] / (q[0] * R);
    double Tg = 2.0 * Tw - Tc;
    const bool floored = Tg < 0.5 * Tw;
    if (floored) Tg = 0.5 * Tw;
    const double dT0 = floored ? 0.0 : Tc / q[0];          // dTg/drho
    const double dT3 = floored ? 0.0 : -1.0 / (q[0] * R);  // dTg/dp
    const double g0 = q[3] / (R * Tg);
    G[0] = -g0 / Tg * dT0;
    G[3] = 1.0 / (R * Tg) - g0 / Tg * dT3;
    G[5] = -1.0; G[10] = -1.0; G[15] = 1.0;
    return;
  }
  G[0] = 1.0; G[5] = 1.0; G[10] = 1.0; G[15] = 1.0;
}

// One face-list entry f: the flux e = AUSM+(fR, fL) enters U' of its cell i
// with -b dt / A_i and, for a mirrored face, of the neighbour j with
// +b dt / A_j; its inputs are qL = Q_i, qR = Q_j and the far nodes' dn, up
// (a quad's: the mean of its two far-side nodes):
//   fL = qL + 0.5 lim(qL - dn, qR - qL),  fR = qR - 0.5 lim(qR - qL, up - qR).
// At a boundary face qR is the ghost of qL at the step's flow time t, dn = qL
// and up = qR, so everything enters through qL; a face of code wall_code adds
// (fw0, fw1) to the adjoints of its momentum fluxes (the wall force's share
// of the objective).  Writes the adjoints of qL, qR, dn and up to four
// [entry][4] arrays (a quad's dn and up halved, for each of its two nodes),
// and adds a fixed boundary state's adjoint to the entry's record of lbc
// ([one-sided entry][4], or nullptr).
__global__ __launch_bounds__(kB) void k_radj_face(MeshView m, gd_params p, const double* Q, const double* Qn,
                                                   const double* lam, double bdt, double t, int32_t wall_code,
                                                   double fw0, double fw1, double* lqo, double* lqn, double* ldn,
                                                   double* lup, double* lbc) {
  const int64_t f = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (f >= m.n_rfaces) return;
  const int64_t nc = m.nc, nn = m.nn;
  const int64_t i = m.rf_cell[f];
  const bool mir = f < m.n_rmirror;
  int64_t j, s = -1;
  int32_t d1, u1;
  double nx, ny, len;
  if (mir) {
    j = m.rf_nb[f]; d1 = m.rf_dn1[f]; u1 = m.rf_up1[f];
    nx = m.rf_nx[f]; ny = m.rf_ny[f]; len = m.rf_len[f];
  } else {
    s = (int64_t)m.rf_slot[f] * nc + i;
    j = m.nb[s]; d1 = m.dn1[s]; u1 = m.up1[s];
    nx = m.fnx[s]; ny = m.fny[s]; len = m.flen[s];
  }
  double w[4], li[4];
  ld2<2>(lam + 4 * i, li);
  const double ci = -bdt * m.rarea[i];
#pragma unroll
  for (int v = 0; v < 4; ++v) w[v] = ci * li[v];
  if (mir) {
    double lj[4];
    ld2<2>(lam + 4 * j, lj);
    const double cj = bdt * m.rarea[j];
#pragma unroll
    for (int v = 0; v < 4; ++v) w[v] = w[v] + cj * lj[v];
  }
  const double eps = p.limiter_eps;
  double qL[4], qR[4], dn[4], up[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) qL[v] = Q[v * nc + i];
  if (j < 0) {
    // a boundary face: first order, everything through qL
    if (wall_code != 0 && j + 1 == wall_code && m.wall_id[s] >= 0) {
      w[1] = w[1] + fw0;
      w[2] = w[2] + fw1;
    }
    const int32_t na = m.ea[s], nb2 = m.eb[s];
    refm::bc_ghost(p.bc, (int)j + 1, qL, nx, ny, m.xc[i], m.yc[i], m.xn[na], m.yn[na], m.xn[nb2], m.yn[nb2], t,
                   p.RGAS, qR);
    double G[16];
    ghost_jac(p.bc, (int)j + 1, qL, nx, ny, p.RGAS, G);
    double fL[4], fR[4], l1y[4], l2x[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double a = qL[v] - qL[v], b = qR[v] - qL[v], c = qR[v] - qR[v];
      const AD2 l1 = refm::limiter(p.limiter, AD2(a, 1.0, 0.0), AD2(b, 0.0, 1.0), eps);
      const AD2 l2 = refm::limiter(p.limiter, AD2(b, 1.0, 0.0), AD2(c, 0.0, 1.0), eps);
      fL[v] = qL[v] + 0.5 * l1.v;
      fR[v] = qR[v] - 0.5 * l2.v;
      l1y[v] = l1.dy;
      l2x[v] = l2.dx;
    }
    double gfR[4], gfL[4];
    ausm_plus_adj(fR, fL, nx, ny, len, p.Gamma, w, gfR, gfL);
    // dfL_v/dqL_u = d_vu + 0.5 l1y_v (G_vu - d_vu),  dfR_v/dqL_u = G_vu - 0.5 l2x_v (G_vu - d_vu)
    double o[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double dvu = v == u ? 1.0 : 0.0, gm = G[4 * v + u] - dvu;
        o[u] = o[u] + gfL[v] * (dvu + 0.5 * l1y[v] * gm) + gfR[v] * (G[4 * v + u] - 0.5 * l2x[v] * gm);
      }
    st2<2>(lqo + 4 * f, o);
    // a fixed state's adjoint (qR = data, G = 0), summed over the steps per
    // one-sided entry (each entry's own record: no two threads share one)
    const int bidx = -((int)j + 1);
    if (lbc && bidx >= 0 && bidx <= GD_MAX_BC && p.bc[bidx].kind == GD_BC_FIXED) {
      double acc[4];
      double* r = lbc + 4 * (f - m.n_rmirror);
      ld2<2>(r, acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] = acc[v] + (gfL[v] * (0.5 * l1y[v]) + gfR[v] * (1.0 - 0.5 * l2x[v]));
      st2<2>(r, acc);
    }
    return;
  }
  // a quad's one-sided entry reads the mean of its two far-side nodes
  const bool quad = !mir && m.has_quads && m.nv[i] == 4;
#pragma unroll
  for (int v = 0; v < 4; ++v) qR[v] = Q[v * nc + j];
  if (quad) {
    const int32_t d2 = m.dn2[s], u2 = m.up2[s];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      up[v] = 0.5 * (Qn[v * nn + u1] + Qn[v * nn + u2]);
      dn[v] = 0.5 * (Qn[v * nn + d1] + Qn[v * nn + d2]);
    }
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v) { dn[v] = Qn[v * nn + d1]; up[v] = Qn[v * nn + u1]; }
  }
  double fL[4], fR[4], dLqL[4], dLqR[4], dLdn[4], dRqR[4], dRqL[4], dRup[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double a = qL[v] - dn[v], b = qR[v] - qL[v], c = up[v] - qR[v];
    const AD2 l1 = refm::limiter(p.limiter, AD2(a, 1.0, 0.0), AD2(b, 0.0, 1.0), eps);
    const AD2 l2 = refm::limiter(p.limiter, AD2(b, 1.0, 0.0), AD2(c, 0.0, 1.0), eps);
    fL[v] = qL[v] + 0.5 * l1.v;
    fR[v] = qR[v] - 0.5 * l2.v;
    dLqL[v] = 1.0 + 0.5 * (l1.dx - l1.dy);
    dLdn[v] = -0.5 * l1.dx;
    dLqR[v] = 0.5 * l1.dy;
    dRqR[v] = 1.0 - 0.5 * (l2.dx - l2.dy);
    dRqL[v] = 0.5 * l2.dx;
    dRup[v] = -0.5 * l2.dy;
  }
  // AUSM+ takes the neighbour side fR as "L"
  double gfR[4], gfL[4];
  ausm_plus_adj(fR, fL, nx, ny, len, p.Gamma, w, gfR, gfL);
  double o[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) o[v] = gfL[v] * dLqL[v] + gfR[v] * dRqL[v];
  st2<2>(lqo + 4 * f, o);
#pragma unroll
  for (int v = 0; v < 4; ++v) o[v] = gfL[v] * dLqR[v] + gfR[v] * dRqR[v];
  st2<2>(lqn + 4 * f, o);
  const double hq = quad ? 0.5 : 1.0;
#pragma unroll
  for (int v = 0; v < 4; ++v) o[v] = hq * (gfL[v] * dLdn[v]);
  st2<2>(ldn + 4 * f, o);
#pragma unroll
  for (int v = 0; v < 4; ++v) o[v] = hq * (gfR[v] * dRup[v]);
  st2<2>(lup + 4 * f, o);
}

// The viscous term's linear chain, per cell i with vertices J:
//   phiN_J = mean over the cluster of J of (rho u, rho v, E),
//   G_i    = sum_J C_iJ phiN_J         (first Green-Gauss level),
//   gN_J   = mean over the cluster of J of G,
//   Lap_i  = sum_J C_iJ . gN_J          (second level, "gradX" of "gradX" + "gradY" of "gradY"),
// C_iJ = sum over the cell's edges at J of 0.5 (-ny, -nx) len / A_i
// (calcGradients.jl:36-139's swapped normals), and Udiff = mu_c Lap.

// Per node: the far-node parts of the entries that read it -> lamQn [nn][4];
// with the viscous term, the second level transposed over the cluster and
// divided by the cluster size (the node mean's transpose, taken here once per
// node) -> lamG6N [nn][6].  NCL > 0: cluster width known (every slot's loads
// first).
template <int NCL>
__global__ __launch_bounds__(kB) void k_radj_nodes(MeshView m, RefAdjMaps a, const double* ldn, const double* lup,
                                                    const double* lam, double bdt, double mu0, double mu1, double mu2,
                                                    int visc, double* lamQn, double* lamG6N) {
  const int64_t J0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  const int64_t J = J0 < m.nn ? J0 : m.nn - 1;
  const int64_t nn = m.nn;
  double q[4] = {0.0, 0.0, 0.0, 0.0};
  row_sum4(a.nq_ref, a.nq_off[J], a.nq_off[J + 1],
           [&](int32_t code) { return (code & 1 ? lup : ldn) + 4 * (int64_t)(code >> 1); }, q);
  __shared__ WaveStage<2> sq;
  sq.store(lamQn, J0, nn, q);
  if (!visc) return;
  const double mu[3] = {mu0 * bdt, mu1 * bdt, mu2 * bdt};
  double g[6] = {0, 0, 0, 0, 0, 0};
  int cnt = 0;
  auto add = [&](double cx, double cy, const double* l) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double t = mu[c] * l[c + 1];
      g[2 * c] = g[2 * c] + cx * t;
      g[2 * c + 1] = g[2 * c + 1] + cy * t;
    }
    ++cnt;
  };
  if constexpr (NCL > 0) {
    int32_t c[NCL];
    double cx[NCL], cy[NCL], l[NCL][4];
#pragma unroll
    for (int s = 0; s < NCL; ++s) {
      c[s] = m.cl[s * nn + J];
      cx[s] = a.ccx[s * nn + J];
      cy[s] = a.ccy[s * nn + J];
      ld2<2>(lam + 4 * (int64_t)(c[s] >= 0 ? c[s] : 0), l[s]);
    }
#pragma unroll
    for (int s = 0; s < NCL; ++s)
      if (c[s] >= 0) add(cx[s], cy[s], l[s]);
  } else {
    for (int s = 0; s < m.ncl; ++s) {
      const int32_t c = m.cl[s * nn + J];
      if (c < 0) continue;
      double l[4];
      ld2<2>(lam + 4 * (int64_t)c, l);
      add(a.ccx[s * nn + J], a.ccy[s * nn + J], l);
    }
  }
  const double ic = cnt ? 1.0 / (double)cnt : 0.0;
#pragma unroll
  for (int t = 0; t < 6; ++t) g[t] = g[t] * ic;
  __shared__ WaveStage<3> sg;
  sg.store(lamG6N, J0, nn, g);
}

// Per cell: the sum over the clusters it belongs to -> tG [nc][6]
__global__ __launch_bounds__(kB) void k_radj_g6(MeshView m, RefAdjMaps a, const double* lamG6N, double* tG) {
  const int64_t i0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  const int64_t i = i0 < m.nc ? i0 : m.nc - 1;
  double g[6] = {0, 0, 0, 0, 0, 0};
  const int64_t r1 = a.cm_off[i + 1];
  for (int64_t r = a.cm_off[i]; r < r1; r += 4) {
    int32_t n[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) n[u] = r + u < r1 ? a.cm_node[r + u] : -1;
    double x[4][6];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (n[u] >= 0) ld2<3>(lamG6N + 6 * (int64_t)n[u], x[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (n[u] >= 0) {
#pragma unroll
        for (int t = 0; t < 6; ++t) g[t] = g[t] + x[u][t];
      }
  }
  __shared__ WaveStage<3> sg;
  sg.store(tG, i0, m.nc, g);
}

// Per node: the first Green-Gauss level transposed over the cluster, divided
// by the cluster size -> lamPhiN [nn][4]
template <int NCL>
__global__ __launch_bounds__(kB) void k_radj_phi(MeshView m, RefAdjMaps a, const double* tG, double* lamPhiN) {
  const int64_t J0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  const int64_t J = J0 < m.nn ? J0 : m.nn - 1;
  const int64_t nn = m.nn;
  double ph[4] = {0, 0, 0, 0};
  int cnt = 0;
  auto add = [&](double cx, double cy, const double* t) {
#pragma unroll
    for (int c = 0; c < 3; ++c) ph[c] = ph[c] + (cx * t[2 * c] + cy * t[2 * c + 1]);
    ++cnt;
  };
  if constexpr (NCL > 0) {
    int32_t c[NCL];
    double cx[NCL], cy[NCL], t[NCL][6];
#pragma unroll
    for (int s = 0; s < NCL; ++s) {
      c[s] = m.cl[s * nn + J];
      cx[s] = a.ccx[s * nn + J];
      cy[s] = a.ccy[s * nn + J];
      ld2<3>(tG + 6 * (int64_t)(c[s] >= 0 ? c[s] : 0), t[s]);
    }
#pragma unroll
    for (int s = 0; s < NCL; ++s)
      if (c[s] >= 0) add(cx[s], cy[s], t[s]);
  } else {
    for (int s = 0; s < m.ncl; ++s) {
      const int32_t c = m.cl[s * nn + J];
      if (c < 0) continue;
      double t[6];
      ld2<3>(tG + 6 * (int64_t)c, t);
      add(a.ccx[s * nn + J], a.ccy[s * nn + J], t);
    }
  }
  const double ic = cnt ? 1.0 / (double)cnt : 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) ph[c] = ph[c] * ic;
  __shared__ WaveStage<2> sp;
  sp.store(lamPhiN, J0, nn, ph);
}

// Per cell: lamQ from the entries that read Q_i and from the node primitives
// (weight / weight sum), then the adjoint of U_s = lam' + P'(U_s)^T lamQ, with
// P' written in the primitives Q_s = P(U_s), plus the cluster means of rho u,
// rho v, E transposed (lamPhiN already carries 1 / cluster size).
__global__ __launch_bounds__(kB) void k_radj_cell(MeshView m, RefAdjMaps a, double gamma, const double* lqo,
                                                   const double* lqn, const double* lamQn, const double* lamPhiN,
                                                   const double* Q, const double* lam, double* lamOut) {
  const int64_t i0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  const int64_t i = i0 < m.nc ? i0 : m.nc - 1;
  const int64_t nc = m.nc;
  double lq[4] = {0, 0, 0, 0};
  row_sum4(a.ce_ref, a.ce_off[i], a.ce_off[i + 1],
           [&](int32_t code) { return (code & 1 ? lqn : lqo) + 4 * (int64_t)(code >> 1); }, lq);
  double lph[3] = {0, 0, 0};
  const int64_t r1 = a.cm_off[i + 1];
  for (int64_t r = a.cm_off[i]; r < r1; r += 4) {
    int32_t n[4];
    double w[4], x[4][4], y[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      n[u] = r + u < r1 ? a.cm_node[r + u] : -1;
      w[u] = r + u < r1 ? a.cm_w[r + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (n[u] >= 0) {
        ld2<2>(lamQn + 4 * (int64_t)n[u], x[u]);
        if (lamPhiN) ld2<2>(lamPhiN + 4 * (int64_t)n[u], y[u]);
      }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (n[u] >= 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) lq[v] = lq[v] + w[u] * x[u][v];
        if (lamPhiN) {
#pragma unroll
          for (int t = 0; t < 3; ++t) lph[t] = lph[t] + y[u][t];
        }
      }
  }
  const double q0 = Q[i], q1 = Q[nc + i], q2 = Q[2 * nc + i];
  const double g1 = gamma - 1.0;
  const double r0 = 1.0 / q0;
  double l[4];
  ld2<2>(lam + 4 * i, l);
  // Q = (U0, U1/U0, U2/U0, (gamma-1)(U3 - (U1^2 + U2^2)/(2 U0)))
  double o[4];
  o[0] = l[0] + lq[0] - lq[1] * q1 * r0 - lq[2] * q2 * r0 + lq[3] * g1 * 0.5 * (q1 * q1 + q2 * q2);
  o[1] = l[1] + lq[1] * r0 - lq[3] * g1 * q1 + lph[0];
  o[2] = l[2] + lq[2] * r0 - lq[3] * g1 * q2 + lph[1];
  o[3] = l[3] + lq[3] * g1 + lph[2];
  __shared__ WaveStage<2> so;
  so.store(lamOut, i0, nc, o);
}

// [4][nc] <-> [nc][4]
__global__ __launch_bounds__(kB) void k_radj_to_rec(int64_t nc, const double* soa, double* rec) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= nc) return;
  const double o[4] = {soa[i], soa[nc + i], soa[2 * nc + i], soa[3 * nc + i]};
  st2<2>(rec + 4 * i, o);
}
__global__ __launch_bounds__(kB) void k_radj_from_rec(int64_t nc, const double* rec, double* soa) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= nc) return;
  double o[4];
  ld2<2>(rec + 4 * i, o);
#pragma unroll
  for (int v = 0; v < 4; ++v) soa[v * nc + i] = o[v];
}

unsigned nb(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

}  // namespace

hipError_t launch_radj_face(const MeshView& m, const gd_params& p, const double* Q, const double* Qn,
                            const double* lam, double bdt, double t, int32_t wall_code, const double fw[2],
                            double* lqo, double* lqn, double* ldn, double* lup, double* lbc, hipStream_t s) {
  if (m.n_rfaces == 0) return hipSuccess;
  hipLaunchKernelGGL(k_radj_face, dim3(nb(m.n_rfaces)), dim3(kB), 0, s, m, p, Q, Qn, lam, bdt, t, wall_code, fw[0],
                     fw[1], lqo, lqn, ldn, lup, lbc);
  return hipGetLastError();
}

#define GD_RADJ_BY_WIDTH(KER, GRID, S, ...)                                                  \
  switch (m.ncl) {                                                                           \
    case 3: hipLaunchKernelGGL(KER<3>, dim3(GRID), dim3(kB), 0, S, __VA_ARGS__); break;      \
    case 4: hipLaunchKernelGGL(KER<4>, dim3(GRID), dim3(kB), 0, S, __VA_ARGS__); break;      \
    case 5: hipLaunchKernelGGL(KER<5>, dim3(GRID), dim3(kB), 0, S, __VA_ARGS__); break;      \
    case 6: hipLaunchKernelGGL(KER<6>, dim3(GRID), dim3(kB), 0, S, __VA_ARGS__); break;      \
    case 7: hipLaunchKernelGGL(KER<7>, dim3(GRID), dim3(kB), 0, S, __VA_ARGS__); break;      \
    case 8: hipLaunchKernelGGL(KER<8>, dim3(GRID), dim3(kB), 0, S, __VA_ARGS__); break;      \
    default: hipLaunchKernelGGL(KER<0>, dim3(GRID), dim3(kB), 0, S, __VA_ARGS__); break;     \
  }

hipError_t launch_radj_nodes(const MeshView& m, const RefAdjMaps& a, const double* ldn, const double* lup,
                             const double* lam, double bdt, const double mu[3], bool visc, double* lamQn,
                             double* lamG6N, hipStream_t s) {
  GD_RADJ_BY_WIDTH(k_radj_nodes, nb(m.nn), s, m, a, ldn, lup, lam, bdt, mu[0], mu[1], mu[2], visc ? 1 : 0, lamQn,
                   lamG6N)
  return hipGetLastError();
}
hipError_t launch_radj_g6(const MeshView& m, const RefAdjMaps& a, const double* lamG6N, double* tG, hipStream_t s) {
  hipLaunchKernelGGL(k_radj_g6, dim3(nb(m.nc)), dim3(kB), 0, s, m, a, lamG6N, tG);
  return hipGetLastError();
}
hipError_t launch_radj_phi(const MeshView& m, const RefAdjMaps& a, const double* tG, double* lamPhiN, hipStream_t s) {
  GD_RADJ_BY_WIDTH(k_radj_phi, nb(m.nn), s, m, a, tG, lamPhiN)
  return hipGetLastError();
}
hipError_t launch_radj_cell(const MeshView& m, const RefAdjMaps& a, double gamma, const double* lqo,
                            const double* lqn, const double* lamQn, const double* lamPhiN, const double* Q,
                            const double* lam, double* lamOut, hipStream_t s) {
  hipLaunchKernelGGL(k_radj_cell, dim3(nb(m.nc)), dim3(kB), 0, s, m, a, gamma, lqo, lqn, lamQn, lamPhiN, Q, lam,
                     lamOut);
  return hipGetLastError();
}
hipError_t launch_radj_to_rec(int64_t nc, const double* soa, double* rec, hipStream_t s) {
  hipLaunchKernelGGL(k_radj_to_rec, dim3(nb(nc)), dim3(kB), 0, s, nc, soa, rec);
  return hipGetLastError();
}
hipError_t launch_radj_from_rec(int64_t nc, const double* rec, double* soa, hipStream_t s) {
  hipLaunchKernelGGL(k_radj_from_rec, dim3(nb(nc)), dim3(kB), 0, s, nc, rec, soa);
  return hipGetLastError();
}

}  // namespace gd
