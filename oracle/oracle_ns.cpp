// CPU oracle of the ns scheme: TEST INFRASTRUCTURE ONLY.
//
// The reference has no cell-gradient MUSCL, no Venkatakrishnan limiter, no
// Sutherland stress fluxes and no isothermal wall (its viscous term is the
// mu*Laplacian of calcDiffTerm.jl); the configurations that use them define
// this scheme, and this file states it plainly, cell by cell, in the mesh
// layout of mesh2d_Int32 (1-based, column-major):
//
//  pass G  least-squares gradients of q = (rho, u, v, p) from face neighbours
//          (ghost values at boundary faces), weights w_k = 1/|d_k|^2;
//          Venkatakrishnan limiter (1993) per variable,
//          eps^2 = (K h)^3 s_v^2 with h = 2A/perimeter and s_v a scale per variable
//  pass F  one flux per face, evaluated for the cell with the lower id (the
//          other cell takes its negative): reconstruction of each side to the
//          face midpoint, q_f = q + phi (grad q . r_f) with that side's own
//          offset r_f, first order where the reconstructed rho or p is not
//          positive; AUSM+ or Roe with the neighbour side as "L" and the
//          owner's inward normal (the reference's convention,
//          computeslope2d.jl:129); viscous flux with the edge-corrected
//          average face gradient, Sutherland mu(T), kappa = mu cp / Pr;
//          isothermal no-slip walls get a one-sided normal gradient; each cell
//          sums its faces in slot order in FP64 and takes an explicit Euler
//          step.
//
// 'mixed' evaluates slopes, limiter, reconstruction and face fluxes in float,
// and keeps geometry, the flux sum, the update and p/T recovery in double.
#include <cmath>
#include <cstdint>
#include <vector>

#include "dual.hpp"
#include "oracle.h"

namespace {

int g_threads = 1;

using std::fabs;
using std::fmax;
using std::fmin;
using std::sqrt;

using orc_ad::Dual;
using orc_ad::val;
using orc_ad::fabs;
using orc_ad::fmax;
using orc_ad::fmin;
using orc_ad::sqrt;

inline int32_t I(const int32_t* a, int64_t n, int64_t i, int k) { return a[(int64_t)(k - 1) * n + (i - 1)]; }
inline double D(const double* a, int64_t n, int64_t i, int k) { return a[(int64_t)(k - 1) * n + (i - 1)]; }

struct Geo {
  int nf;
  int32_t nb[4];       // 0-based neighbour or -1
  int32_t kn[4];       // the neighbour's slot of the shared face
  int32_t code[4];
  double nx[4], ny[4], len[4], rx[4], ry[4], dx[4], dy[4], wx[4], wy[4];
  double rnx[4], rny[4];  // face midpoint - neighbour centroid, from the neighbour's own vertices
  double area, rarea, hmin;
};

void face_mid(const orc_mesh* m, int64_t i1, int k, double* r) {
  const int64_t n = m->nCells;
  const int32_t a = I(m->cells2nodes, n, i1, 2 * k - 1), b = I(m->cells2nodes, n, i1, 2 * k);
  const int nv = I(m->mesh_connectivity, n, i1, 3);
  double xa = 0, ya = 0, xb = 0, yb = 0;
  for (int v = 0; v < nv; ++v) {
    const int32_t id = I(m->mesh_connectivity, n, i1, 4 + v);
    if (id == a) { xa = D(m->cell_nodes_X, n, i1, v + 1); ya = D(m->cell_nodes_Y, n, i1, v + 1); }
    if (id == b) { xb = D(m->cell_nodes_X, n, i1, v + 1); yb = D(m->cell_nodes_Y, n, i1, v + 1); }
  }
  r[0] = 0.5 * (xa + xb) - D(m->cell_mid_points, n, i1, 1);
  r[1] = 0.5 * (ya + yb) - D(m->cell_mid_points, n, i1, 2);
}

std::vector<Geo> geometry(const orc_mesh* m) {
  const int64_t n = m->nCells;
  std::vector<Geo> g(n);
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int64_t i1 = 1; i1 <= n; ++i1) {
    Geo& G = g[i1 - 1];
    G.nf = I(m->mesh_connectivity, n, i1, 3);
    G.area = m->cell_areas[i1 - 1];
    G.rarea = 1.0 / G.area;
    double perim = 0.0, M11 = 0, M12 = 0, M22 = 0, wk[4];
    for (int k = 1; k <= G.nf; ++k) {
      const int q = k - 1;
      const int32_t e = I(m->cell_stiffness, n, i1, k);
      G.nx[q] = D(m->cell_edges_Nx, n, i1, k);
      G.ny[q] = D(m->cell_edges_Ny, n, i1, k);
      G.len[q] = D(m->cell_edges_length, n, i1, k);
      perim += G.len[q];
      double r[2];
      face_mid(m, i1, k, r);
      G.rx[q] = r[0]; G.ry[q] = r[1];
      if (e >= 1) {
        G.nb[q] = e - 1; G.code[q] = 0;
        int kj = 0;
        for (int t = 1; t <= I(m->mesh_connectivity, n, e, 3); ++t)
          if (I(m->cell_stiffness, n, e, t) == i1) kj = t;
        double rj[2];
        face_mid(m, e, kj, rj);
        G.kn[q] = kj - 1;
        G.rnx[q] = rj[0]; G.rny[q] = rj[1];
        G.dx[q] = r[0] - rj[0]; G.dy[q] = r[1] - rj[1];
      } else {
        G.nb[q] = -1; G.code[q] = e; G.kn[q] = -1; G.rnx[q] = 0.0; G.rny[q] = 0.0;
        const double rn = r[0] * G.nx[q] + r[1] * G.ny[q];
        G.dx[q] = 2.0 * rn * G.nx[q]; G.dy[q] = 2.0 * rn * G.ny[q];
      }
      const double w = 1.0 / (G.dx[q] * G.dx[q] + G.dy[q] * G.dy[q]);
      wk[q] = w;
      M11 += w * G.dx[q] * G.dx[q];
      M12 += w * G.dx[q] * G.dy[q];
      M22 += w * G.dy[q] * G.dy[q];
    }
    G.hmin = 2.0 * G.area / perim;
    const double det = M11 * M22 - M12 * M12;
    for (int q = 0; q < G.nf; ++q) {
      const double ix = (M22 * G.dx[q] - M12 * G.dy[q]) / det;
      const double iy = (-M12 * G.dx[q] + M11 * G.dy[q]) / det;
      G.wx[q] = wk[q] * ix; G.wy[q] = wk[q] * iy;
    }
  }
  return g;
}

// geometry of the last mesh seen (keyed by its arrays), so that repeated steps
// on one mesh do not recompute it
const std::vector<Geo>& cached_geometry(const orc_mesh* m) {
  static std::vector<Geo> g;
  static const double* key_area = nullptr;
  static const int32_t* key_stiff = nullptr;
  static int64_t key_n = -1;
  if (key_area != m->cell_areas || key_stiff != m->cell_stiffness || key_n != m->nCells) {
    g = geometry(m);
    key_area = m->cell_areas; key_stiff = m->cell_stiffness; key_n = m->nCells;
  }
  return g;
}

template <class D>
inline void prims(const D* U, int64_t n, int64_t i, double g, D* q) {
  const D r = U[i], mx = U[n + i], my = U[2 * n + i], E = U[3 * n + i];
  q[0] = r;
  q[1] = mx / r;
  q[2] = my / r;
  q[3] = (g - 1.0) * (E - 0.5 * (mx * mx + my * my) / r);
}

// Sutherland: mu_ref (T/T_ref)^(3/2) (T_ref + S)/(T + S), x^(3/2) = x sqrt(x)
template <class R>
inline R sutherland(R T, const orc_ns_params* p) {
  const R Tr = (R)p->T_ref, S = (R)p->S;
  const R t = T / Tr;
  return (R)p->mu_ref * (t * sqrt(t)) * (Tr + S) / (T + S);
}

// ghost state of a boundary face for the ns scheme
// (bcf[idx] = the fixed state of code -idx at the compute type)
template <class R>
inline void ghost(const orc_ns_params* p, const R (*bcf)[4], int code, const R* q, R nx, R ny, R* g) {
  const int idx = -code;
  const orc_bc* b = (idx >= 0 && idx <= ORC_MAX_BC) ? &p->bc[idx] : nullptr;
  const int kind = b ? b->kind : ORC_BC_EXTRAPOLATE;
  if (kind == ORC_BC_FIXED) {
    for (int v = 0; v < 4; ++v) g[v] = bcf[idx][v];
  } else if (kind == ORC_BC_SLIP_WALL || kind == ORC_BC_ISOTHERMAL_WALL) {
    const R un = q[1] * nx + q[2] * ny;
    g[0] = q[0]; g[1] = q[1] - (R)2 * un * nx; g[2] = q[2] - (R)2 * un * ny; g[3] = q[3];
  } else {
    for (int v = 0; v < 4; ++v) g[v] = q[v];
  }
}

// AUSM+ per unit length, neighbour side "L", normal into the own cell:
// returns -(F . n) (same splitting as the reference scheme)
template <class R>
inline void ausm(const R* L, const R* Rr, R nx, R ny, R g, R* out) {
  auto mp = [](R M) { const R d = M * M - 1; return fabs(M) >= 1 ? (R)0.5 * (M + fabs(M)) : (R)0.25 * (M + 1) * (M + 1) + (R)0.125 * d * d; };
  auto mm = [](R M) { const R d = M * M - 1; return fabs(M) >= 1 ? (R)0.5 * (M - fabs(M)) : -(R)0.25 * (M - 1) * (M - 1) - (R)0.125 * d * d; };
  auto pp = [](R M) { const R d = M * M - 1; return fabs(M) >= 1 ? (M > 0 ? (R)1 : (R)0) : (R)0.25 * (M + 1) * (M + 1) * (2 - M) + (R)0.1875 * M * d * d; };
  auto pm = [](R M) { const R d = M * M - 1; return fabs(M) >= 1 ? (M < 0 ? (R)1 : (R)0) : (R)0.25 * (M - 1) * (M - 1) * (2 + M) - (R)0.1875 * M * d * d; };
  const R rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const R rR = Rr[0], uR = Rr[1], vR = Rr[2], pR = Rr[3];
  const R htL = (R)0.5 * (uL * uL + vL * vL) + g / (g - 1) * pL / rL;
  const R htR = (R)0.5 * (uR * uR + vR * vR) + g / (g - 1) * pR / rR;
  const R c = 2 * (g - 1) / (g + 1);
  const R unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
  const R asL = sqrt(c * htL), asR = sqrt(c * htR);
  const R aL = asL * fmin((R)1, asL / fabs(unL));
  const R aR = asR * fmin((R)1, asR / fabs(unR));
  const R a12 = fmin(aL, aR);
  const R ML = unL / a12, MR = unR / a12;
  const R m = mp(ML) + mm(MR);
  const R p12 = pp(ML) * pL + pm(MR) * pR;
  const R mpl = (R)0.5 * (m + fabs(m)), mmi = (R)0.5 * (m - fabs(m));
  out[0] = -(a12 * (mpl * rL + mmi * rR));
  out[1] = -(a12 * (mpl * rL * uL + mmi * rR * uR) + p12 * nx);
  out[2] = -(a12 * (mpl * rL * vL + mmi * rR * vR) + p12 * ny);
  out[3] = -(a12 * (mpl * rL * htL + mmi * rR * htR));
}

// Roe per unit length with Harten's fix (width 0.2 of the wave speed scale, as
// RoeFlux2dFast.jl:85-96 but relative to the Roe sound speed so that it is
// independent of the units)
template <class R>
inline void roe(const R* L, const R* Rr, R nx, R ny, R g, R* out) {
  const R rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const R rR = Rr[0], uR = Rr[1], vR = Rr[2], pR = Rr[3];
  const R mx = -ny, my = nx;
  const R unL = uL * nx + vL * ny, umL = uL * mx + vL * my;
  const R unR = uR * nx + vR * ny, umR = uR * mx + vR * my;
  const R HL = g / (g - 1) * pL / rL + (R)0.5 * (uL * uL + vL * vL);
  const R HR = g / (g - 1) * pR / rR + (R)0.5 * (uR * uR + vR * vR);
  const R RT = sqrt(rR / rL);
  const R rho = RT * rL;
  const R u = (uL + RT * uR) / (1 + RT), v = (vL + RT * vR) / (1 + RT), H = (HL + RT * HR) / (1 + RT);
  const R a = sqrt((g - 1) * (H - (R)0.5 * (u * u + v * v)));
  const R un = u * nx + v * ny, um = u * mx + v * my;
  const R dp = pR - pL, dun = unR - unL, dum = umR - umL, drho = rR - rL;
  const R w1 = (dp - rho * a * dun) / (2 * a * a), w2 = rho * dum, w3 = drho - dp / (a * a),
          w4 = (dp + rho * a * dun) / (2 * a * a);
  R s1 = fabs(un - a), s4 = fabs(un + a);
  const R s2 = fabs(un);
  const R dw = (R)0.2 * a;
  if (s1 < dw) s1 = (R)0.5 * (s1 * s1 / dw + dw);
  if (s4 < dw) s4 = (R)0.5 * (s4 * s4 / dw + dw);
  const R d1 = s1 * w1 + s2 * w3 + s4 * w4;
  const R d2 = s1 * w1 * (u - a * nx) + s2 * w2 * mx + s2 * w3 * u + s4 * w4 * (u + a * nx);
  const R d3 = s1 * w1 * (v - a * ny) + s2 * w2 * my + s2 * w3 * v + s4 * w4 * (v + a * ny);
  const R d4 = s1 * w1 * (H - un * a) + s2 * w2 * um + s2 * w3 * (R)0.5 * (u * u + v * v) + s4 * w4 * (H + un * a);
  out[0] = -(R)0.5 * (rL * unL + rR * unR - d1);
  out[1] = -(R)0.5 * (rL * unL * uL + pL * nx + rR * unR * uR + pR * nx - d2);
  out[2] = -(R)0.5 * (rL * unL * vL + pL * ny + rR * unR * vR + pR * ny - d3);
  out[3] = -(R)0.5 * (rL * unL * HL + rR * unR * HR - d4);
}

// viscous flux per unit length through a face with inward normal n:
// (0, tau.n, (tau.u).n + kappa grad T.n), kappa = mu cp / Pr
template <class R>
inline void visc_face(const orc_ns_params* p, R cp, R uf, R vf, R Tf, const R (*gf)[2], R nx, R ny, R* gv) {
  const R mu = sutherland<R>(Tf, p);
  const R kap = mu * cp / (R)p->Pr;
  const R div = gf[0][0] + gf[1][1];
  const R txx = mu * (2 * gf[0][0] - (R)(2.0 / 3.0) * div);
  const R tyy = mu * (2 * gf[1][1] - (R)(2.0 / 3.0) * div);
  const R txy = mu * (gf[0][1] + gf[1][0]);
  gv[0] = 0;
  gv[1] = txx * nx + txy * ny;
  gv[2] = txy * nx + tyy * ny;
  gv[3] = uf * gv[1] + vf * gv[2] + kap * (gf[2][0] * nx + gf[2][1] * ny);
}

// R: slope/limiter/flux arithmetic; D: state, flux sum and update (double, or
// Dual for the tangent-linear step).  bcf: fixed states per code.
template <class R, class D>
int ns_step(const orc_mesh* m, const orc_ns_params* p, const R (*bcf)[4], double dt, const D* U, D* Un,
            double* grad8, double* phi4, D* force2) {
  const int64_t n = m->nCells;
  const std::vector<Geo>& G = cached_geometry(m);
  const double g = p->gamma;
  const R gR = (R)g, RR = (R)p->R;
  std::vector<R> q(4 * n), gr(8 * n), ph(4 * n);
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    D qq[4];
    prims(U, n, i, g, qq);
    for (int v = 0; v < 4; ++v) q[v * n + i] = (R)qq[v];
  }
  // ---- pass G
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const Geo& c = G[i];
    R qi[4], gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0}, qmx[4], qmn[4];
    for (int v = 0; v < 4; ++v) { qi[v] = q[v * n + i]; qmx[v] = qi[v]; qmn[v] = qi[v]; }
    for (int k = 0; k < c.nf; ++k) {
      R qk[4];
      if (c.nb[k] >= 0) for (int v = 0; v < 4; ++v) qk[v] = q[v * n + c.nb[k]];
      else ghost<R>(p, bcf, c.code[k], qi, (R)c.nx[k], (R)c.ny[k], qk);
      for (int v = 0; v < 4; ++v) {
        const R d = qk[v] - qi[v];
        gx[v] += (R)c.wx[k] * d;
        gy[v] += (R)c.wy[k] * d;
        qmx[v] = fmax(qmx[v], qk[v]);
        qmn[v] = fmin(qmn[v], qk[v]);
      }
    }
    const R kh = (R)(p->venkat_K * c.hmin);
    const R kh3 = kh * kh * kh;
    for (int v = 0; v < 4; ++v) {
      const R e2 = kh3 * (R)(p->venkat_scale[v] * p->venkat_scale[v]);
      R phi = 1;
      for (int k = 0; k < c.nf; ++k) {
        const R d2 = gx[v] * (R)c.rx[k] + gy[v] * (R)c.ry[k];
        R d1;
        if (d2 > 0) d1 = qmx[v] - qi[v];
        else if (d2 < 0) d1 = qmn[v] - qi[v];
        else continue;
        const R psi = ((d1 * d1 + e2) * d2 + 2 * d2 * d2 * d1) / (d1 * d1 + 2 * d2 * d2 + d1 * d2 + e2) / d2;
        phi = fmin(phi, psi);
      }
      ph[v * n + i] = fmax(phi, (R)0);
      gr[(2 * v) * n + i] = gx[v];
      gr[(2 * v + 1) * n + i] = gy[v];
    }
  }
  if (grad8) for (int64_t t = 0; t < 8 * n; ++t) grad8[t] = val(gr[t]);
  if (phi4) for (int64_t t = 0; t < 4 * n; ++t) phi4[t] = val(ph[t]);
  // ---- pass F: face fluxes (owner = lower cell id), then per-cell sums
  const R cp = gR * RR / (gR - 1);
  std::vector<D> ff((size_t)16 * n, D(0.0));  // [slot][var][cell]: flux leaving the cell through the face
  std::vector<D> wall(force2 ? 2 * n : 0, D(0.0));
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const Geo& c = G[i];
    R qi[4];
    for (int v = 0; v < 4; ++v) qi[v] = q[v * n + i];
    for (int k = 0; k < c.nf; ++k) {
      const int32_t j = c.nb[k];
      if (j >= 0 && j < i) continue;  // evaluated by the neighbour
      const R nx = (R)c.nx[k], ny = (R)c.ny[k];
      R qa[4];  // owner side at the face
      for (int v = 0; v < 4; ++v)
        qa[v] = qi[v] + ph[v * n + i] * (gr[(2 * v) * n + i] * (R)c.rx[k] + gr[(2 * v + 1) * n + i] * (R)c.ry[k]);
      if (!(qa[0] > 0) || !(qa[3] > 0)) for (int v = 0; v < 4; ++v) qa[v] = qi[v];
      R qb[4];  // neighbour side at the face, from the neighbour's own offset
      if (j >= 0) {
        const R rxj = (R)c.rnx[k], ryj = (R)c.rny[k];
        for (int v = 0; v < 4; ++v)
          qb[v] = q[v * n + j] + ph[v * n + j] * (gr[(2 * v) * n + j] * rxj + gr[(2 * v + 1) * n + j] * ryj);
        if (!(qb[0] > 0) || !(qb[3] > 0)) for (int v = 0; v < 4; ++v) qb[v] = q[v * n + j];
      } else {
        ghost<R>(p, bcf, c.code[k], qa, nx, ny, qb);
      }
      R e[4];
      if (p->flux == ORC_FLUX_ROE) roe<R>(qb, qa, nx, ny, gR, e);
      else ausm<R>(qb, qa, nx, ny, gR, e);
      R gv[4] = {0, 0, 0, 0};
      if (p->viscous) {
        const int idx = j >= 0 ? 0 : -c.code[k];
        const int kind = (j < 0 && idx <= ORC_MAX_BC) ? p->bc[idx].kind : -1;
        if (j >= 0) {
          // edge-corrected average of the two cell gradients of u, v, T
          R qj[4];
          for (int v = 0; v < 4; ++v) qj[v] = q[v * n + j];
          const R rRi = 1 / (qi[0] * RR), rRj = 1 / (qj[0] * RR);
          const R Ti = qi[3] * rRi, Tj = qj[3] * rRj;
          R g_i[3][2], g_j[3][2];
          for (int s = 0; s < 2; ++s) {
            g_i[0][s] = gr[(2 + s) * n + i]; g_i[1][s] = gr[(4 + s) * n + i];
            g_j[0][s] = gr[(2 + s) * n + j]; g_j[1][s] = gr[(4 + s) * n + j];
            g_i[2][s] = (gr[(6 + s) * n + i] - Ti * RR * gr[s * n + i]) * rRi;
            g_j[2][s] = (gr[(6 + s) * n + j] - Tj * RR * gr[s * n + j]) * rRj;
          }
          const R ddx = (R)c.dx[k], ddy = (R)c.dy[k];
          const R rdl = 1 / sqrt(ddx * ddx + ddy * ddy);
          const R tx = ddx * rdl, ty = ddy * rdl;
          const R dphi[3] = {qj[1] - qi[1], qj[2] - qi[2], Tj - Ti};
          R gf[3][2];
          for (int s = 0; s < 3; ++s) {
            const R ax = (R)0.5 * (g_i[s][0] + g_j[s][0]), ay = (R)0.5 * (g_i[s][1] + g_j[s][1]);
            const R corr = dphi[s] * rdl - (ax * tx + ay * ty);
            gf[s][0] = ax + corr * tx; gf[s][1] = ay + corr * ty;
          }
          visc_face<R>(p, cp, (R)0.5 * (qi[1] + qj[1]), (R)0.5 * (qi[2] + qj[2]), (R)0.5 * (Ti + Tj), gf, nx, ny, gv);
        } else if (kind == ORC_BC_ISOTHERMAL_WALL) {
          // one-sided normal gradient to the wall (u = 0, T = T_w); the fluid's
          // outward normal is -n
          const R Tw = (R)p->bc[idx].data[0];
          const R Ti = qi[3] * (1 / (qi[0] * RR));
          const R rdp = -1 / ((R)c.rx[k] * nx + (R)c.ry[k] * ny);
          const R dudn = (0 - qi[1]) * rdp, dvdn = (0 - qi[2]) * rdp, dTdn = (Tw - Ti) * rdp;
          const R ox = -nx, oy = -ny;
          const R gf[3][2] = {{dudn * ox, dudn * oy}, {dvdn * ox, dvdn * oy}, {dTdn * ox, dTdn * oy}};
          visc_face<R>(p, cp, (R)0, (R)0, Tw, gf, nx, ny, gv);
        }
      }
      D fk[4];
      for (int v = 0; v < 4; ++v) fk[v] = ((D)e[v] + (D)gv[v]) * c.len[k];
      for (int v = 0; v < 4; ++v) ff[((size_t)k * 4 + v) * n + i] = fk[v];
      if (j >= 0) {
        for (int v = 0; v < 4; ++v) ff[((size_t)c.kn[k] * 4 + v) * n + j] = -fk[v];
      } else if (force2) {
        const int idx = -c.code[k];
        if (idx <= ORC_MAX_BC && (p->bc[idx].kind == ORC_BC_ISOTHERMAL_WALL || p->bc[idx].kind == ORC_BC_SLIP_WALL)) {
          wall[2 * i] += fk[1]; wall[2 * i + 1] += fk[2];
        }
      }
    }
  }
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const Geo& c = G[i];
    D F[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < c.nf; ++k)
      for (int v = 0; v < 4; ++v) F[v] += ff[((size_t)k * 4 + v) * n + i];
    for (int v = 0; v < 4; ++v) Un[v * n + i] = U[v * n + i] - F[v] * dt * c.rarea;
  }
  if (force2) {
    D fx = 0.0, fy = 0.0;
    for (int64_t i = 0; i < n; ++i) { fx += wall[2 * i]; fy += wall[2 * i + 1]; }
    force2[0] = fx; force2[1] = fy;
  }
  return 0;
}

}  // namespace

extern "C" {

void orc_set_threads(int32_t n) { g_threads = n > 0 ? n : 1; }

int orc_ns_step(const orc_mesh* m, const orc_ns_params* p, double dt, const double* U, double* Un, double* grad8,
                double* phi4, double* force2) {
  if (p->mixed) {
    float bcf[ORC_MAX_BC + 1][4];
    for (int c = 0; c <= ORC_MAX_BC; ++c)
      for (int v = 0; v < 4; ++v) bcf[c][v] = (float)p->bc[c].data[v];
    return ns_step<float, double>(m, p, bcf, dt, U, Un, grad8, phi4, force2);
  }
  double bcf[ORC_MAX_BC + 1][4];
  for (int c = 0; c <= ORC_MAX_BC; ++c)
    for (int v = 0; v < 4; ++v) bcf[c][v] = p->bc[c].data[v];
  return ns_step<double, double>(m, p, bcf, dt, U, Un, grad8, phi4, force2);
}

// Tangent-linear FP64 step: (U, dU) -> (Un, dUn) and the wall force with its
// derivative; dbc (may be NULL) is the tangent of the fixed boundary states,
// [ORC_MAX_BC + 1][4].
int orc_ns_step_tl(const orc_mesh* m, const orc_ns_params* p, double dt, const double* U, const double* dU,
                   const double* dbc, double* Un, double* dUn, double* force2, double* dforce2) {
  if (p->mixed) return -1;
  const int64_t n = m->nCells;
  std::vector<Dual> u(4 * n), un(4 * n);
  for (int64_t t = 0; t < 4 * n; ++t) u[t] = Dual(U[t], dU ? dU[t] : 0.0);
  Dual bcf[ORC_MAX_BC + 1][4];
  for (int c = 0; c <= ORC_MAX_BC; ++c)
    for (int v = 0; v < 4; ++v) bcf[c][v] = Dual(p->bc[c].data[v], dbc ? dbc[4 * c + v] : 0.0);
  Dual f2[2];
  const int rc = ns_step<Dual, Dual>(m, p, bcf, dt, u.data(), un.data(), nullptr, nullptr, f2);
  for (int64_t t = 0; t < 4 * n; ++t) { Un[t] = un[t].v; dUn[t] = un[t].d; }
  if (force2) { force2[0] = f2[0].v; force2[1] = f2[1].v; }
  if (dforce2) { dforce2[0] = f2[0].d; dforce2[1] = f2[1].d; }
  return rc;
}

// CFL dt of the ns scheme: cfl * min_i h_i / (|u|+a + 2 gamma mu/(Pr rho h_i))
double orc_ns_dt(const orc_mesh* m, const orc_ns_params* p, const double* U, double cfl) {
  const int64_t n = m->nCells;
  const std::vector<Geo>& G = cached_geometry(m);
  double best = INFINITY;
  for (int64_t i = 0; i < n; ++i) {
    double q[4];
    prims<double>(U, n, i, p->gamma, q);
    const double a = sqrt(p->gamma * q[3] / q[0]);
    double lam = sqrt(q[1] * q[1] + q[2] * q[2]) + a;
    if (p->viscous) {
      const double T = q[3] / (q[0] * p->R);
      lam += 2.0 * p->gamma * sutherland<double>(T, p) / (p->Pr * q[0] * G[i].hmin);
    }
    best = fmin(best, G[i].hmin / lam);
  }
  return cfl * best;
}

// the ns scheme's AUSM+ alone (q8 = L then R primitives, per unit length;
// the neighbour side is "L"), for the flux's known-answer tests
void orc_ns_ausm_flux(const double* q8, double nx, double ny, double gamma, double* flux4) {
  ausm<double>(q8, q8 + 4, nx, ny, gamma, flux4);
}

// the ns scheme's Roe alone (same convention as orc_ns_ausm_flux)
void orc_ns_roe_flux(const double* q8, double nx, double ny, double gamma, double* flux4) {
  roe<double>(q8, q8 + 4, nx, ny, gamma, flux4);
}

}  // extern "C"
