// Discrete adjoint of one ns step (AUSM+ or Roe), templated on the face arithmetic
// type like the forward kernels (double: FP64, float: FP32-mixed).
//
// Forward step n, dt fixed:  Q = prims(U^n);  rec = (q, grad q, phi)(Q);
//   f_face = Phi(rec_owner, rec_neighbour);  U^{n+1} = U^n - dt/A sum(+-f_face);
//   J += w . (wall force of step n) = sum over wall faces of w . (f_1, f_2).
// Reverse step n, given lamU^{n+1} and lamF = -dt/A lamU^{n+1} per cell:
//   k_adj_cell   per cell: for each of its faces, lam_f = lamF(owner) -
//                lamF(neighbour) (+ w on wall faces) and the transposed
//                Jacobian of the face flux with respect to this cell's record
//                (the face is evaluated once from each side, in the owner's
//                orientation, so nothing per face is stored); then the
//                transposed Jacobian of the gradient + limiter with respect to
//                the stencil's primitives -> own part, and one part per
//                neighbour slot (nrev); ghost states chained back to the cell
//                or, for fixed states, to the boundary data
//   k_adj_state  per cell: lamU^n = lamU^{n+1} + own part + the neighbours'
//                parts for this cell (the cell pass sends every part already
//                through (dQ/dU)^T of the receiving cell); next lamF
// Every derivative follows the forward code's branches as a forward-mode dual
// number would (fabs'(0) = +1; fmin/fmax take the first argument on ties), so
// the transposed step is exactly the transpose of the tangent-linear step.
#pragma once
#include <hip/hip_runtime.h>

#include "gd_ns_impl.hpp"

namespace gd {
namespace adj {

using ns::FaceGeo;
using ns::kBlock;
using ns::ld16;
using ns::ld4;
using ns::st4;

template <class T>
struct AdjKArgs {
  MeshView m;
  gd_params prm;
  const T* rec;        // [nc][16] forward record of the step
  T* lamF;             // [nc][4]  -dt/A lamU^{n+1}
  T* selfq;            // [nc][4] own part, conservative variables
  T* nrev;             // [nc][nf][4] part for each neighbour, its conservative variables
  const T* sgeo;       // [5][nf][nc] each slot's face geometry (nx, ny, len, other cell's offset x, y)
  double* bcsum;       // [n_wall][4] fixed-state adjoint summed over steps
  const double* lamU_in;
  double* lamU_out;
  double dt;
  int32_t wall_code;
  double fw[2];
};

template <class T>
AdjKArgs<T> kargs(const AdjStepArgs& a) {
  AdjKArgs<T> k;
  k.m = a.m;
  k.prm = a.prm;
  k.rec = static_cast<const T*>(a.rec);
  k.lamF = static_cast<T*>(a.lamF);
  k.selfq = static_cast<T*>(a.selfq);
  k.nrev = static_cast<T*>(a.nrev);
  k.sgeo = static_cast<const T*>(a.sgeo);
  k.bcsum = a.bcsum;
  k.lamU_in = a.lamU_in;
  k.lamU_out = a.lamU_out;
  k.dt = a.dt;
  k.wall_code = a.wall_code;
  k.fw[0] = a.fw[0];
  k.fw[1] = a.fw[1];
  return k;
}

// (dQ/dU)^T l for Q = (rho, u, v, p) of U = (rho, rho u, rho v, E) at the
// primitive state q, in T (the cell pass applies it to every part it sends)
template <class T>
__device__ __forceinline__ void prims_rev(const T* q, T gm1, const T* l, T* o) {
  const T ri = 1 / q[0], u = q[1], v = q[2];
  o[0] = l[0] - (u * l[1] + v * l[2]) * ri + gm1 * (T)0.5 * (u * u + v * v) * l[3];
  o[1] = l[1] * ri - gm1 * u * l[3];
  o[2] = l[2] * ri - gm1 * v * l[3];
  o[3] = gm1 * l[3];
}

// derivative sign of fabs as a dual number takes it
template <class T>
__device__ __forceinline__ T dsgn(T x) { return x < 0 ? (T)-1 : (T)1; }

// Transposed Jacobian of ns::ausm(L, R, n): lo = adjoint of the 4 outputs;
// adds to lL, lR.
template <class T>
__device__ __forceinline__ void ausm_rev(const T* L, const T* R, T nx, T ny, T g, const T* lo, T* lL, T* lR) {
  const T rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const T rR = R[0], uR = R[1], vR = R[2], pR = R[3];
  const T kg = g / (g - 1);
  const T htL = (T)0.5 * (uL * uL + vL * vL) + kg * pL / rL;
  const T htR = (T)0.5 * (uR * uR + vR * vR) + kg * pR / rR;
  const T c = 2 * (g - 1) / (g + 1);
  const T unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
  const T asL = sqrt(c * htL), asR = sqrt(c * htR);
  const T xL = asL / fabs(unL), xR = asR / fabs(unR);
  const T aL = asL * fmin((T)1, xL);
  const T aR = asR * fmin((T)1, xR);
  const bool pickR = aR < aL;  // fmin(aL, aR) as a dual number picks
  const T a12 = pickR ? aR : aL;
  const T ML = unL / a12, MR = unR / a12;
  T mp, pp, dmp, dpp, mm, pm, dmm, dpm;
  {
    const T M = ML, d = M * M - 1;
    if (fabs(M) >= 1) {
      mp = (T)0.5 * (M + fabs(M)); pp = M > 0 ? (T)1 : (T)0;
      dmp = M > 0 ? (T)1 : (T)0; dpp = 0;
    } else {
      mp = (T)0.25 * (M + 1) * (M + 1) + (T)0.125 * d * d;
      pp = (T)0.25 * (M + 1) * (M + 1) * (2 - M) + (T)0.1875 * M * d * d;
      dmp = (T)0.5 * (M + 1) + (T)0.5 * M * d;
      dpp = (T)0.5 * (M + 1) * (2 - M) - (T)0.25 * (M + 1) * (M + 1) + (T)0.1875 * d * d + (T)0.75 * M * M * d;
    }
  }
  {
    const T M = MR, d = M * M - 1;
    if (fabs(M) >= 1) {
      mm = (T)0.5 * (M - fabs(M)); pm = M < 0 ? (T)1 : (T)0;
      dmm = M < 0 ? (T)1 : (T)0; dpm = 0;
    } else {
      mm = -(T)0.25 * (M - 1) * (M - 1) - (T)0.125 * d * d;
      pm = (T)0.25 * (M - 1) * (M - 1) * (2 + M) - (T)0.1875 * M * d * d;
      dmm = -(T)0.5 * (M - 1) - (T)0.5 * M * d;
      dpm = (T)0.5 * (M - 1) * (2 + M) + (T)0.25 * (M - 1) * (M - 1) - (T)0.1875 * d * d - (T)0.75 * M * M * d;
    }
  }
  const T m = mp + mm;
  const bool mpos = !(m < 0);  // fabs'(0) = +1
  const T mpl = mpos ? m : (T)0, mmi = mpos ? (T)0 : m;
  const T A = mpl * rL, B = mmi * rR;
  // out = -(a12 (A psiL + B psiR) + p12 (0, nx, ny, 0)), psi = (1, u, v, ht)
  T la12 = -(lo[0] * (A + B) + lo[1] * (A * uL + B * uR) + lo[2] * (A * vL + B * vR) + lo[3] * (A * htL + B * htR));
  const T lA = -a12 * (lo[0] + lo[1] * uL + lo[2] * vL + lo[3] * htL);
  const T lB = -a12 * (lo[0] + lo[1] * uR + lo[2] * vR + lo[3] * htR);
  T luL = -a12 * A * lo[1], lvL = -a12 * A * lo[2], lhtL = -a12 * A * lo[3];
  T luR = -a12 * B * lo[1], lvR = -a12 * B * lo[2], lhtR = -a12 * B * lo[3];
  const T lp12 = -(lo[1] * nx + lo[2] * ny);
  T lrL = lA * mpl, lrR = lB * mmi;
  const T lm = mpos ? lA * rL : lB * rR;
  T lpL = lp12 * pp, lpR = lp12 * pm;
  const T lML = lm * dmp + lp12 * pL * dpp;
  const T lMR = lm * dmm + lp12 * pR * dpm;
  T lunL = lML / a12, lunR = lMR / a12;
  la12 -= (lML * ML + lMR * MR) / a12;
  const T laL = pickR ? (T)0 : la12, laR = pickR ? la12 : (T)0;
  T lasL, lasR;
  if (xL < 1) {  // aL = asL^2 / |unL|
    lasL = laL * 2 * xL;
    lunL += -laL * xL * xL * dsgn(unL);
  } else {
    lasL = laL;
  }
  if (xR < 1) {
    lasR = laR * 2 * xR;
    lunR += -laR * xR * xR * dsgn(unR);
  } else {
    lasR = laR;
  }
  lhtL += lasL * c / (2 * asL);
  lhtR += lasR * c / (2 * asR);
  luL += lhtL * uL + lunL * nx;
  lvL += lhtL * vL + lunL * ny;
  lpL += lhtL * kg / rL;
  lrL -= lhtL * kg * pL / (rL * rL);
  luR += lhtR * uR + lunR * nx;
  lvR += lhtR * vR + lunR * ny;
  lpR += lhtR * kg / rR;
  lrR -= lhtR * kg * pR / (rR * rR);
  lL[0] += lrL; lL[1] += luL; lL[2] += lvL; lL[3] += lpL;
  lR[0] += lrR; lR[1] += luR; lR[2] += lvR; lR[3] += lpR;
}

// Transposed Jacobian of ns::roe(L, R, n) (Roe averages, wave strengths,
// Harten's fix of width 0.2 a on the acoustic waves): lo = adjoint of the 4
// outputs; adds to lL, lR.  The fix's branches are taken as the forward takes
// them; fabs'(0) = +1.
template <class T>
__device__ __forceinline__ void roe_rev(const T* L, const T* R, T nx, T ny, T g, const T* lo, T* lL, T* lR) {
  // forward recompute
  const T rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const T rR = R[0], uR = R[1], vR = R[2], pR = R[3];
  const T mx = -ny, my = nx;
  const T kg = g / (g - 1);
  const T unL = uL * nx + vL * ny, umL = uL * mx + vL * my;
  const T unR = uR * nx + vR * ny, umR = uR * mx + vR * my;
  const T HL = kg * pL / rL + (T)0.5 * (uL * uL + vL * vL);
  const T HR = kg * pR / rR + (T)0.5 * (uR * uR + vR * vR);
  const T RT = sqrt(rR / rL);
  const T rho = RT * rL;
  const T den = 1 + RT;
  const T u = (uL + RT * uR) / den, v = (vL + RT * vR) / den, H = (HL + RT * HR) / den;
  const T q2 = u * u + v * v;
  const T a = sqrt((g - 1) * (H - (T)0.5 * q2));
  const T un = u * nx + v * ny, um = u * mx + v * my;
  const T dp = pR - pL, dun = unR - unL, dum = umR - umL, drho = rR - rL;
  const T A2 = 2 * a * a;
  const T w1 = (dp - rho * a * dun) / A2, w2 = rho * dum, w3 = drho - dp / (a * a), w4 = (dp + rho * a * dun) / A2;
  const T s1r = fabs(un - a), s4r = fabs(un + a), s2 = fabs(un);
  const T dw = (T)0.2 * a;
  const bool f1 = s1r < dw, f4 = s4r < dw;
  const T s1 = f1 ? (T)0.5 * (s1r * s1r / dw + dw) : s1r;
  const T s4 = f4 ? (T)0.5 * (s4r * s4r / dw + dw) : s4r;
  const T t1 = s1 * w1, t2 = s2 * w2, t3 = s2 * w3, t4 = s4 * w4;
  // out = -1/2 (central part - d)
  const T c = (T)-0.5;
  T lrL = c * (lo[0] * unL + lo[1] * unL * uL + lo[2] * unL * vL + lo[3] * unL * HL);
  T lrR = c * (lo[0] * unR + lo[1] * unR * uR + lo[2] * unR * vR + lo[3] * unR * HR);
  T lunL = c * rL * (lo[0] + lo[1] * uL + lo[2] * vL + lo[3] * HL);
  T lunR = c * rR * (lo[0] + lo[1] * uR + lo[2] * vR + lo[3] * HR);
  T luL = c * lo[1] * rL * unL, lvL = c * lo[2] * rL * unL, lHL = c * lo[3] * rL * unL;
  T luR = c * lo[1] * rR * unR, lvR = c * lo[2] * rR * unR, lHR = c * lo[3] * rR * unR;
  T lpL = c * (lo[1] * nx + lo[2] * ny), lpR = lpL;
  const T l1 = (T)0.5 * lo[0], l2 = (T)0.5 * lo[1], l3 = (T)0.5 * lo[2], l4 = (T)0.5 * lo[3];
  // d1..d4
  const T lt1 = l1 + l2 * (u - a * nx) + l3 * (v - a * ny) + l4 * (H - un * a);
  const T lt2 = l2 * mx + l3 * my + l4 * um;
  const T lt3 = l1 + l2 * u + l3 * v + l4 * (T)0.5 * q2;
  const T lt4 = l1 + l2 * (u + a * nx) + l3 * (v + a * ny) + l4 * (H + un * a);
  T lu = l2 * (t1 + t3 + t4), lv = l3 * (t1 + t3 + t4);
  T la = (t4 - t1) * (l2 * nx + l3 * ny + l4 * un);
  T lH = l4 * (t1 + t4);
  T lun = l4 * a * (t4 - t1);
  T lum = l4 * t2;
  T lq2 = (T)0.5 * l4 * t3;
  // t = s w
  T ls1 = lt1 * w1, ls4 = lt4 * w4;
  const T ls2 = lt2 * w2 + lt3 * w3;
  const T lw1 = lt1 * s1, lw2 = lt2 * s2, lw3 = lt3 * s2, lw4 = lt4 * s4;
  // Harten's fix
  T ldw = 0;
  if (f1) { ldw += ls1 * (T)0.5 * (1 - s1r * s1r / (dw * dw)); ls1 = ls1 * s1r / dw; }
  if (f4) { ldw += ls4 * (T)0.5 * (1 - s4r * s4r / (dw * dw)); ls4 = ls4 * s4r / dw; }
  la += (T)0.2 * ldw;
  const T g1 = ls1 * dsgn(un - a), g4 = ls4 * dsgn(un + a);
  lun += g1 + g4 + ls2 * dsgn(un);
  la += g4 - g1;
  // wave strengths
  const T ln1 = lw1 / A2, ln4 = lw4 / A2;
  const T lA2 = -(lw1 * w1 + lw4 * w4) / A2;
  T ldp = ln1 + ln4 - lw3 / (a * a);
  T lrho = (ln4 - ln1) * a * dun + lw2 * dum;
  la += (ln4 - ln1) * rho * dun + lA2 * 4 * a + lw3 * 2 * dp / (a * a * a);
  const T ldun = (ln4 - ln1) * rho * a;
  const T ldum = lw2 * rho;
  const T ldrho = lw3;
  lpR += ldp; lpL -= ldp;
  lunR += ldun; lunL -= ldun;
  T lumR = ldum, lumL = -ldum;
  lrR += ldrho; lrL -= ldrho;
  // averaged state
  lu += lun * nx + lum * mx;
  lv += lun * ny + lum * my;
  const T lX = la * (g - 1) / (2 * a);  // a = sqrt((g - 1) X), X = H - q2 / 2
  lH += lX;
  lq2 -= (T)0.5 * lX;
  lu += 2 * u * lq2;
  lv += 2 * v * lq2;
  T lRT = (lu * (uR - u) + lv * (vR - v) + lH * (HR - H)) / den + lrho * rL;
  luL += lu / den; luR += lu * RT / den;
  lvL += lv / den; lvR += lv * RT / den;
  lHL += lH / den; lHR += lH * RT / den;
  lrL += lrho * RT;
  const T lrat = lRT / (2 * RT);  // RT = sqrt(rR / rL)
  lrR += lrat / rL;
  lrL -= lrat * rR / (rL * rL);
  // H of each side, then the normal and tangential velocities
  lpL += lHL * kg / rL; lrL -= lHL * kg * pL / (rL * rL); luL += lHL * uL; lvL += lHL * vL;
  lpR += lHR * kg / rR; lrR -= lHR * kg * pR / (rR * rR); luR += lHR * uR; lvR += lHR * vR;
  luL += lunL * nx + lumL * mx; lvL += lunL * ny + lumL * my;
  luR += lunR * nx + lumR * mx; lvR += lunR * ny + lumR * my;
  lL[0] += lrL; lL[1] += luL; lL[2] += lvL; lL[3] += lpL;
  lR[0] += lrR; lR[1] += luR; lR[2] += lvR; lR[3] += lpR;
}

// Transposed Jacobian of ns::viscous_face with respect to (uf, vf, Tf, gf);
// lg = adjoint of gv[1..3] (gv[0] = 0).
template <class T>
__device__ __forceinline__ void viscous_face_rev(const gd_params& P, T cp, T uf, T vf, T Tf, const T (*gf)[2], T nx,
                                                 T ny, const T* lg, T* luf, T* lvf, T* lTf, T (*lgf)[2]) {
  const T mu = ns::sutherland<T>(Tf, P);
  const T kap = mu * cp / (T)P.Pr;
  const T div = gf[0][0] + gf[1][1];
  const T sxx = 2 * gf[0][0] - (T)(2.0 / 3.0) * div, syy = 2 * gf[1][1] - (T)(2.0 / 3.0) * div;
  const T sxy = gf[0][1] + gf[1][0];
  const T gv1 = mu * sxx * nx + mu * sxy * ny, gv2 = mu * sxy * nx + mu * syy * ny;
  const T l1 = lg[1] + uf * lg[3], l2 = lg[2] + vf * lg[3];
  *luf = gv1 * lg[3];
  *lvf = gv2 * lg[3];
  const T ltxx = l1 * nx, ltxy = l1 * ny + l2 * nx, ltyy = l2 * ny;
  const T qn = gf[2][0] * nx + gf[2][1] * ny;
  const T lkap = lg[3] * qn;
  lgf[2][0] = lg[3] * kap * nx;
  lgf[2][1] = lg[3] * kap * ny;
  const T lmu = ltxx * sxx + ltyy * syy + ltxy * sxy + lkap * cp / (T)P.Pr;
  const T ltr = (T)(2.0 / 3.0) * (ltxx + ltyy);
  lgf[0][0] = mu * (2 * ltxx - ltr);
  lgf[1][1] = mu * (2 * ltyy - ltr);
  lgf[0][1] = mu * ltxy;
  lgf[1][0] = mu * ltxy;
  const T S = (T)P.S_suth;
  *lTf = lmu * mu * ((T)1.5 / Tf - 1 / (Tf + S));
}

// adjoint of T = p / (rho R) and of the T gradient (gp - T R grho) / (rho R)
// of one cell record, added into its record adjoint lr
template <class T>
__device__ __forceinline__ void temp_rev(const T* r, T Rg, T lT, const T* lgT, T* lr) {
  const T rho = r[0], p = r[3];
  const T rR = 1 / (rho * Rg);
  const T Tc = p * rR;
  T lrR = 0;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const T grho = r[4 + c], gp = r[10 + c];
    lr[10 + c] += lgT[c] * rR;
    lr[4 + c] -= lgT[c] * Tc * Rg * rR;
    lT -= lgT[c] * Rg * grho * rR;
    lrR += lgT[c] * (gp - Tc * Rg * grho);
  }
  lr[3] += lT * rR;
  lrR += lT * p;
  lr[0] -= lrR * rR / rho;
}

// Geometry of one face as the transposed flux needs it (loaded ahead of use)
template <class T>
struct FaceIn {
  T nx, ny, len;  // owner's inward normal, length
  T ox, oy;       // face midpoint - the other cell's centroid (0 at a boundary)
  bool on_wall;   // its force enters J
};

// Each cell slot's face geometry copied out of the face list into the slot's
// place, so that the cell pass reads it unit-stride instead of gathering it
// through the face ids (the values are the flux pass's, bit for bit): the
// owner's normal and the length, and the other cell's offset to the face
// midpoint.  The cell's own offset is its slot offset r_f, which the face list
// holds as an exact copy (both are computed from the same cell's vertices).
template <class T>
__global__ __launch_bounds__(kBlock) void k_slot_geo(MeshView m, T* sg) {
  using G = FaceGeo<T>;
  const int64_t nc = m.nc, ns = (int64_t)m.nf * nc;
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= nc) return;
  for (int k = 0; k < m.nf; ++k) {
    const int32_t cf = m.cface[k * nc + i];
    const int64_t f = cf >> 1;
    const bool nbr = (cf & 1) != 0;  // this cell is the face's neighbour
    T v[kAdjSlotGeo] = {0, 0, 0, 0, 0};
    if (cf >= 0) {
      v[0] = G::fnx(m, f); v[1] = G::fny(m, f); v[2] = G::flen(m, f);
      v[3] = nbr ? G::fr0(m, 0, f) : G::fr1(m, 0, f);
      v[4] = nbr ? G::fr0(m, 1, f) : G::fr1(m, 1, f);
    }
#pragma unroll
    for (int q = 0; q < kAdjSlotGeo; ++q) sg[q * ns + k * nc + i] = v[q];
  }
}

template <class T>
hipError_t launch_slot_geo(const MeshView& m, T* sg, hipStream_t s) {
  hipLaunchKernelGGL((k_slot_geo<T>), dim3(ns::nblocks(m.nc)), dim3(kBlock), 0, s, m, sg);
  return hipGetLastError();
}

// slot k of cell i; c1 < 0: the boundary code - 1; f: the face (wall flag)
template <class T>
__device__ __forceinline__ FaceIn<T> load_face(const AdjKArgs<T>& a, int k, int64_t i, int64_t f, int32_t c1) {
  const MeshView& m = a.m;
  const int64_t ns = (int64_t)m.nf * m.nc, s = k * m.nc + i;
  const T* g0 = a.sgeo + s;
  FaceIn<T> g;
  g.nx = ns::ntl(g0); g.ny = ns::ntl(g0 + ns);
  g.len = ns::ntl(g0 + 2 * ns);
  g.ox = 0; g.oy = 0;
  g.on_wall = false;
  if (c1 >= 0) {
    g.ox = ns::ntl(g0 + 3 * ns); g.oy = ns::ntl(g0 + 4 * ns);
  } else {
    g.on_wall = a.wall_code == c1 + 1 && m.fwall[f] >= 0;
  }
  return g;
}

// Transposed face flux for one side of a face, from that side.  (sx, sy):
// face midpoint - this cell's centroid; ri / rj:
// this cell's and the other cell's records (a boundary face passes this
// cell's twice; this cell then is the owner); li / lj: their lamF (lj = 0 at
// a boundary); sd: this cell is the face's neighbour; c1 < 0: the boundary
// code - 1.  The face is evaluated in the owner's orientation exactly as the
// flux pass evaluates it: the owner-minus-neighbour differences are negated
// when this cell is the neighbour (an exact operation) and the two-sided
// averages are sums, which do not depend on the order of their operands.
// Adds the adjoint of this cell's record to lr; at a boundary face returns the
// fixed-state adjoint in bq (zero unless the code is GD_BC_FIXED).
template <class T, bool VISC, bool ROE>
__device__ __forceinline__ void face_rev(const AdjKArgs<T>& a, const FaceIn<T>& fg, T sx, T sy, int32_t c1, bool sd,
                                         const T* ri, const T* rj, const T* li, const T* lj, T* lr, double* bq) {
  const gd_params& P = a.prm;
  const T nx = fg.nx, ny = fg.ny, len = fg.len;
  const T g = (T)P.Gamma, Rg = (T)P.RGAS;
  T le[4];  // lamF(owner) - lamF(neighbour)
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const T d = li[v] - lj[v];
    le[v] = sd ? -d : d;
  }
  if (fg.on_wall) { le[1] += (T)a.fw[0]; le[2] += (T)a.fw[1]; }
#pragma unroll
  for (int v = 0; v < 4; ++v) le[v] *= len;
  // forward recompute of both face states, each from its own record and offset
  const T ox = fg.ox, oy = fg.oy;
  const T* gs = ri + 4;
  T qs[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) qs[v] = ri[v] + ri[12 + v] * (gs[2 * v] * sx + gs[2 * v + 1] * sy);
  const bool fs = !(qs[0] > 0) || !(qs[3] > 0);
  if (fs) {
#pragma unroll
    for (int v = 0; v < 4; ++v) qs[v] = ri[v];
  }
  const T* go = rj + 4;
  T qo[4];
  if (c1 >= 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) qo[v] = rj[v] + rj[12 + v] * (go[2 * v] * ox + go[2 * v + 1] * oy);
    if (!(qo[0] > 0) || !(qo[3] > 0)) {
#pragma unroll
      for (int v = 0; v < 4; ++v) qo[v] = rj[v];
    }
  } else {
    ns::ghost<T>(P, c1 + 1, qs, nx, ny, qo);
  }
  T qa[4], qb[4];  // owner's and neighbour's
#pragma unroll
  for (int v = 0; v < 4; ++v) { qa[v] = sd ? qo[v] : qs[v]; qb[v] = sd ? qs[v] : qo[v]; }
  T lqa[4] = {0, 0, 0, 0}, lqb[4] = {0, 0, 0, 0};
  if (ROE) roe_rev<T>(qb, qa, nx, ny, g, le, lqb, lqa);
  else ausm_rev<T>(qb, qa, nx, ny, g, le, lqb, lqa);
  const int bkind = c1 < 0 ? (-(c1 + 1) <= GD_MAX_BC ? P.bc[-(c1 + 1)].kind : GD_BC_EXTRAPOLATE) : -1;
  // written unconditionally so that bq stays in registers
#pragma unroll
  for (int v = 0; v < 4; ++v) bq[v] = bkind == GD_BC_FIXED ? (double)lqb[v] : 0.0;
  if (c1 < 0) {
    const int kind = bkind;
    if (kind == GD_BC_FIXED) {
    } else if (kind == GD_BC_SLIP_WALL || kind == GD_BC_ISOTHERMAL_WALL) {
      const T sn = lqb[1] * nx + lqb[2] * ny;
      lqa[0] += lqb[0];
      lqa[1] += lqb[1] - 2 * nx * sn;
      lqa[2] += lqb[2] - 2 * ny * sn;
      lqa[3] += lqb[3];
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) lqa[v] += lqb[v];
    }
  }
  T lqs[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) lqs[v] = sd ? lqb[v] : lqa[v];
  if (VISC) {
    T lgv[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) lgv[v] = le[v];
    const T cp = g * Rg / (g - 1);
    const T rRs = 1 / (ri[0] * Rg);
    const T Ts = ri[3] * rRs;
    if (c1 >= 0) {
      const T rRo = 1 / (rj[0] * Rg);
      const T To = rj[3] * rRo;
      T g_s[3][2], g_o[3][2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        g_s[0][c] = gs[2 + c]; g_s[1][c] = gs[4 + c];
        g_s[2][c] = (gs[6 + c] - Ts * Rg * gs[c]) * rRs;
        g_o[0][c] = go[2 + c]; g_o[1][c] = go[4 + c];
        g_o[2][c] = (go[6 + c] - To * Rg * go[c]) * rRo;
      }
      // centroid neighbour - centroid owner (an exact negation from the neighbour)
      const T ex0 = sx - ox, ey0 = sy - oy;
      const T ddx = sd ? -ex0 : ex0, ddy = sd ? -ey0 : ey0;
      const T rdl = 1 / sqrt(ddx * ddx + ddy * ddy);
      const T tx = ddx * rdl, ty = ddy * rdl;
      // neighbour minus owner of u, v, T
      const T d3[3] = {rj[1] - ri[1], rj[2] - ri[2], To - Ts};
      T dphi[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) dphi[c] = sd ? -d3[c] : d3[c];
      T gf[3][2];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T ax = (T)0.5 * (g_s[c][0] + g_o[c][0]), ay = (T)0.5 * (g_s[c][1] + g_o[c][1]);
        const T corr = dphi[c] * rdl - (ax * tx + ay * ty);
        gf[c][0] = ax + corr * tx; gf[c][1] = ay + corr * ty;
      }
      T luf, lvf, lTf, lgf[3][2];
      viscous_face_rev<T>(P, cp, (T)0.5 * (ri[1] + rj[1]), (T)0.5 * (ri[2] + rj[2]), (T)0.5 * (Ts + To), gf, nx, ny,
                          lgv, &luf, &lvf, &lTf, lgf);
      // both sides take half the averaged-gradient adjoint; the difference
      // term has opposite signs
      const T sg = sd ? (T)1 : (T)-1;
      T lg[3][2], ldphi[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T tl = tx * lgf[c][0] + ty * lgf[c][1];
        lg[c][0] = (T)0.5 * (lgf[c][0] - tx * tl);
        lg[c][1] = (T)0.5 * (lgf[c][1] - ty * tl);
        ldphi[c] = rdl * tl;
      }
      lr[1] += (T)0.5 * luf + sg * ldphi[0];
      lr[2] += (T)0.5 * lvf + sg * ldphi[1];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        lr[6 + c] += lg[0][c];
        lr[8 + c] += lg[1][c];
      }
      const T lgT[2] = {lg[2][0], lg[2][1]};
      temp_rev<T>(ri, Rg, (T)0.5 * lTf + sg * ldphi[2], lgT, lr);
    } else {
      const int idx = -(c1 + 1);
      if (idx <= GD_MAX_BC && P.bc[idx].kind == GD_BC_ISOTHERMAL_WALL) {
        const T Tw = (T)P.bc[idx].data[0];
        const T rdp = -1 / (sx * nx + sy * ny);
        const T dudn = (0 - ri[1]) * rdp, dvdn = (0 - ri[2]) * rdp, dTdn = (Tw - Ts) * rdp;
        const T ex = -nx, ey = -ny;
        const T gf[3][2] = {{dudn * ex, dudn * ey}, {dvdn * ex, dvdn * ey}, {dTdn * ex, dTdn * ey}};
        T luf, lvf, lTf, lgf[3][2];
        viscous_face_rev<T>(P, cp, (T)0, (T)0, Tw, gf, nx, ny, lgv, &luf, &lvf, &lTf, lgf);
        const T ldn[3] = {lgf[0][0] * ex + lgf[0][1] * ey, lgf[1][0] * ex + lgf[1][1] * ey,
                          lgf[2][0] * ex + lgf[2][1] * ey};
        lr[1] -= ldn[0] * rdp;
        lr[2] -= ldn[1] * rdp;
        const T lT = -ldn[2] * rdp;  // T_i = p rR
        lr[3] += lT * rRs;
        lr[0] -= lT * Ts / ri[0];
      }
    }
  }
  // reconstruction q_f = q + phi (grad q . r), first order where it fell back
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    lr[v] += lqs[v];
    if (!fs) {
      lr[4 + 2 * v] += lqs[v] * ri[12 + v] * sx;
      lr[5 + 2 * v] += lqs[v] * ri[12 + v] * sy;
      lr[12 + v] += lqs[v] * (gs[2 * v] * sx + gs[2 * v + 1] * sy);
    }
  }
}

// ---- cell pass, one thread per cell: every load of the cell's faces is
// issued before the first transposed flux, then the four variables' transposed
// gradient + limiter run in the same thread.
template <class T, bool QUADS, bool VISC, bool ROE>
__global__ __launch_bounds__(kBlock) void k_adj_cell(AdjKArgs<T> a) {
  constexpr int NF = QUADS ? 4 : 3;
  using G = FaceGeo<T>;
  const MeshView& m = a.m;
  const int64_t nc = m.nc;
  const int64_t i0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  // lanes past the end stand in for the last cell in the cooperative gather
  // and store nothing
  const bool live = i0 < nc;
  const int64_t i = live ? i0 : nc - 1;
  const int nf = QUADS ? m.nv[i] : 3;
  int32_t jn[NF], fc[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) { jn[k] = ns::ntl(&m.nb[k * nc + i]); fc[k] = ns::ntl(&m.cface[k * nc + i]); }
  // the cell's adjoint flux inputs and geometry, loaded before the record
  // gather so that they arrive during it (loads after its LDS fences would
  // wait on their own)
  T li[4], lj[NF][4];
  ld4(a.lamF + 4 * i, li);
  FaceIn<T> fg[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    const int64_t jj = jn[k] >= 0 ? jn[k] : i;
    ld4(a.lamF + 4 * jj, lj[k]);
    if (k < nf) fg[k] = load_face(a, k, i, fc[k] >> 1, jn[k] >= 0 ? 0 : jn[k]);
  }
  T wx[NF], wy[NF], rx[NF], ry[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    wx[k] = G::lsq(m, k, 0, i); wy[k] = G::lsq(m, k, 1, i);
    rx[k] = G::rf(m, k, 0, i); ry[k] = G::rf(m, k, 1, i);
  }
  const T kh = (T)a.prm.venkat_K * G::hmin(m, i);
  T lr[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) lr[t] = 0;
  double bqk[NF][4];
#pragma unroll
  for (int k = 0; k < NF; ++k)
#pragma unroll
    for (int v = 0; v < 4; ++v) bqk[k][v] = 0.0;
  // own and neighbour records (a boundary or absent slot reads the own): the
  // own record arrives first and is kept; each face is transposed as soon as
  // its neighbour's record has arrived, after which only that neighbour's
  // state is kept (faces in slot order, as the sums require)
  __shared__ typename ns::RecGather<T, NF + 1>::Shared sh;
  T ri[16], qk[NF][4];
  {
    int32_t cells[NF + 1];
    cells[0] = (int32_t)i;
#pragma unroll
    for (int k = 0; k < NF; ++k) cells[1 + k] = jn[k] >= 0 ? jn[k] : (int32_t)i;
    ns::RecGather<T, NF + 1>::gather_each(sh, a.rec, cells, [&](int r, const T (&rec)[16]) {
      if (r == 0) {
#pragma unroll
        for (int t = 0; t < 16; ++t) ri[t] = rec[t];
        return;
      }
      const int k = r - 1;
#pragma unroll
      for (int v = 0; v < 4; ++v) qk[k][v] = rec[v];
      if (live && k < nf) {
        const int32_t c1 = jn[k] >= 0 ? 0 : jn[k];
        if (c1 < 0) {
#pragma unroll
          for (int v = 0; v < 4; ++v) lj[k][v] = 0;
        }
        face_rev<T, VISC, ROE>(a, fg[k], rx[k], ry[k], c1, (fc[k] & 1) != 0, ri, rec, li, lj[k], lr, bqk[k]);
      }
    });
  }
  if (!live) return;
  const T* qi = ri;
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    if (k < nf && jn[k] < 0) {
      const int64_t s = k * nc + i;
      ns::ghost<T>(a.prm, jn[k] + 1, qi, G::nx(m, s), G::ny(m, s), qk[k]);
    }
  }
  T lq[NF + 1][4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    lq[0][v] = lr[v];
#pragma unroll
    for (int k = 0; k < NF; ++k) lq[1 + k][v] = 0;
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const T gx = ri[4 + 2 * v], gy = ri[5 + 2 * v];
    T lgx = lr[4 + 2 * v], lgy = lr[5 + 2 * v];
    const T lph = lr[12 + v];
    T qmx = qi[v], qmn = qi[v];
    int amx = 0, amn = 0;
#pragma unroll
    for (int k = 0; k < NF; ++k)
      if (k < nf) {
        if (qk[k][v] > qmx) { qmx = qk[k][v]; amx = 1 + k; }
        if (qk[k][v] < qmn) { qmn = qk[k][v]; amn = 1 + k; }
      }
    const T e2 = kh * kh * kh * (T)(a.prm.venkat_scale[v] * a.prm.venkat_scale[v]);
    T phi = 1;
    int ks = -1;
#pragma unroll
    for (int k = 0; k < NF; ++k)
      if (k < nf) {
        const T d2 = gx * rx[k] + gy * ry[k];
        if (d2 != 0) {
          const T d1 = d2 > 0 ? qmx - qi[v] : qmn - qi[v];
          const T psi = ((d1 * d1 + e2) * d2 + 2 * d2 * d2 * d1) / (d1 * d1 + 2 * d2 * d2 + d1 * d2 + e2) / d2;
          if (psi < phi) { phi = psi; ks = k; }
        }
      }
    if (ks >= 0 && !(phi < 0)) {
      T rxs = 0, rys = 0;
#pragma unroll
      for (int k = 0; k < NF; ++k)
        if (k == ks) { rxs = rx[k]; rys = ry[k]; }
      const T d2 = gx * rxs + gy * rys;
      const int ae = d2 > 0 ? amx : amn;
      const T d1 = (d2 > 0 ? qmx : qmn) - qi[v];
      const T D = d1 * d1 + 2 * d2 * d2 + d1 * d2 + e2;
      const T dd1 = ((2 * d1 + 2 * d2) - phi * (2 * d1 + d2)) / D;
      const T dd2 = (2 * d1 - phi * (4 * d2 + d1)) / D;
      const T ld1 = lph * dd1, ld2 = lph * dd2;
      lgx += ld2 * rxs;
      lgy += ld2 * rys;
#pragma unroll
      for (int p = 0; p <= NF; ++p)
        if (p == ae) lq[p][v] += ld1;
      lq[0][v] -= ld1;
    }
#pragma unroll
    for (int k = 0; k < NF; ++k)
      if (k < nf) {
        const T t = wx[k] * lgx + wy[k] * lgy;
        lq[1 + k][v] += t;
        lq[0][v] -= t;
      }
  }
  const int32_t nfs = m.nf;
  const T gm1 = (T)(a.prm.Gamma - 1.0);
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    if (k >= nf) continue;
    if (jn[k] >= 0) {
      T o[4];
      prims_rev<T>(qk[k], gm1, lq[1 + k], o);  // in the neighbour's conservative variables
      st4(a.nrev + 4 * (i * nfs + k), o[0], o[1], o[2], o[3]);
    } else {
      const int idx = -(jn[k] + 1);
      const int kind = idx <= GD_MAX_BC ? a.prm.bc[idx].kind : GD_BC_EXTRAPOLATE;
      const int64_t s = k * nc + i;
      if (kind == GD_BC_FIXED) {
        const int32_t wid = m.wall_id[s];
        if (wid >= 0) {
#pragma unroll
          for (int v = 0; v < 4; ++v) a.bcsum[4 * (int64_t)wid + v] += bqk[k][v] + (double)lq[1 + k][v];
        }
      } else if (kind == GD_BC_SLIP_WALL || kind == GD_BC_ISOTHERMAL_WALL) {
        const T nx = G::nx(m, s), ny = G::ny(m, s);
        const T sn = lq[1 + k][1] * nx + lq[1 + k][2] * ny;
        lq[0][0] += lq[1 + k][0];
        lq[0][1] += lq[1 + k][1] - 2 * nx * sn;
        lq[0][2] += lq[1 + k][2] - 2 * ny * sn;
        lq[0][3] += lq[1 + k][3];
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) lq[0][v] += lq[1 + k][v];
      }
    }
  }
  T o[4];
  prims_rev<T>(qi, gm1, lq[0], o);
  st4(a.selfq + 4 * i, o[0], o[1], o[2], o[3]);
}

// ---- state pass: gather the primitive adjoint, transpose the primitive
// recovery, add to lamU; seed the next step's lamF.  The slots are unrolled
// (NF = 3 or 4), so that every slot's ids and then every neighbour part are
// loaded together rather than one dependent pair per slot.
template <class T, bool GATHER, bool QUADS>
__global__ __launch_bounds__(kBlock) void k_adj_state(AdjKArgs<T> a) {
  constexpr int NF = QUADS ? 4 : 3;
  const MeshView& m = a.m;
  const int64_t nc = m.nc;
  const int64_t j = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (j >= nc) return;
  double lU[4];
#pragma unroll
  // lamU, the ids and 1/A are read once per reverse step and lamU is
  // written for the next one only: non-temporal both ways, leaving the caches
  // to the parts and lamF that the cell pass gathers (200-step adjoint
  // 86.4 -> 83.8 ms on one box)
  for (int v = 0; v < 4; ++v) lU[v] = ns::ntl(&a.lamU_in[v * nc + j]);
  if (GATHER) {
    const int nf = QUADS ? m.nv[j] : 3;
    int32_t jn[NF], kk[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      jn[k] = -1; kk[k] = 0;
      if (k < nf) { jn[k] = ns::ntl(&m.nb[k * nc + j]); kk[k] = ns::ntl(&m.kn[k * nc + j]); }
    }
    T s4[4], b[NF][4];
    ld4(a.selfq + 4 * j, s4);
#pragma unroll
    for (int k = 0; k < NF; ++k)
      if (jn[k] >= 0) ld4(a.nrev + 4 * ((int64_t)jn[k] * m.nf + kk[k]), b[k]);
#pragma unroll
    for (int v = 0; v < 4; ++v) lU[v] += (double)s4[v];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      if (jn[k] >= 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) lU[v] += (double)b[k][v];
      }
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) __builtin_nontemporal_store(lU[w], &a.lamU_out[w * nc + j]);
  }
  const double s = -a.dt * ns::ntl(&m.rarea[j]);
  st4(a.lamF + 4 * j, (T)(s * lU[0]), (T)(s * lU[1]), (T)(s * lU[2]), (T)(s * lU[3]));
}

template <class T, bool ROE>
void launch_cell_flux(const AdjKArgs<T>& a, hipStream_t s) {
  const dim3 g(ns::nblocks(a.m.nc)), b(kBlock);
  const bool visc = a.prm.viscous != 0;
  if (a.m.has_quads) {
    if (visc) hipLaunchKernelGGL((k_adj_cell<T, true, true, ROE>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_adj_cell<T, true, false, ROE>), g, b, 0, s, a);
  } else {
    if (visc) hipLaunchKernelGGL((k_adj_cell<T, false, true, ROE>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_adj_cell<T, false, false, ROE>), g, b, 0, s, a);
  }
}

template <class T>
hipError_t launch_cell(const AdjKArgs<T>& a, hipStream_t s) {
  if (a.prm.flux == GD_FLUX_ROE) launch_cell_flux<T, true>(a, s);
  else launch_cell_flux<T, false>(a, s);
  return hipGetLastError();
}

template <class T>
hipError_t launch_state(const AdjKArgs<T>& a, bool gather, hipStream_t s) {
  const dim3 g(ns::nblocks(a.m.nc)), b(kBlock);
  if (!gather) hipLaunchKernelGGL((k_adj_state<T, false, false>), g, b, 0, s, a);
  else if (a.m.has_quads) hipLaunchKernelGGL((k_adj_state<T, true, true>), g, b, 0, s, a);
  else hipLaunchKernelGGL((k_adj_state<T, true, false>), g, b, 0, s, a);
  return hipGetLastError();
}

}  // namespace adj
}  // namespace gd
