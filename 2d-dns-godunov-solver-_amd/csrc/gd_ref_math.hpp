// Device arithmetic of the reference scheme (FP64).
//
// This file is compiled with -ffp-contract=off: every a*b+c below is a
// multiply and an add, in the order the reference writes it, so the FP64
// kernels reproduce the reference arithmetic; fma() appears only where the
// reference itself calls muladd.  Division and sqrt are IEEE correctly rounded
// (HIP's default lowering for double).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/godunov.h"

namespace gd {
namespace refm {

// Julia's sign(x): +-1, or x itself for +-0 and NaN
__device__ __forceinline__ double jsign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }
// Julia's min/max propagate NaN (fmin/fmax would drop it)
__device__ __forceinline__ double jmin(double x, double y) {
  return (x != x || y != y) ? x + y : (y < x ? y : x);
}
__device__ __forceinline__ double jmax(double x, double y) {
  return (x != x || y != y) ? x + y : (y > x ? y : x);
}

// The limiter, splittings and AUSM+ below are templates on the scalar: with
// double they are the forward kernels' arithmetic; the reference-scheme
// adjoint instantiates the limiter with the dual number of gd_ad.hpp (found by
// argument-dependent lookup: fabs, jsign, jmin), which also carries its two
// partial derivatives.

// limiters.jl:33-56 and the commented true minmod (limiters.jl:45-47)
template <class T>
__device__ __forceinline__ T limiter(int kind, T x, T y, double c) {
  if (kind == GD_LIM_VAN_LEER_A) return (x * y + fabs(x * y)) / (x + y + c);
  if (kind == GD_LIM_VAN_ALBADA_A) return (x * (y * y + c) + y * (x * x + c)) / (x * x + y * y + 2.0 * c);
  if (kind == GD_LIM_MINMOD) return 0.5 * (jsign(x) + jsign(y)) * jmin(fabs(x), fabs(y));
  // Minmod_Limiter: (sign x + sign y) |x y| / (|x| + |y| + c), a harmonic mean
  return (jsign(x) + jsign(y)) * fabs(x * y) / (fabs(x) + fabs(y) + c);
}

// AUSM+ Mach and pressure splittings (Liou 1996, beta = 1/8, alpha = 3/16,
// 0.25 in M_beta): AUSMflux2dFast.jl:23-53
template <class T>
__device__ __forceinline__ T mach_plus(T M) {
  const T d = M * M - 1.0;
  return fabs(M) >= 1.0 ? 0.5 * (M + fabs(M)) : 0.25 * (M + 1.0) * (M + 1.0) + (1.0 / 8.0) * d * d;
}
template <class T>
__device__ __forceinline__ T mach_minus(T M) {
  const T d = M * M - 1.0;
  return fabs(M) >= 1.0 ? 0.5 * (M - fabs(M)) : -0.25 * (M - 1.0) * (M - 1.0) - (1.0 / 8.0) * d * d;
}
template <class T>
__device__ __forceinline__ T pres_plus(T M) {
  const T d = M * M - 1.0;
  return fabs(M) >= 1.0 ? 0.5 * (1.0 + jsign(M)) : 0.25 * (M + 1.0) * (M + 1.0) * (2.0 - M) + (3.0 / 16.0) * M * d * d;
}
template <class T>
__device__ __forceinline__ T pres_minus(T M) {
  const T d = M * M - 1.0;
  return fabs(M) >= 1.0 ? 0.5 * (1.0 - jsign(M)) : 0.25 * (M - 1.0) * (M - 1.0) * (2.0 + M) - (3.0 / 16.0) * M * d * d;
}

// AUSM+ face flux (AUSMflux2dFast.jl:87-175).  "L" is the state upstream of
// the normal; the result is -(F . n) * len.
template <class T>
__device__ __forceinline__ void ausm_plus(const T* L, const T* R, double nx, double ny, double len, double g, T* out) {
  const T rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const T rR = R[0], uR = R[1], vR = R[2], pR = R[3];
  const T htL = fma(0.5, (uL * uL + vL * vL), pL / rL + pL / rL / (g - 1.0));
  const T htR = fma(0.5, (uR * uR + vR * vR), pR / rR + pR / rR / (g - 1.0));
  const double cst = 2.0 * (g - 1.0) / (g + 1.0);
  const T unL = uL * nx + vL * ny;
  const T unR = uR * nx + vR * ny;
  const T asL = sqrt(cst * htL);
  const T asR = sqrt(cst * htR);
  const T aL = asL * jmin(1.0, asL / fabs(unL));
  const T aR = asR * jmin(1.0, asR / fabs(unR));
  const T a12 = jmin(aL, aR);
  const T ML = unL / a12, MR = unR / a12;
  const T m = mach_plus(ML) + mach_minus(MR);
  const T p12 = pres_plus(ML) * pL + pres_minus(MR) * pR;
  const T mp = 0.5 * (m + fabs(m));
  const T mm = 0.5 * (m - fabs(m));
  out[0] = -(a12 * (mp * rL + mm * rR) + 0.0) * len;
  out[1] = -(a12 * (mp * rL * uL + mm * rR * uR) + p12 * nx) * len;
  out[2] = -(a12 * (mp * rL * vL + mm * rR * vR) + p12 * ny) * len;
  out[3] = -(a12 * (mp * rL * htL + mm * rR * htR) + 0.0) * len;
}

// AUSM+M (Chen et al. 2020): AUSMflux2dFast.jl:234-284
__device__ __forceinline__ void ausm_plus_m2020(const double* L, const double* R, double nx, double ny, double len,
                                                double g, double* out) {
  const double rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const double rR = R[0], uR = R[1], vR = R[2], pR = R[3];
  const double htL = fma(0.5, (uL * uL + vL * vL), pL / rL + pL / rL / (g - 1.0));
  const double htR = fma(0.5, (uR * uR + vR * vR), pR / rR + pR / rR / (g - 1.0));
  const double tL = uL * ny - vL * nx, tR = uR * ny - vR * nx;
  const double htN = 0.5 * (htL + htR - 0.5 * (tL * tL + tR * tR));
  const double as = sqrt(2.0 * (g - 1.0) / (g + 1.0) * htN);
  const double unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
  const double a12 = (0.5 * (uL * nx + vL * ny + uR * nx + vR * ny) >= 0) ? as * as / jmax(fabs(unL), as)
                                                                         : as * as / jmax(fabs(unR), as);
  const double f = 0.5 * (1.0 - cos(M_PI * jmin(1.0, jmax(fabs(unL / a12), fabs(unR / a12)))));
  const double gg = 0.5 * (1.0 + cos(M_PI * jmin(pL / pR, pR / pL)));
  const double p12 = -gg * g * (pL + pR) * 0.5 / a12 * (pres_plus(unL / a12) * pres_minus(unR / a12));
  const double pux = p12 * (uR - uL);
  const double puy = p12 * (vR - vL);
  const double ps = 0.5 * (pL + pR) - 0.5 * 0.5 * (rL + rR) * a12 * (uR - uL);
  const double M12 = mach_plus(unL / a12) + mach_minus(unR / a12) -
                     0.5 * (1.0 - f) * (pR - pL) * 2.0 / (rL + rR) / a12 / a12 * (1.0 - gg);
  const double md = M12 >= 0.0 ? M12 * a12 * rL : M12 * a12 * rR;
  const double mp = 0.5 * (md + fabs(md)), mm = 0.5 * (md - fabs(md));
  out[0] = -((0.5 * (md + fabs(md) * 1.0) + mm * 1.0) + 0.0) * len;
  out[1] = -((mp * uL + mm * uR) + ps * nx + pux) * len;
  out[2] = -((mp * vL + mm * vR) + ps * ny + puy) * len;
  out[3] = -((mp * htL + mm * htR) + 0.0) * len;
}

// Roe flux with Harten's entropy fix on the acoustic waves (absolute width
// 0.2): RoeFlux2dFast.jl:1-198
__device__ __forceinline__ void roe(const double* L, const double* R, double nx, double ny, double len, double g,
                                    double* out) {
  const double rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const double rR = R[0], uR = R[1], vR = R[2], pR = R[3];
  const double mx = -ny, my = nx;
  const double unL = uL * nx + vL * ny, umL = uL * mx + vL * my;
  const double aL = sqrt(g * pL / rL);
  const double HL = aL * aL / (g - 1.0) + 0.5 * (uL * uL + vL * vL);
  const double unR = uR * nx + vR * ny, umR = uR * mx + vR * my;
  const double aR = sqrt(g * pR / rR);
  const double HR = aR * aR / (g - 1.0) + 0.5 * (uR * uR + vR * vR);
  const double RT = sqrt(rR / rL);
  const double rho = RT * rL;
  const double u = (uL + RT * uR) / (1.0 + RT);
  const double v = (vL + RT * vR) / (1.0 + RT);
  const double H = (HL + RT * HR) / (1.0 + RT);
  const double a = sqrt((g - 1.0) * (H - 0.5 * (u * u + v * v)));
  const double un = u * nx + v * ny, um = u * mx + v * my;
  const double dp = pR - pL, dun = unR - unL, dum = umR - umL, drho = rR - rL;
  const double w1 = (dp - rho * a * dun) / (2.0 * a * a);
  const double w2 = rho * dum;
  const double w3 = drho - dp / (a * a);
  const double w4 = (dp + rho * a * dun) / (2.0 * a * a);
  double s1 = fabs(un - a);
  const double s2 = fabs(un);
  double s4 = fabs(un + a);
  const double dw = 1.0 / 5.0;
  if (s1 < dw) s1 = 0.5 * (s1 * s1 / dw + dw);
  if (s4 < dw) s4 = 0.5 * (s4 * s4 / dw + dw);
  // dissipation sum_k |lambda_k| w_k r_k, accumulated wave by wave
  double d1 = 0.0, d2 = 0.0, d3 = 0.0, d4 = 0.0;
  d1 = d1 + s1 * w1 * 1.0;            d2 = d2 + s1 * w1 * (u - a * nx);
  d3 = d3 + s1 * w1 * (v - a * ny);   d4 = d4 + s1 * w1 * (H - un * a);
  d1 = d1 + s2 * w2 * 0.0;            d2 = d2 + s2 * w2 * mx;
  d3 = d3 + s2 * w2 * my;             d4 = d4 + s2 * w2 * um;
  d1 = d1 + s2 * w3 * 1.0;            d2 = d2 + s2 * w3 * u;
  d3 = d3 + s2 * w3 * v;              d4 = d4 + s2 * w3 * (0.5 * (u * u + v * v));
  d1 = d1 + s4 * w4 * 1.0;            d2 = d2 + s4 * w4 * (u + a * nx);
  d3 = d3 + s4 * w4 * (v + a * ny);   d4 = d4 + s4 * w4 * (H + un * a);
  out[0] = -0.5 * (rL * unL + rR * unR - d1) * len;
  out[1] = -0.5 * ((rL * unL * uL + pL * nx) + (rR * unR * uR + pR * nx) - d2) * len;
  out[2] = -0.5 * ((rL * unL * vL + pL * ny) + (rR * unR * vR + pR * ny) - d3) * len;
  out[3] = -0.5 * (rL * unL * HL + rR * unR * HR - d4) * len;
}

__device__ __forceinline__ void face_flux(int kind, const double* L, const double* R, double nx, double ny,
                                          double len, double g, double* out) {
  if (kind == GD_FLUX_ROE) roe(L, R, nx, ny, len, g, out);
  else if (kind == GD_FLUX_AUSM_PLUS_M2020) ausm_plus_m2020(L, R, nx, ny, len, g, out);
  else ausm_plus(L, R, nx, ny, len, g, out);
}

// air conductivity k(T) of thermo.jl:13-60 (power-law mu, piecewise cp)
__device__ __forceinline__ double air_cp_poly(double T, bool high) {
  const double i = 1.0 / T;
  if (!high)
    return 2898903.0 * (i * i) + -56496.26 * i + 1437.799 + -1.653609 * T + 0.003062254 * (T * T) +
           -2.279138e-6 * (T * T * T) + 6.272365e-10 * pow(T, 4.0);
  return 6.932494e+7 * (i * i) + -361053.2 * i + 1476.665 + -0.06138349 * T + 2.027963e-5 * (T * T) +
         -3.075525e-9 * (T * T * T) + 1.888054e-13 * pow(T, 4.0);
}
__device__ __forceinline__ double air_k(double T) {
  double cp;
  if (T < 200.0) cp = air_cp_poly(200.0, false);
  else if (T > 200.0 && T < 1000.0) cp = air_cp_poly(T, false);
  else if (T >= 1000.0 && T < 6000.0) cp = air_cp_poly(T, true);
  else cp = air_cp_poly(6000.0, true);
  const double mu = 1.716e-5 * pow(T / 273.11, 2.0 / 3.0);
  return 15.0 / 4.0 * 8.314462 / 28.9647 * mu * (4.0 / 15.0 * cp * 28.9647 / 8.314462 + 1.0 / 3.0);
}

// Ghost primitive state for a boundary face.
//   fixed / zero-gradient / slip wall: boundaryConditions2d.jl:29-57,
//     bcInviscidWall.jl:109-126;
//   TGV mirror: analytic TGV (initfields_TGV2d.jl:9-27) at the centroid
//     mirrored across the edge (calcGradients.jl:21-34),
//     boundaryConditions_TGV2d.jl:16-30;
//   isothermal no-slip wall: u_g = -u, T_g = 2 T_w - T (floored at T_w/2), p_g = p.
__device__ __forceinline__ void bc_ghost(const gd_bc* table, int code, const double* q, double nx, double ny,
                                         double xc, double yc, double xa, double ya, double xb, double yb, double t,
                                         double R, double* g) {
  const int idx = -code;
  const gd_bc b = (idx >= 0 && idx <= GD_MAX_BC) ? table[idx] : gd_bc{GD_BC_EXTRAPOLATE, 0, {0, 0, 0, 0}};
  if (b.kind == GD_BC_FIXED) {
    g[0] = b.data[0]; g[1] = b.data[1]; g[2] = b.data[2]; g[3] = b.data[3];
  } else if (b.kind == GD_BC_SLIP_WALL) {
    g[0] = q[0];
    g[1] = q[1] * (ny * ny - nx * nx) - 2.0 * nx * ny * q[2];
    g[2] = q[2] * (nx * nx - ny * ny) - 2.0 * nx * ny * q[1];
    g[3] = q[3];
  } else if (b.kind == GD_BC_TGV_MIRROR) {
    const double dx = xb - xa, dy = yb - ya;
    const double a = (dx * dx - dy * dy) / (dx * dx + dy * dy);
    const double bb = 2.0 * dx * dy / (dx * dx + dy * dy);
    const double sx = a * (xc - xa) + bb * (yc - ya) + xa;
    const double sy = bb * (xc - xa) - a * (yc - ya) + ya;
    const double P = b.data[0], U = b.data[1], L = b.data[2], mu = b.data[3];
    const double T = P / R;
    const double F = exp(-2.0 * mu * t / L / L);
    const double Fp = exp(-4.0 * mu * t / L / L);
    g[0] = P / R / T;
    g[1] = U * sin(sx / L) * cos(sy / L) * F;
    g[2] = -U * cos(sx / L) * sin(sy / L) * F;
    g[3] = P + 0.25 * U * U * (cos(2 * sx / L) + cos(2 * sy / L)) * Fp;
  } else if (b.kind == GD_BC_ISOTHERMAL_WALL) {
    const double Tc = q[3] / (q[0] * R);
    double Tg = 2.0 * b.data[0] - Tc;
    if (Tg < 0.5 * b.data[0]) Tg = 0.5 * b.data[0];
    g[0] = q[3] / (R * Tg); g[1] = -q[1]; g[2] = -q[2]; g[3] = q[3];
  } else {
    g[0] = q[0]; g[1] = q[1]; g[2] = q[2]; g[3] = q[3];
  }
}

}  // namespace refm
}  // namespace gd
