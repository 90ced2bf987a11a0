// CPU oracle for the explicit Godunov step: TEST INFRASTRUCTURE ONLY.
//
// A scalar restatement of the reference's Julia step (reference mode) used as
// the checker by tests, the smoke check and the bench's CPU-baseline leg.
// The product library never links it.  Built with -ffp-contract=off so that
// every a*b+c is two roundings, as in the Julia source, except where the
// reference itself writes muladd (then fma is used).
//
// Layout: Julia column-major (nCells, k) arrays, 1-based ids, as in
// mesh2d_Int32 (primeObjects.jl:109-135) and fields2d (primeObjects.jl:85-106).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// 1-based column-major accessors
inline int32_t I(const int32_t* a, int64_t n, int64_t i, int k) { return a[(int64_t)(k - 1) * n + (i - 1)]; }
inline double D(const double* a, int64_t n, int64_t i, int k) { return a[(int64_t)(k - 1) * n + (i - 1)]; }

// Julia sign(x) for Float64: +-1, or x itself for +-0 (and NaN)
inline double jsign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }
// Julia min(x, y) for Float64 propagates NaN
inline double jmin(double x, double y) {
  if (std::isnan(x) || std::isnan(y)) return x + y;
  return y < x ? y : x;
}
inline double jmax(double x, double y) {
  if (std::isnan(x) || std::isnan(y)) return x + y;
  return y > x ? y : x;
}

// ---------------------------------------------------------------- limiters.jl
// Minmod_Limiter, limiters.jl:33-35 (a harmonic/van Leer mean, despite the name)
inline double minmod_limiter_ref(double x, double y, double c) {
  return (jsign(x) + jsign(y)) * std::fabs(x * y) / (std::fabs(x) + std::fabs(y) + c);
}
// vanLeer_LimiterA, limiters.jl:41-43
inline double van_leer_a(double a, double b, double c) { return (a * b + std::fabs(a * b)) / (a + b + c); }
// vanAlbada_LimiterA, limiters.jl:54-56
inline double van_albada_a(double a, double b, double c) {
  return (a * (b * b + c) + b * (a * a + c)) / (a * a + b * b + 2.0 * c);
}
// true minmod, limiters.jl:45-47 (commented there as Minmod_LimiterA)
inline double minmod_true(double x, double y, double) {
  return 0.5 * (jsign(x) + jsign(y)) * jmin(std::fabs(x), std::fabs(y));
}
inline double limiter(int kind, double x, double y, double c) {
  switch (kind) {
    case ORC_LIM_VAN_LEER_A: return van_leer_a(x, y, c);
    case ORC_LIM_VAN_ALBADA_A: return van_albada_a(x, y, c);
    case ORC_LIM_MINMOD: return minmod_true(x, y, c);
    default: return minmod_limiter_ref(x, y, c);
  }
}

// ----------------------------------------------------------- AUSMflux2dFast.jl
// splittings, AUSMflux2dFast.jl:23-53
inline double Palfa_m(double M, double A) { return 0.25 * (M - 1.0) * (M - 1.0) * (2.0 + M) - A * M * (M * M - 1.0) * (M * M - 1.0); }
inline double Palfa_p(double M, double A) { return 0.25 * (M + 1.0) * (M + 1.0) * (2.0 - M) + A * M * (M * M - 1.0) * (M * M - 1.0); }
inline double Mbetta_p(double M, double B) { return 0.25 * (M + 1.0) * (M + 1.0) + B * (M * M - 1.0) * (M * M - 1.0); }
inline double Mbetta_m(double M, double B) { return -0.25 * (M - 1.0) * (M - 1.0) - B * (M * M - 1.0) * (M * M - 1.0); }
inline double P_m(double M, double A) { return (std::fabs(M) >= 1.0) ? 0.5 * (1.0 - jsign(M)) : Palfa_m(M, A); }
inline double P_p(double M, double A) { return (std::fabs(M) >= 1.0) ? 0.5 * (1.0 + jsign(M)) : Palfa_p(M, A); }
inline double M_p(double M, double B) { return (std::fabs(M) >= 1.0) ? 0.5 * (M + std::fabs(M)) : Mbetta_p(M, B); }
inline double M_m(double M, double B) { return (std::fabs(M) >= 1.0) ? 0.5 * (M - std::fabs(M)) : Mbetta_m(M, B); }

// AUSMplusFlux2dFast, AUSMflux2dFast.jl:87-175
void ausm_plus(double rhoL, double UL, double VL, double PL, double rhoR, double UR, double VR, double PR, double nx,
               double ny, double side, double gamma, double* flux) {
  // muladd in the source (:118-119) -> fused multiply-add
  const double htL = std::fma(0.5, (UL * UL + VL * VL), PL / rhoL + PL / rhoL / (gamma - 1.0));
  const double htR = std::fma(0.5, (UR * UR + VR * VR), PR / rhoR + PR / rhoR / (gamma - 1.0));
  const double aL_tilda = std::sqrt(2.0 * (gamma - 1.0) / (gamma + 1.0) * htL) *
                          jmin(1.0, (std::sqrt(2.0 * (gamma - 1.0) / (gamma + 1.0) * htL)) / std::fabs(UL * nx + VL * ny));
  const double aR_tilda = std::sqrt(2.0 * (gamma - 1.0) / (gamma + 1.0) * htR) *
                          jmin(1.0, (std::sqrt(2.0 * (gamma - 1.0) / (gamma + 1.0) * htR)) / std::fabs(UR * nx + VR * ny));
  const double a12 = jmin(aL_tilda, aR_tilda);
  const double m_dot12 = M_p((UL * nx + VL * ny) / a12, 1.0 / 8.0) + M_m((UR * nx + VR * ny) / a12, 1.0 / 8.0);
  const double p12 = P_p((UL * nx + VL * ny) / a12, 3.0 / 16.0) * PL + P_m((UR * nx + VR * ny) / a12, 3.0 / 16.0) * PR;
  flux[0] = -(a12 * (0.5 * (m_dot12 + std::fabs(m_dot12)) * rhoL + 0.5 * (m_dot12 - std::fabs(m_dot12)) * rhoR) + 0.0) * side;
  flux[1] = -(a12 * (0.5 * (m_dot12 + std::fabs(m_dot12)) * rhoL * UL + 0.5 * (m_dot12 - std::fabs(m_dot12)) * rhoR * UR) + p12 * nx) * side;
  flux[2] = -(a12 * (0.5 * (m_dot12 + std::fabs(m_dot12)) * rhoL * VL + 0.5 * (m_dot12 - std::fabs(m_dot12)) * rhoR * VR) + p12 * ny) * side;
  flux[3] = -(a12 * (0.5 * (m_dot12 + std::fabs(m_dot12)) * rhoL * htL + 0.5 * (m_dot12 - std::fabs(m_dot12)) * rhoR * htR) + 0.0) * side;
}

// AUSMplusM2020Flux2dFast, AUSMflux2dFast.jl:234-284
void ausm_plus_m2020(double rhoL, double UL, double VL, double PL, double rhoR, double UR, double VR, double PR,
                     double nx, double ny, double side, double gamma, double* flux) {
  const double htL = std::fma(0.5, (UL * UL + VL * VL), PL / rhoL + PL / rhoL / (gamma - 1.0));
  const double htR = std::fma(0.5, (UR * UR + VR * VR), PR / rhoR + PR / rhoR / (gamma - 1.0));
  const double htN = 0.5 * (htL + htR - 0.5 * ((UL * ny - VL * nx) * (UL * ny - VL * nx) + (UR * ny - VR * nx) * (UR * ny - VR * nx)));
  const double as = std::sqrt(2.0 * (gamma - 1.0) / (gamma + 1.0) * htN);
  double a12 = 0.0;
  if (0.5 * (UL * nx + VL * ny + UR * nx + VR * ny) >= 0)
    a12 = as * as / jmax(std::fabs(UL * nx + VL * ny), as);
  else
    a12 = as * as / jmax(std::fabs(UR * nx + VR * ny), as);
  const double f = 0.5 * (1.0 - std::cos(M_PI * jmin(1.0, jmax(std::fabs((UL * nx + VL * ny) / a12), std::fabs((UR * nx + VR * ny) / a12)))));
  const double g = 0.5 * (1.0 + std::cos(M_PI * jmin(PL / PR, PR / PL)));
  const double p12 = -g * gamma * (PL + PR) * 0.5 / a12 * (P_p((UL * nx + VL * ny) / a12, 3.0 / 16.0) * P_m((UR * nx + VR * ny) / a12, 3.0 / 16.0));
  const double pux = p12 * (UR - UL);
  const double puy = p12 * (VR - VL);
  const double ps = 0.5 * (PL + PR) - 0.5 * 0.5 * (rhoL + rhoR) * a12 * (UR - UL);
  const double M12 = M_p((UL * nx + VL * ny) / a12, 1.0 / 8.0) + M_m((UR * nx + VR * ny) / a12, 1.0 / 8.0) -
                     0.5 * (1.0 - f) * (PR - PL) * 2.0 / (rhoL + rhoR) / a12 / a12 * (1.0 - g);
  double m_dot12 = 0.0;
  if (M12 >= 0.0) m_dot12 = M12 * a12 * rhoL; else m_dot12 = M12 * a12 * rhoR;
  flux[0] = -((0.5 * (m_dot12 + std::fabs(m_dot12) * 1.0) + 0.5 * (m_dot12 - std::fabs(m_dot12)) * 1.0) + 0.0) * side;
  flux[1] = -((0.5 * (m_dot12 + std::fabs(m_dot12)) * UL + 0.5 * (m_dot12 - std::fabs(m_dot12)) * UR) + ps * nx + pux) * side;
  flux[2] = -((0.5 * (m_dot12 + std::fabs(m_dot12)) * VL + 0.5 * (m_dot12 - std::fabs(m_dot12)) * VR) + ps * ny + puy) * side;
  flux[3] = -((0.5 * (m_dot12 + std::fabs(m_dot12)) * htL + 0.5 * (m_dot12 - std::fabs(m_dot12)) * htR) + 0.0) * side;
}

// ------------------------------------------------------------ RoeFlux2dFast.jl
// RoeFlux2dFast, RoeFlux2dFast.jl:1-198
void roe(double rhoL, double uL, double vL, double pL, double rhoR, double uR, double vR, double pR, double nx,
         double ny, double side, double gammaV, double* flux) {
  const double mx = -ny;
  const double my = nx;
  const double unL = uL * nx + vL * ny;
  const double umL = uL * mx + vL * my;
  const double aL = std::sqrt(gammaV * pL / rhoL);
  const double HL = aL * aL / (gammaV - 1.0) + 0.5 * (uL * uL + vL * vL);
  const double unR = uR * nx + vR * ny;
  const double umR = uR * mx + vR * my;
  const double aR = std::sqrt(gammaV * pR / rhoR);
  const double HR = aR * aR / (gammaV - 1.0) + 0.5 * (uR * uR + vR * vR);
  const double RT = std::sqrt(rhoR / rhoL);
  const double rho = RT * rhoL;
  const double u = (uL + RT * uR) / (1.0 + RT);
  const double v = (vL + RT * vR) / (1.0 + RT);
  const double H = (HL + RT * HR) / (1.0 + RT);
  const double a = std::sqrt((gammaV - 1.0) * (H - 0.5 * (u * u + v * v)));
  const double un = u * nx + v * ny;
  const double um = u * mx + v * my;
  const double drho = rhoR - rhoL;
  const double dp = pR - pL;
  const double dun = unR - unL;
  const double dum = umR - umL;
  const double LdU1 = (dp - rho * a * dun) / (2.0 * a * a);
  const double LdU2 = rho * dum;
  const double LdU3 = drho - dp / (a * a);
  const double LdU4 = (dp + rho * a * dun) / (2.0 * a * a);
  double ws1 = std::fabs(un - a);
  const double ws2 = std::fabs(un);
  const double ws3 = std::fabs(un);
  double ws4 = std::fabs(un + a);
  const double dws1 = 1.0 / 5.0;  // Harten fix, absolute width (:85-96)
  if (ws1 < dws1) ws1 = 0.5 * (ws1 * ws1 / dws1 + dws1);
  const double dws4 = 1.0 / 5.0;
  if (ws4 < dws4) ws4 = 0.5 * (ws4 * ws4 / dws4 + dws4);
  const double Rv11 = 1.0, Rv21 = u - a * nx, Rv31 = v - a * ny, Rv41 = H - un * a;
  const double Rv12 = 0.0, Rv22 = mx, Rv32 = my, Rv42 = um;
  const double Rv13 = 1.0, Rv23 = u, Rv33 = v, Rv43 = 0.5 * (u * u + v * v);
  const double Rv14 = 1.0, Rv24 = u + a * nx, Rv34 = v + a * ny, Rv44 = H + un * a;
  double diss1 = 0.0, diss2 = 0.0, diss3 = 0.0, diss4 = 0.0;
  diss1 = diss1 + ws1 * LdU1 * Rv11;
  diss1 = diss1 + ws2 * LdU2 * Rv12;
  diss1 = diss1 + ws3 * LdU3 * Rv13;
  diss1 = diss1 + ws4 * LdU4 * Rv14;
  diss2 = diss2 + ws1 * LdU1 * Rv21;
  diss2 = diss2 + ws2 * LdU2 * Rv22;
  diss2 = diss2 + ws3 * LdU3 * Rv23;
  diss2 = diss2 + ws4 * LdU4 * Rv24;
  diss3 = diss3 + ws1 * LdU1 * Rv31;
  diss3 = diss3 + ws2 * LdU2 * Rv32;
  diss3 = diss3 + ws3 * LdU3 * Rv33;
  diss3 = diss3 + ws4 * LdU4 * Rv34;
  diss4 = diss4 + ws1 * LdU1 * Rv41;
  diss4 = diss4 + ws2 * LdU2 * Rv42;
  diss4 = diss4 + ws3 * LdU3 * Rv43;
  diss4 = diss4 + ws4 * LdU4 * Rv44;
  const double fL1 = rhoL * unL;
  const double fL2 = rhoL * unL * uL + pL * nx;
  const double fL3 = rhoL * unL * vL + pL * ny;
  const double fL4 = rhoL * unL * HL;
  const double fR1 = rhoR * unR;
  const double fR2 = rhoR * unR * uR + pR * nx;
  const double fR3 = rhoR * unR * vR + pR * ny;
  const double fR4 = rhoR * unR * HR;
  flux[0] = -0.5 * (fL1 + fR1 - diss1) * side;
  flux[1] = -0.5 * (fL2 + fR2 - diss2) * side;
  flux[2] = -0.5 * (fL3 + fR3 - diss3) * side;
  flux[3] = -0.5 * (fL4 + fR4 - diss4) * side;
}

inline void face_flux(int kind, const double* L, const double* R, double nx, double ny, double side, double gamma,
                      double* out) {
  if (kind == ORC_FLUX_ROE)
    roe(L[0], L[1], L[2], L[3], R[0], R[1], R[2], R[3], nx, ny, side, gamma, out);
  else if (kind == ORC_FLUX_AUSM_PLUS_M2020)
    ausm_plus_m2020(L[0], L[1], L[2], L[3], R[0], R[1], R[2], R[3], nx, ny, side, gamma, out);
  else
    ausm_plus(L[0], L[1], L[2], L[3], R[0], R[1], R[2], R[3], nx, ny, side, gamma, out);
}

// ---------------------------------------------------------------- thermo.jl
// calcAirViscosityPowerLawFromT, thermo.jl:13-15
inline double air_mu_power_law(double T) { return 1.716e-5 * std::pow(T / 273.11, 2.0 / 3.0); }
// Julia literal powers: T^(-2) = (inv(T))^2, T^(-1) = inv(T), T^2 = T*T, T^3 = T*T*T, T^4 generic pow
inline double lit_pow_m2(double T) { const double i = 1.0 / T; return i * i; }
inline double lit_pow_4(double T) { return std::pow(T, 4.0); }
// calcAirCpLow / calcAirCpHigh, thermo.jl:17-41
inline double air_cp_low(double T) {
  return 2898903.0 * lit_pow_m2(T) + -56496.26 * (1.0 / T) + 1437.799 + -1.653609 * T + 0.003062254 * (T * T) +
         -2.279138e-6 * (T * T * T) + 6.272365e-10 * lit_pow_4(T);
}
inline double air_cp_high(double T) {
  return 6.932494e+7 * lit_pow_m2(T) + -361053.2 * (1.0 / T) + 1476.665 + -0.06138349 * T + 2.027963e-5 * (T * T) +
         -3.075525e-9 * (T * T * T) + 1.888054e-13 * lit_pow_4(T);
}
// calcAirHeatCapacityFromT, thermo.jl:43-54 (T == 200 exactly falls through to the last branch)
inline double air_cp(double T) {
  if (T < 200.0) return air_cp_low(200.0);
  else if (T > 200.0 && T < 1000.0) return air_cp_low(T);
  else if (T >= 1000.0 && T < 6000.0) return air_cp_high(T);
  else return air_cp_high(6000.0);
}
// calcAirConductivityFromT, thermo.jl:57-60
inline double air_k(double T) {
  return 15.0 / 4.0 * 8.314462 / 28.9647 * air_mu_power_law(T) * (4.0 / 15.0 * air_cp(T) * 28.9647 / 8.314462 + 1.0 / 3.0);
}

// --------------------------------------------------------------- boundaries
// getSymmetricalPointAboutCellEdge, calcGradients.jl:21-34
inline void mirror_point(double px, double py, double x0, double y0, double x1, double y1, double* sp) {
  const double dx = x1 - x0;
  const double dy = y1 - y0;
  const double a = (dx * dx - dy * dy) / (dx * dx + dy * dy);
  const double b = 2.0 * dx * dy / (dx * dx + dy * dy);
  sp[0] = a * (px - x0) + b * (py - y0) + x0;
  sp[1] = b * (px - x0) - a * (py - y0) + y0;
}
// computeTGV2d, initfields_TGV2d.jl:9-27 (P, U, L, mu from the BC data)
inline void tgv2d(double x, double y, double time, double RGAS, const double* d, double* q) {
  const double P = d[0], U = d[1], L = d[2], mu = d[3];
  const double T = P / RGAS;
  const double F = std::exp(-2.0 * mu * time / L / L);
  const double Fp = std::exp(-4.0 * mu * time / L / L);
  q[0] = P / RGAS / T;
  q[1] = U * std::sin(x / L) * std::cos(y / L) * F;
  q[2] = -U * std::cos(x / L) * std::sin(y / L) * F;
  q[3] = P + 0.25 * U * U * (std::cos(2 * x / L) + std::cos(2 * y / L)) * Fp;
}

// Ghost primitive state for face k of cell i (index = 2k-1):
// ComputeUPhysFromBoundaries, boundaryConditions2d.jl:29-57 (fixed inflow /
// slip wall / zero-gradient) and boundaryConditions_TGV2d.jl:16-30 (TGV mirror);
// the wall reflection is updateVelocityFromCurvWall, bcInviscidWall.jl:109-126.
void bc_ghost(const orc_mesh* m, const orc_thermo* th, const orc_bc* table, int64_t i, int k, const double* cur,
              double t, double* bnd) {
  const int64_t n = m->nCells;
  const int32_t code = I(m->cell_stiffness, n, i, k);
  const int32_t idx = -code;
  const orc_bc* b = (idx >= 0 && idx <= ORC_MAX_BC) ? &table[idx] : nullptr;
  const int kind = b ? b->kind : ORC_BC_EXTRAPOLATE;
  const double nx = D(m->cell_edges_Nx, n, i, k), ny = D(m->cell_edges_Ny, n, i, k);
  switch (kind) {
    case ORC_BC_FIXED:
      bnd[0] = b->data[0]; bnd[1] = b->data[1]; bnd[2] = b->data[2]; bnd[3] = b->data[3];
      break;
    case ORC_BC_SLIP_WALL:
      bnd[0] = cur[0];
      bnd[1] = cur[1] * (ny * ny - nx * nx) - 2.0 * nx * ny * cur[2];
      bnd[2] = cur[2] * (nx * nx - ny * ny) - 2.0 * nx * ny * cur[1];
      bnd[3] = cur[3];
      break;
    case ORC_BC_TGV_MIRROR: {
      const int index = 2 * k - 1;
      const double x0 = D(m->cell_mid_points, n, i, 1), y0 = D(m->cell_mid_points, n, i, 2);
      const int32_t n1 = I(m->cells2nodes, n, i, index), n2 = I(m->cells2nodes, n, i, index + 1);
      double sp[2];
      mirror_point(x0, y0, m->xNodes[n1 - 1], m->yNodes[n1 - 1], m->xNodes[n2 - 1], m->yNodes[n2 - 1], sp);
      tgv2d(sp[0], sp[1], t, th->RGAS, b->data, bnd);
      break;
    }
    case ORC_BC_ISOTHERMAL_WALL: {
      // no-slip isothermal ghost: u_g = -u_c, T_g = 2 T_w - T_c, p_g = p_c
      const double Tc = cur[3] / (cur[0] * th->RGAS);
      double Tg = 2.0 * b->data[0] - Tc;
      if (Tg < 0.5 * b->data[0]) Tg = 0.5 * b->data[0];
      bnd[0] = cur[3] / (th->RGAS * Tg); bnd[1] = -cur[1]; bnd[2] = -cur[2]; bnd[3] = cur[3];
      break;
    }
    default:
      bnd[0] = cur[0]; bnd[1] = cur[1]; bnd[2] = cur[2]; bnd[3] = cur[3];
  }
}

// computeInterfaceSlope, computeslope2d.jl:3-136: reconstruct the two face
// states of face k of cell i and evaluate the face flux with the neighbour
// side passed as "L" (normals point into cell i).
void interface_flux(const orc_mesh* m, const orc_fields* f, const orc_thermo* th, const orc_params* prm, int64_t i,
                    int k, const double* uLeftp, double flowTime, double* flux, double* UpLeftOut = nullptr,
                    double* UpRightOut = nullptr) {
  const int64_t n = m->nCells;
  const int32_t ek = I(m->cell_stiffness, n, i, k);
  const int32_t ek_type = I(m->mesh_connectivity, n, i, 2);
  const double side = D(m->cell_edges_length, n, i, k);
  const double nx = D(m->cell_edges_Nx, n, i, k);
  const double ny = D(m->cell_edges_Ny, n, i, k);
  double uUpp[4] = {0, 0, 0, 0}, uDownp[4] = {0, 0, 0, 0}, uRightp[4] = {0, 0, 0, 0};
  const int index = 2 * k - 1;
  if (ek >= 1 && ek <= n) {
    if (ek_type == 3) {
      const int32_t pDown1 = I(m->node2cellsL2down, n, i, index);
      const int32_t pUp1 = I(m->node2cellsL2up, n, i, index);
      uUpp[0] = f->densityNodes[pUp1 - 1]; uUpp[1] = f->UxNodes[pUp1 - 1];
      uUpp[2] = f->UyNodes[pUp1 - 1]; uUpp[3] = f->pressureNodes[pUp1 - 1];
      uDownp[0] = f->densityNodes[pDown1 - 1]; uDownp[1] = f->UxNodes[pDown1 - 1];
      uDownp[2] = f->UyNodes[pDown1 - 1]; uDownp[3] = f->pressureNodes[pDown1 - 1];
    } else if (ek_type == 2) {
      const int32_t pDown1 = I(m->node2cellsL2down, n, i, index), pDown2 = I(m->node2cellsL2down, n, i, index + 1);
      const int32_t pUp1 = I(m->node2cellsL2up, n, i, index), pUp2 = I(m->node2cellsL2up, n, i, index + 1);
      uUpp[0] = 0.5 * (f->densityNodes[pUp1 - 1] + f->densityNodes[pUp2 - 1]);
      uUpp[1] = 0.5 * (f->UxNodes[pUp1 - 1] + f->UxNodes[pUp2 - 1]);
      uUpp[2] = 0.5 * (f->UyNodes[pUp1 - 1] + f->UyNodes[pUp2 - 1]);
      uUpp[3] = 0.5 * (f->pressureNodes[pUp1 - 1] + f->pressureNodes[pUp2 - 1]);
      uDownp[0] = 0.5 * (f->densityNodes[pDown1 - 1] + f->densityNodes[pDown2 - 1]);
      uDownp[1] = 0.5 * (f->UxNodes[pDown1 - 1] + f->UxNodes[pDown2 - 1]);
      uDownp[2] = 0.5 * (f->UyNodes[pDown1 - 1] + f->UyNodes[pDown2 - 1]);
      uDownp[3] = 0.5 * (f->pressureNodes[pDown1 - 1] + f->pressureNodes[pDown2 - 1]);
    }
    uRightp[0] = f->densityCells[ek - 1]; uRightp[1] = f->UxCells[ek - 1];
    uRightp[2] = f->UyCells[ek - 1]; uRightp[3] = f->pressureCells[ek - 1];
  } else {
    bc_ghost(m, th, prm->bc, i, k, uLeftp, flowTime, uRightp);
    for (int v = 0; v < 4; ++v) { uDownp[v] = uLeftp[v]; uUpp[v] = uRightp[v]; }
  }
  const double ksi = prm->limiter_eps;
  double UpLeft[4], UpRight[4];
  for (int v = 0; v < 4; ++v)
    UpLeft[v] = uLeftp[v] + 0.5 * limiter(prm->limiter, uLeftp[v] - uDownp[v], uRightp[v] - uLeftp[v], ksi);
  for (int v = 0; v < 4; ++v)
    UpRight[v] = uRightp[v] - 0.5 * limiter(prm->limiter, uRightp[v] - uLeftp[v], uUpp[v] - uRightp[v], ksi);
  if (UpLeftOut) for (int v = 0; v < 4; ++v) UpLeftOut[v] = UpLeft[v];
  if (UpRightOut) for (int v = 0; v < 4; ++v) UpRightOut[v] = UpRight[v];
  if (flux) face_flux(prm->flux, UpRight, UpLeft, nx, ny, side, th->Gamma, flux);
}

template <class F>
int for_partition(const int32_t* table, int32_t nThreads, F&& body) {
  if (nThreads < 1) return -1;
  for (int p = 0; p < nThreads; ++p)
    if (table[p] < 1 || table[nThreads + p] < table[p] - 1) return -2;  // 0:0 ranges index cell 0 in Julia
#pragma omp parallel for num_threads(nThreads) schedule(static, 1)
  for (int p = 0; p < nThreads; ++p) {
    const int64_t b = table[p], e = table[nThreads + p];
    for (int64_t i = b; i <= e; ++i) body(i);
  }
  return 0;
}

// calcPhiInCellCluster, calcGradients.jl:3-18: unweighted mean over a node's cluster
inline double phi_in_cluster(const orc_mesh* m, int32_t node, const double* field) {
  double phi = 0.0;
  int64_t counter = 0;
  for (int s = 1; s <= m->nNeibCells; ++s) {
    const int32_t cell = I(m->cell_clusters, m->nNodes, node, s);
    if (cell > 0 && cell <= m->nCells) { phi += field[cell - 1]; counter += 1; }
  }
  return counter != 0 ? phi / (double)counter : 0.0;
}

// calcScalarFieldGreenGaussNodesBasedGradient, calcGradients.jl:36-139.  The
// boundary branch (:100-113) is overwritten by :116-126 and has no effect; the
// normals are used swapped, so "gradX" is d/dy and "gradY" is d/dx.
void green_gauss_cell(const orc_mesh* m, const double* phi, int64_t C, double* gx, double* gy) {
  const int64_t n = m->nCells;
  double side[4] = {0, 0, 0, 0}, nx[4] = {0, 0, 0, 0}, ny[4] = {0, 0, 0, 0};
  double pl[4] = {0, 0, 0, 0}, pr[4] = {0, 0, 0, 0};
  const int nf = I(m->mesh_connectivity, n, C, 3) == 4 ? 4 : 3;
  for (int k = 1; k <= nf; ++k) {
    side[k - 1] = D(m->cell_edges_length, n, C, k);
    nx[k - 1] = D(m->cell_edges_Nx, n, C, k);
    ny[k - 1] = D(m->cell_edges_Ny, n, C, k);
    pl[k - 1] = phi_in_cluster(m, I(m->cells2nodes, n, C, 2 * k - 1), phi);
    pr[k - 1] = phi_in_cluster(m, I(m->cells2nodes, n, C, 2 * k), phi);
  }
  double fx[4], fy[4];
  for (int k = 0; k < 4; ++k) {
    fx[k] = 0.5 * (pl[k] + pr[k]) * -ny[k] * side[k];
    fy[k] = 0.5 * (pl[k] + pr[k]) * -nx[k] * side[k];
  }
  const double A = m->cell_areas[C - 1];
  *gx = (fx[0] + fx[1] + fx[2] + fx[3]) / A;
  *gy = (fy[0] + fy[1] + fy[2] + fy[3]) / A;
}

}  // namespace

extern "C" {

double orc_limiter(int32_t kind, double x, double y, double c) { return limiter(kind, x, y, c); }

// the reference scheme's ghost state of face k of cell i (both 1-based) from
// the cell's primitives, for the ghost's known-answer tests
void orc_ref_bc_ghost(const orc_mesh* m, const orc_thermo* th, const orc_params* prm, int64_t i, int32_t k,
                      const double* cur, double t, double* out) {
  bc_ghost(m, th, prm->bc, i, k, cur, t, out);
}

void orc_ausm_plus_flux(const double* q, double nx, double ny, double side, double gamma, double* flux) {
  ausm_plus(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], nx, ny, side, gamma, flux);
}
void orc_ausm_plus_m2020_flux(const double* q, double nx, double ny, double side, double gamma, double* flux) {
  ausm_plus_m2020(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], nx, ny, side, gamma, flux);
}
void orc_roe_flux(const double* q, double nx, double ny, double side, double gamma, double* flux) {
  roe(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], nx, ny, side, gamma, flux);
}
double orc_air_conductivity(double T) { return air_k(T); }
// calcSutherlandViscosityOF, sutherland.jl:2-6
double orc_sutherland_viscosity(double T) { return 1.4792e-6 * std::sqrt(T) / (1.0 + 116.0 / T); }

// distibuteCellsInThreadsSA, partMesh2d.jl:2-34
void orc_distribute_cells_in_threads(int32_t nThreads, int64_t nCells, int32_t* t) {
  for (int i = 0; i < 2 * nThreads; ++i) t[i] = 0;
  if (nThreads > 1) {
    const int64_t nParts = nCells / nThreads;
    for (int i = 1; i <= nThreads; ++i) t[nThreads + i - 1] = (int32_t)(nCells - nParts * (nThreads - i));
    for (int i = 1; i <= nThreads; ++i) t[i - 1] = (int32_t)(t[nThreads + i - 1] - nParts + 1);
    t[0] = 1;
  }
}

// phs2dcns2dcellsSA, utilsFVM2dp.jl:4-17
void orc_phs2dcns(const orc_fields* f, int64_t n, double gamma, double* U) {
  for (int64_t i = 0; i < n; ++i) {
    U[i] = f->densityCells[i];
    U[n + i] = f->densityCells[i] * f->UxCells[i];
    U[2 * n + i] = f->densityCells[i] * f->UyCells[i];
    U[3 * n + i] = f->pressureCells[i] / (gamma - 1.0) +
                   0.5 * f->densityCells[i] * (f->UxCells[i] * f->UxCells[i] + f->UyCells[i] * f->UyCells[i]);
  }
}

// cells2nodesSolutionReconstructionWithStencils, utilsFVM2dp.jl:299-365
int orc_cells2nodes(const int32_t* nodesThreads, int32_t nThreads, const orc_mesh* m, const orc_fields* f,
                    const double* Uc, double* Un) {
  const int64_t nc = m->nCells, nn = m->nNodes;
  return for_partition(nodesThreads, nThreads, [&](int64_t J) {
    double det = 0.0;
    double r = 0.0, u = 0.0, v = 0.0, p = 0.0, s[4] = {0, 0, 0, 0};
    for (int j = 1; j <= 8; ++j) {
      const int32_t c = I(m->cell_clusters, nn, J, j);
      if (c != 0) {
        const double wi = D(m->node_stencils, nn, J, j);
        if (Un) for (int q = 0; q < 4; ++q) s[q] += Uc[q * nc + c - 1] * wi;
        r += f->densityCells[c - 1] * wi;
        u += f->UxCells[c - 1] * wi;
        v += f->UyCells[c - 1] * wi;
        p += f->pressureCells[c - 1] * wi;
        det += wi;
      }
    }
    if (det != 0) {
      r = r / det; u = u / det; v = v / det; p = p / det;
      for (int q = 0; q < 4; ++q) s[q] = s[q] / det;
    }
    f->densityNodes[J - 1] = r; f->UxNodes[J - 1] = u; f->UyNodes[J - 1] = v; f->pressureNodes[J - 1] = p;
    if (Un) for (int q = 0; q < 4; ++q) Un[q * nn + J - 1] = s[q];
  });
}

// calcOneStageMUSCLcpuFast, calcOneStageMUSCLcpuFast.jl:3-107 (per-thread body
// SecondOrderUpwindM2, SOUscheme.jl:3-118)
int orc_calc_one_stage_muscl_cpu_fast(const int32_t* cellsThreads, int32_t nThreads, double bettaKJ, double dt,
                                      double flowTime, const orc_mesh* m, const orc_fields* f, const orc_thermo* th,
                                      const orc_params* prm, const double* UOld, double* FLUXES, const double* Udiff,
                                      double* UNew) {
  const int64_t n = m->nCells;
  return for_partition(cellsThreads, nThreads, [&](int64_t i) {
    const double uLeftp[4] = {f->densityCells[i - 1], f->UxCells[i - 1], f->UyCells[i - 1], f->pressureCells[i - 1]};
    double e1[4] = {0, 0, 0, 0}, e2[4] = {0, 0, 0, 0}, e3[4] = {0, 0, 0, 0}, e4[4] = {0, 0, 0, 0};
    double F[4];
    const int32_t nv = I(m->mesh_connectivity, n, i, 3);
    if (nv == 3) {
      interface_flux(m, f, th, prm, i, 1, uLeftp, flowTime, e1);
      interface_flux(m, f, th, prm, i, 2, uLeftp, flowTime, e2);
      interface_flux(m, f, th, prm, i, 3, uLeftp, flowTime, e3);
      for (int v = 0; v < 4; ++v) F[v] = e1[v] + e2[v] + e3[v];
    } else {
      interface_flux(m, f, th, prm, i, 1, uLeftp, flowTime, e1);
      interface_flux(m, f, th, prm, i, 2, uLeftp, flowTime, e2);
      interface_flux(m, f, th, prm, i, 3, uLeftp, flowTime, e3);
      interface_flux(m, f, th, prm, i, 4, uLeftp, flowTime, e4);
      for (int v = 0; v < 4; ++v) F[v] = e1[v] + e2[v] + e3[v] + e4[v];
    }
    const double Rarea = 1.0 / m->cell_areas[i - 1];
    for (int v = 0; v < 4; ++v) {
      if (FLUXES) FLUXES[v * n + i - 1] = F[v];
      const double ud = (prm->include_udiff && Udiff) ? Udiff[v * n + i - 1] : 0.0;
      if (prm->include_udiff && Udiff)
        UNew[v * n + i - 1] = (UOld[v * n + i - 1] - F[v] * bettaKJ * dt * Rarea + bettaKJ * dt * ud);
      else
        UNew[v * n + i - 1] = (UOld[v * n + i - 1] - F[v] * bettaKJ * dt * Rarea);
    }
  });
}

// calcOneStageMUSCL, calcOneStageMUSCL.jl:4-106 with its face-state phase
// computeMUSCLStencilsCUDA / computeInterfaceSlopeCUDA (computeMUSCLStencilsCUDA.jl:3-28,
// computeslope2d.jl:139-269); no Udiff (commented out in the source).
int orc_calc_one_stage_muscl(const int32_t* cellsThreads, int32_t nThreads, double betta, double dt, double flowTime,
                             const orc_mesh* m, const orc_fields* f, const orc_thermo* th, const orc_params* prm,
                             const double* UOld, double* UNew, double* FLUXES, double* ULeftOut, double* URightOut) {
  const int64_t n = m->nCells;
  std::vector<double> UL(16 * n, 0.0), UR(16 * n, 0.0);  // (4, n) per face slot: ULeft1..4, URight1..4
  int rc = for_partition(cellsThreads, nThreads, [&](int64_t i) {
    const double uLeftp[4] = {f->densityCells[i - 1], f->UxCells[i - 1], f->UyCells[i - 1], f->pressureCells[i - 1]};
    const int nf = I(m->mesh_connectivity, n, i, 3) == 4 ? 4 : 3;
    for (int k = 1; k <= nf; ++k)
      interface_flux(m, f, th, prm, i, k, uLeftp, flowTime, nullptr, &UL[(k - 1) * 4 * n + 4 * (i - 1)],
                     &UR[(k - 1) * 4 * n + 4 * (i - 1)]);
  });
  if (rc) return rc;
  if (ULeftOut) std::copy(UL.begin(), UL.end(), ULeftOut);
  if (URightOut) std::copy(UR.begin(), UR.end(), URightOut);
  return for_partition(cellsThreads, nThreads, [&](int64_t C) {
    double e[4][4] = {{0}};
    const int nf = I(m->mesh_connectivity, n, C, 3) == 4 ? 4 : 3;
    for (int k = 1; k <= nf; ++k) {
      const double* l = &UL[(k - 1) * 4 * n + 4 * (C - 1)];
      const double* r = &UR[(k - 1) * 4 * n + 4 * (C - 1)];
      face_flux(prm->flux, r, l, D(m->cell_edges_Nx, n, C, k), D(m->cell_edges_Ny, n, C, k),
                D(m->cell_edges_length, n, C, k), th->Gamma, e[k - 1]);
    }
    double F[4];
    for (int v = 0; v < 4; ++v) F[v] = nf == 4 ? e[0][v] + e[1][v] + e[2][v] + e[3][v] : e[0][v] + e[1][v] + e[2][v];
    const double Rarea = 1.0 / m->cell_areas[C - 1];
    for (int v = 0; v < 4; ++v) {
      if (FLUXES) FLUXES[v * n + C - 1] = F[v];
      UNew[v * n + C - 1] = (UOld[v * n + C - 1] - F[v] * betta * dt * Rarea);
    }
  });
}

int orc_green_gauss_gradient(const int32_t* cellsThreads, int32_t nThreads, const orc_mesh* m, const double* phi,
                             double* gx, double* gy) {
  return for_partition(cellsThreads, nThreads, [&](int64_t C) { green_gauss_cell(m, phi, C, &gx[C - 1], &gy[C - 1]); });
}

// calcDiffTerm, calcDiffTerm.jl:31-121.  The node interpolations at :69-76 do
// not influence the result (the gradient routine never reads its node-field
// argument) and are skipped; Udiff[:,1] is left as the caller passed it.
int orc_calc_diff_term(const int32_t* cellsThreads, int32_t nThreads, const orc_mesh* m, const orc_fields* f,
                       const orc_thermo* th, const orc_params* prm, const double* UC, double* Udiff, double* lap3) {
  const int64_t n = m->nCells;
  std::vector<double> g[3][2], dd[3][2];
  for (int q = 0; q < 3; ++q)
    for (int c = 0; c < 2; ++c) { g[q][c].assign(n, 0.0); dd[q][c].assign(n, 0.0); }
  std::vector<double> dummy(n, 0.0);
  int rc = 0;
  for (int q = 0; q < 3; ++q)
    rc |= orc_green_gauss_gradient(cellsThreads, nThreads, m, UC + (q + 1) * n, g[q][0].data(), g[q][1].data());
  for (int q = 0; q < 3; ++q) {
    rc |= orc_green_gauss_gradient(cellsThreads, nThreads, m, g[q][0].data(), dd[q][0].data(), dummy.data());
    rc |= orc_green_gauss_gradient(cellsThreads, nThreads, m, g[q][1].data(), dummy.data(), dd[q][1].data());
  }
  if (rc) return rc;
  return for_partition(cellsThreads, nThreads, [&](int64_t i) {
    const double lu = dd[0][0][i - 1] + dd[0][1][i - 1];
    const double lv = dd[1][0][i - 1] + dd[1][1][i - 1];
    const double le = dd[2][0][i - 1] + dd[2][1][i - 1];
    if (lap3) { lap3[i - 1] = lu; lap3[n + i - 1] = lv; lap3[2 * n + i - 1] = le; }
    const double T = f->pressureCells[i - 1] / f->densityCells[i - 1] / th->RGAS;
    const double mu = prm->mu_const;
    const double k = prm->energy_law == ORC_ENERGY_CONST ? prm->k_const : air_k(T);
    Udiff[n + i - 1] = mu * lu;
    Udiff[2 * n + i - 1] = mu * lv;
    Udiff[3 * n + i - 1] = mu * k * le;
  });
}

// updateVariablesSA, evaluate2d.jl:52-110 (the density clamp at :63-77 refers
// to an undefined variable in the source and is not restated)
int orc_update_variables(const int32_t* cellsThreads, int32_t nThreads, int64_t n, double Gamma, const double* Un,
                         double* Uo, double* Delta, const orc_fields* f) {
  return for_partition(cellsThreads, nThreads, [&](int64_t i1) {
    const int64_t i = i1 - 1;
    f->densityCells[i] = Un[i];
    f->UxCells[i] = Un[n + i] / Un[i];
    f->UyCells[i] = Un[2 * n + i] / Un[i];
    f->pressureCells[i] = (Gamma - 1.0) * (Un[3 * n + i] - 0.5 * (Un[n + i] * Un[n + i] + Un[2 * n + i] * Un[2 * n + i]) / Un[i]);
    f->aSoundCells[i] = std::sqrt(Gamma * f->pressureCells[i] / f->densityCells[i]);
    f->VMAXCells[i] = std::sqrt(f->UxCells[i] * f->UxCells[i] + f->UyCells[i] * f->UyCells[i]) + f->aSoundCells[i];
    for (int q = 0; q < 4; ++q) {
      if (Delta) Delta[q * n + i] = Un[q * n + i] - Uo[q * n + i];
      Uo[q * n + i] = Un[q * n + i];
    }
  });
}

// dynControls.tau = CFL * maxEdge / max(velmax, 1e-6), testViscousFluxesTGV2d.jl:420-423
double orc_ref_tau(const orc_mesh* m, const orc_fields* f, double cfl) {
  double maxEdge = m->HX[0], vmax = f->VMAXCells[0];
  for (int64_t i = 1; i < m->nCells; ++i) {
    if (m->HX[i] > maxEdge) maxEdge = m->HX[i];
    if (f->VMAXCells[i] > vmax) vmax = f->VMAXCells[i];
  }
  return cfl * maxEdge / jmax(vmax, 1.0e-6);
}

// CFL-limited step: cfl * min_i h_i / VMAX_i with h = 2 A / perimeter (the
// inscribed-circle radius of a triangle); the reference's own tau uses the
// longest edge instead, which is not a stable bound on triangles.
double orc_local_dt(const orc_mesh* m, const orc_fields* f, double cfl) {
  const int64_t n = m->nCells;
  double best = INFINITY;
  for (int64_t i = 1; i <= n; ++i) {
    const int nv = I(m->mesh_connectivity, n, i, 3);
    double perim = 0.0;
    for (int k = 1; k <= nv; ++k) perim += D(m->cell_edges_length, n, i, k);
    const double h = 2.0 * m->cell_areas[i - 1] / perim;
    const double r = h / f->VMAXCells[i - 1];
    if (r < best) best = r;
  }
  return cfl * best;
}

}  // extern "C"

// ------------------------------------------------------ the reference's GPU stages
// kernel_AUSM2d of AUSMflux2dCUDA.jl:56-98 (FP64) and of AUSMflux2dCUDAx32.jl:56-98
// (Float32 arrays), restated at the level of the source expressions with
// Julia's promotion rules: in the Float32 kernel the split functions
// (AUSMflux2dCUDAx32.jl:19-48) take Float32 arguments but their literals
// (0.5, 0.25, 1.0, 2.0) are Float64, so M + 1.0 and the products with them are
// Float64, M*M and AUSM_ALFA*M stay Float32, and each function returns Float32;
// the kernel body is Float32 throughout.  L = the neighbour side (uRight).
namespace {

struct CudaSplitF64 {
  static double M_p(double M, double B) {
    return std::fabs(M) >= 1.0 ? 0.5 * (M + std::fabs(M)) : 0.25 * (M + 1.0) * (M + 1.0) + B * (M * M - 1.0) * (M * M - 1.0);
  }
  static double M_m(double M, double B) {
    return std::fabs(M) >= 1.0 ? 0.5 * (M - std::fabs(M))
                               : -0.25 * (M - 1.0) * (M - 1.0) - B * (M * M - 1.0) * (M * M - 1.0);
  }
  static double P_p(double M, double A) {
    return std::fabs(M) >= 1.0 ? 0.5 * (1.0 + jsign(M))
                               : 0.25 * (M + 1.0) * (M + 1.0) * (2.0 - M) + A * M * (M * M - 1.0) * (M * M - 1.0);
  }
  static double P_m(double M, double A) {
    return std::fabs(M) >= 1.0 ? 0.5 * (1.0 - jsign(M))
                               : 0.25 * (M - 1.0) * (M - 1.0) * (2.0 + M) - A * M * (M * M - 1.0) * (M * M - 1.0);
  }
};

inline float fsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : x); }

struct CudaSplitX32 {  // Float32 in and out, Float64 where a literal meets the argument
  static float M_p(float M, float B) {
    if (std::fabs(M) >= 1.0) return (float)(0.5 * (double)(M + std::fabs(M)));
    const double a = (double)M + 1.0, d = (double)(M * M) - 1.0;
    return (float)(0.25 * a * a + (double)B * d * d);
  }
  static float M_m(float M, float B) {
    if (std::fabs(M) >= 1.0) return (float)(0.5 * (double)(M - std::fabs(M)));
    const double a = (double)M - 1.0, d = (double)(M * M) - 1.0;
    return (float)(-0.25 * a * a - (double)B * d * d);
  }
  static float P_p(float M, float A) {
    if (std::fabs(M) >= 1.0) return (float)(0.5 * (1.0 + (double)fsign(M)));
    const double a = (double)M + 1.0, d = (double)(M * M) - 1.0;
    return (float)(0.25 * a * a * (2.0 - (double)M) + (double)(A * M) * d * d);
  }
  static float P_m(float M, float A) {
    if (std::fabs(M) >= 1.0) return (float)(0.5 * (1.0 - (double)fsign(M)));
    const double a = (double)M - 1.0, d = (double)(M * M) - 1.0;
    return (float)(0.25 * a * a * (2.0 + (double)M) - (double)(A * M) * d * d);
  }
};

// one kernel_AUSM2d evaluation added into acc (flux_k[i] = flux_k[i] - (...) * side)
template <class T, class Split>
void cuda_ausm(const T* L, const T* R, T nx, T ny, T side, T g, T* acc) {
  const T rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const T rR = R[0], uR = R[1], vR = R[2], pR = R[3];
  const T one = (T)1.0, two = (T)2.0, half = (T)0.5;
  const T htL = pL / rL / (g - one) + half * (uL * uL + vL * vL) + pL / rL;
  const T htR = pR / rR / (g - one) + half * (uR * uR + vR * vR) + pR / rR;
  const T aL = std::sqrt(two * (g - one) / (g + one) * htL) *
               std::min(one, std::sqrt(two * (g - one) / (g + one) * htL) / std::fabs(uL * nx + vL * ny));
  const T aR = std::sqrt(two * (g - one) / (g + one) * htR) *
               std::min(one, std::sqrt(two * (g - one) / (g + one) * htR) / std::fabs(uR * nx + vR * ny));
  const T a12 = std::min(aL, aR);
  const T ML = (uL * nx + vL * ny) / a12, MR = (uR * nx + vR * ny) / a12;
  const T B = (T)(1.0 / 8.0), A = (T)(3.0 / 16.0);
  const T m = Split::M_p(ML, B) + Split::M_m(MR, B);
  const T p12 = Split::P_p(ML, A) * pL + Split::P_m(MR, A) * pR;
  const T mp = half * (m + std::fabs(m)), mm = half * (m - std::fabs(m));
  acc[0] = acc[0] - (a12 * (mp * rL + mm * rR)) * side;
  acc[1] = acc[1] - (a12 * (mp * rL * uL + mm * rR * uR) + p12 * nx) * side;
  acc[2] = acc[2] - (a12 * (mp * rL * vL + mm * rR * vR) + p12 * ny) * side;
  acc[3] = acc[3] - (a12 * (mp * rL * htL + mm * rR * htR)) * side;
}

template <class T, class Split>
void cuda_stage(const orc_mesh* m, const orc_thermo* th, double betta, double dt, const double* UOld,
                const double* ULeft, const double* URight, double* UNew, double* iFlux) {
  const int64_t n = m->nCells;
  const T g = (T)th->Gamma;
  for (int64_t i = 0; i < n; ++i) {
    T acc[4] = {0, 0, 0, 0};
    // slots 1..3 only: the 4th face's launch is commented out
    // (calcOneStageCUDA.jl:89-103, calcOneStageCUDAx32.jl:111-124)
    for (int k = 0; k < 3; ++k) {
      T L[4], R[4];
      for (int v = 0; v < 4; ++v) {
        L[v] = (T)URight[(int64_t)k * 4 * n + 4 * i + v];  // copyto! rounds Float64 to T
        R[v] = (T)ULeft[(int64_t)k * 4 * n + 4 * i + v];
      }
      cuda_ausm<T, Split>(L, R, (T)m->cell_edges_Nx[(int64_t)k * n + i], (T)m->cell_edges_Ny[(int64_t)k * n + i],
                          (T)m->cell_edges_length[(int64_t)k * n + i], g, acc);
    }
    const double Rarea = 1.0 / m->cell_areas[i];
    for (int v = 0; v < 4; ++v) {
      const double fl = (double)acc[v];
      if (iFlux) iFlux[v * n + i] = fl;
      // calcOneStageCUDAx32.jl:138-153: the update on the host in Float64, no Udiff
      UNew[v * n + i] = UOld[v * n + i] - fl * betta * dt * Rarea;
    }
  }
}

}  // namespace

// calcOneStageCUDA (calcOneStageCUDA.jl:4-135) / calcOneStageCUDAx32
// (calcOneStageCUDAx32.jl:5-155) from the caller's face states ULeft1..4 /
// URight1..4 (four (4, nCells) column-major arrays each, slot after slot).
int orc_calc_one_stage_cuda(int32_t x32, const orc_mesh* m, const orc_thermo* th, double betta, double dt,
                            const double* UOld, const double* ULeft, const double* URight, double* UNew,
                            double* iFlux) {
  if (!m || !th || !UOld || !ULeft || !URight || !UNew) return -1;
  if (x32) cuda_stage<float, CudaSplitX32>(m, th, betta, dt, UOld, ULeft, URight, UNew, iFlux);
  else cuda_stage<double, CudaSplitF64>(m, th, betta, dt, UOld, ULeft, URight, UNew, iFlux);
  return 0;
}

// the split functions alone, for the tests: which = 0 M+, 1 M-, 2 P+, 3 P-
double orc_cuda_split(int32_t x32, int32_t which, double M) {
  if (x32) {
    const float f = (float)M, B = 0.125f, A = 0.1875f;
    switch (which) {
      case 0: return CudaSplitX32::M_p(f, B);
      case 1: return CudaSplitX32::M_m(f, B);
      case 2: return CudaSplitX32::P_p(f, A);
      default: return CudaSplitX32::P_m(f, A);
    }
  }
  switch (which) {
    case 0: return CudaSplitF64::M_p(M, 0.125);
    case 1: return CudaSplitF64::M_m(M, 0.125);
    case 2: return CudaSplitF64::P_p(M, 0.1875);
    default: return CudaSplitF64::P_m(M, 0.1875);
  }
}
