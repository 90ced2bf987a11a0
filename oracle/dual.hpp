// Forward-mode dual number for the oracle (TEST INFRASTRUCTURE ONLY): a value
// and one directional derivative.  A step instantiated with it gives the exact
// tangent-linear step, the reference against which a discrete adjoint is
// checked (<adjoint, dU0> = dJ[dU0]).  Branches follow the value (fabs'(0) =
// +1; min and max keep the first argument on ties), and the value part is the
// plain double computation, operation for operation.
#pragma once
#include <cmath>

namespace orc_ad {

struct Dual {
  double v, d;
  Dual() : v(0.0), d(0.0) {}
  Dual(double x) : v(x), d(0.0) {}  // NOLINT: constants enter the scheme as duals
  Dual(double x, double dx) : v(x), d(dx) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
inline bool operator!=(Dual a, Dual b) { return a.v != b.v; }
inline Dual sqrt(Dual a) {
  const double r = std::sqrt(a.v);
  return {r, a.d / (2.0 * r)};
}
inline Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
inline Dual fmin(Dual a, Dual b) { return b.v < a.v ? b : a; }
inline Dual fmax(Dual a, Dual b) { return b.v > a.v ? b : a; }
// a*b + c with one rounding of the value (the source's muladd)
inline Dual fma(Dual a, Dual b, Dual c) { return {std::fma(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d}; }
inline double val(double x) { return x; }
inline double val(float x) { return x; }
inline double val(Dual x) { return x.v; }

}  // namespace orc_ad
