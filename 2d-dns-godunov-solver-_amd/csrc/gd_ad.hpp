// Forward-mode dual number for the reference-scheme adjoint: a value and one
// directional derivative.  The reference scheme's templated limiters and AUSM+
// (gd_ref_math.hpp) instantiated with it give, per seeded input, the exact
// derivative of the arithmetic the forward kernels perform; the face adjoint
// kernel seeds each input in turn and contracts with the incoming adjoint.
// Branches follow the value, as in the forward code (fabs'(0) = +1; min keeps
// its first argument on ties), which is also the convention of the oracle's
// tangent-linear loop the adjoint is checked against.
#pragma once
#include <hip/hip_runtime.h>

namespace gd {
// The dual number's arithmetic lives in its own namespace so that argument-
// dependent lookup finds it for AD operands while ordinary lookup from the
// double code of the enclosing namespaces keeps finding ::fma, ::sqrt, ...
namespace ad {

struct AD {
  double v, d;
  __device__ AD() : v(0.0), d(0.0) {}
  __device__ AD(double x) : v(x), d(0.0) {}  // NOLINT: constants enter the arithmetic as duals
  __device__ AD(double x, double dx) : v(x), d(dx) {}
};
__device__ __forceinline__ AD operator+(AD a, AD b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ AD operator-(AD a, AD b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ AD operator-(AD a) { return {-a.v, -a.d}; }
__device__ __forceinline__ AD operator*(AD a, AD b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ AD operator/(AD a, AD b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ bool operator<(AD a, AD b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(AD a, AD b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(AD a, AD b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>=(AD a, AD b) { return a.v >= b.v; }
__device__ __forceinline__ AD fabs(AD a) { return a.v < 0.0 ? -a : a; }
__device__ __forceinline__ AD sqrt(AD a) {
  const double r = ::sqrt(a.v);
  return {r, a.d / (2.0 * r)};
}
// a*b + c with the value rounded once (the reference's muladd)
__device__ __forceinline__ AD fma(AD a, AD b, AD c) { return {::fma(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d}; }
// Julia's sign (derivative 0) and NaN-propagating min, as refm::jsign / jmin
__device__ __forceinline__ AD jsign(AD x) { return AD(x.v > 0.0 ? 1.0 : (x.v < 0.0 ? -1.0 : x.v)); }
__device__ __forceinline__ AD jmin(AD x, AD y) { return (x.v != x.v || y.v != y.v) ? x + y : (y.v < x.v ? y : x); }

}  // namespace ad
using ad::AD;
}  // namespace gd
