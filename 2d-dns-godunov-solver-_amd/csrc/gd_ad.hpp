// Forward-mode dual number for the reference-scheme adjoint: a value and two
// directional derivatives.  The reference scheme's templated limiter
// (gd_ref_math.hpp) instantiated with it gives the limiter's value, with the
// forward kernels' double arithmetic, and its two partial derivatives.
// Branches follow the value, as in the forward code (fabs'(0) = +1; min keeps
// its first argument on ties), which is also the convention of the oracle's
// tangent-linear loop the adjoint is checked against.
#pragma once
#include <hip/hip_runtime.h>

namespace gd {
// The dual number's arithmetic lives in its own namespace so that argument-
// dependent lookup finds it for AD2 operands while ordinary lookup from the
// double code of the enclosing namespaces keeps finding ::fma, ::sqrt, ...
namespace ad {

// Two directional derivatives at once (the limiter's partials in x and y):
// the value arithmetic is the double arithmetic, each division rounds its
// value once and scales the derivatives by one reciprocal.
struct AD2 {
  double v, dx, dy;
  __device__ AD2() : v(0.0), dx(0.0), dy(0.0) {}
  __device__ AD2(double x) : v(x), dx(0.0), dy(0.0) {}  // NOLINT
  __device__ AD2(double x, double a, double b) : v(x), dx(a), dy(b) {}
};
__device__ __forceinline__ AD2 operator+(AD2 a, AD2 b) { return {a.v + b.v, a.dx + b.dx, a.dy + b.dy}; }
__device__ __forceinline__ AD2 operator-(AD2 a, AD2 b) { return {a.v - b.v, a.dx - b.dx, a.dy - b.dy}; }
__device__ __forceinline__ AD2 operator-(AD2 a) { return {-a.v, -a.dx, -a.dy}; }
__device__ __forceinline__ AD2 operator*(AD2 a, AD2 b) {
  return {a.v * b.v, a.dx * b.v + a.v * b.dx, a.dy * b.v + a.v * b.dy};
}
__device__ __forceinline__ AD2 operator/(AD2 a, AD2 b) {
  const double q = a.v / b.v, r = 1.0 / b.v;
  return {q, (a.dx - q * b.dx) * r, (a.dy - q * b.dy) * r};
}
__device__ __forceinline__ bool operator<(AD2 a, AD2 b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(AD2 a, AD2 b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(AD2 a, AD2 b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>=(AD2 a, AD2 b) { return a.v >= b.v; }
__device__ __forceinline__ AD2 fabs(AD2 a) { return a.v < 0.0 ? -a : a; }
__device__ __forceinline__ AD2 jsign(AD2 x) { return AD2(x.v > 0.0 ? 1.0 : (x.v < 0.0 ? -1.0 : x.v)); }
__device__ __forceinline__ AD2 jmin(AD2 x, AD2 y) { return (x.v != x.v || y.v != y.v) ? x + y : (y.v < x.v ? y : x); }

}  // namespace ad
using ad::AD2;
}  // namespace gd
