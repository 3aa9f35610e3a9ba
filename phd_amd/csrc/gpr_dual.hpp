// Forward-mode dual numbers for the local derivatives of the adjoint: a value
// and one directional derivative, with the arithmetic and functions the
// pointwise physics uses.  Comparisons and branches follow the value, so a
// derivative is that of the branch the value takes (max, min, |x| at a kink).
#pragma once

#include <hip/hip_runtime.h>

namespace gpr {
// The type and its functions live in their own namespace, found by
// argument-dependent lookup: calls such as sqrt(x) on a double inside gpr
// still resolve to the standard functions.
namespace ad {

struct Dual {
  double v, d;
  __device__ __forceinline__ Dual() : v(0.0), d(0.0) {}
  __device__ __forceinline__ Dual(double x) : v(x), d(0.0) {}
  __device__ __forceinline__ Dual(double x, double dx) : v(x), d(dx) {}
  __device__ __forceinline__ Dual &operator+=(const Dual &b) { v += b.v; d += b.d; return *this; }
  __device__ __forceinline__ Dual &operator-=(const Dual &b) { v -= b.v; d -= b.d; return *this; }
  __device__ __forceinline__ Dual &operator*=(const Dual &b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
  __device__ __forceinline__ Dual &operator/=(const Dual &b) {
    const double r = v / b.v;
    d = (d - r * b.d) / b.v;
    v = r;
    return *this;
  }
};

__device__ __forceinline__ Dual operator-(const Dual &a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator+(Dual a, const Dual &b) { return a += b; }
__device__ __forceinline__ Dual operator-(Dual a, const Dual &b) { return a -= b; }
__device__ __forceinline__ Dual operator*(Dual a, const Dual &b) { return a *= b; }
__device__ __forceinline__ Dual operator/(Dual a, const Dual &b) { return a /= b; }
__device__ __forceinline__ Dual operator+(Dual a, double b) { a.v += b; return a; }
__device__ __forceinline__ Dual operator+(double b, Dual a) { a.v += b; return a; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { a.v -= b; return a; }
__device__ __forceinline__ Dual operator-(double b, const Dual &a) { return Dual(b - a.v, -a.d); }
__device__ __forceinline__ Dual operator*(const Dual &a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(double b, const Dual &a) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator/(const Dual &a, double b) { return Dual(a.v / b, a.d / b); }
__device__ __forceinline__ Dual operator/(double b, const Dual &a) {
  const double r = b / a.v;
  return Dual(r, -r * a.d / a.v);
}

#define GPR_DUAL_CMP(op)                                                                             \
  __device__ __forceinline__ bool operator op(const Dual &a, const Dual &b) { return a.v op b.v; } \
  __device__ __forceinline__ bool operator op(const Dual &a, double b) { return a.v op b; }        \
  __device__ __forceinline__ bool operator op(double a, const Dual &b) { return a op b.v; }
GPR_DUAL_CMP(<)
GPR_DUAL_CMP(>)
GPR_DUAL_CMP(<=)
GPR_DUAL_CMP(>=)
GPR_DUAL_CMP(==)
GPR_DUAL_CMP(!=)
#undef GPR_DUAL_CMP

// At a zero argument with a zero derivative (sqrt of a sum of squares at
// its minimum, e.g. the stress norm of a body at rest) the derivative is
// taken as 0, the value of the one-sided limits' average.
__device__ __forceinline__ Dual sqrt(const Dual &a) {
  const double s = ::sqrt(a.v);
  return Dual(s, (s == 0.0 && a.d == 0.0) ? 0.0 : a.d / (2.0 * s));
}
__device__ __forceinline__ Dual cbrt(const Dual &a) {
  const double c = ::cbrt(a.v);
  return Dual(c, a.d / (3.0 * c * c));
}
__device__ __forceinline__ Dual exp(const Dual &a) {
  const double e = ::exp(a.v);
  return Dual(e, e * a.d);
}
__device__ __forceinline__ Dual log(const Dual &a) { return Dual(::log(a.v), a.d / a.v); }
// a^b; the exponent's derivative term only where it is non-zero (a constant
// exponent of a negative base stays finite)
__device__ __forceinline__ double exponent_value(const Dual &b) { return b.v; }
__device__ __forceinline__ Dual pow(const Dual &a, const Dual &b) {
  const double p = ::pow(a.v, b.v);
  double d = b.v * ::pow(a.v, b.v - 1.0) * a.d;
  if (b.d != 0.0) d += p * ::log(a.v) * b.d;
  return Dual(p, d);
}
__device__ __forceinline__ Dual fabs(const Dual &a) { return a.v < 0 ? -a : a; }
__device__ __forceinline__ Dual fmin(const Dual &a, const Dual &b) { return b.v < a.v ? b : a; }
__device__ __forceinline__ Dual fmax(const Dual &a, const Dual &b) { return b.v > a.v ? b : a; }
__device__ __forceinline__ Dual fmin(const Dual &a, double b) { return b < a.v ? Dual(b) : a; }
__device__ __forceinline__ Dual fmax(const Dual &a, double b) { return b > a.v ? Dual(b) : a; }
__device__ __forceinline__ Dual acos(const Dual &a) {
  return Dual(::acos(a.v), -a.d / ::sqrt(1.0 - a.v * a.v));
}
__device__ __forceinline__ Dual cos(const Dual &a) { return Dual(::cos(a.v), -::sin(a.v) * a.d); }

}  // namespace ad

using ad::Dual;

}  // namespace gpr
