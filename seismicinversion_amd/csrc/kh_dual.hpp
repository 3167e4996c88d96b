// Forward-mode dual numbers over complex doubles: z = v + e*d with e^2 = 0.
//
// Every operation propagates the tangent d by the derivative of the primal
// operation (holomorphic for +, -, *, /, sqrt, exp), so a formula written
// once as a template over the scalar type yields, for
//   Z = cd              its value,
//   Z = Dual<cd>        its value and one directional derivative,
//   Z = Dual<Dual<cd>>  additionally a second directional derivative
//                       (used for d(dc/dx)[dm] in Hessian-vector products).
// Primal parts are computed by exactly the operations of the plain `cd`
// code, so a dual evaluation reproduces the plain value.
#pragma once

#include "kh_complex.hpp"

namespace kh {

template <class T>
struct Dual {
  T v, d;
};
using dcd = Dual<cd>;
using ddcd = Dual<dcd>;

template <class T>
__host__ __device__ __forceinline__ Dual<T> operator+(const Dual<T>& a, const Dual<T>& b) {
  return {a.v + b.v, a.d + b.d};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> operator-(const Dual<T>& a, const Dual<T>& b) {
  return {a.v - b.v, a.d - b.d};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> operator-(const Dual<T>& a) {
  return {-a.v, -a.d};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> operator*(const Dual<T>& a, const Dual<T>& b) {
  return {a.v * b.v, cfma(a.v, b.d, a.d * b.v)};
}
// c + a*b and c - a*b with fused tangents (see cfma / cfms on cd).
template <class T>
__host__ __device__ __forceinline__ Dual<T> cfma(const Dual<T>& a, const Dual<T>& b,
                                                 const Dual<T>& c) {
  return {cfma(a.v, b.v, c.v), cfma(a.v, b.d, cfma(a.d, b.v, c.d))};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> cfms(const Dual<T>& a, const Dual<T>& b,
                                                 const Dual<T>& c) {
  return {cfms(a.v, b.v, c.v), cfms(a.v, b.d, cfms(a.d, b.v, c.d))};
}
// A dual times (or fused with) a complex value without tangent: the full
// dual operation with that tangent 0, less the products by 0.  Each tangent
// is the expression the full operation reduces to (an fma by 0 returns its
// addend), so for finite operands the results are the full operation's bit
// for bit (only the sign of an exact zero may differ).
__host__ __device__ __forceinline__ dcd operator*(const dcd& a, const cd& b) {
  return {a.v * b, a.d * b};
}
__host__ __device__ __forceinline__ dcd operator*(const cd& a, const dcd& b) {
  return {a * b.v, cfma(a, b.d, cd{0.0, 0.0})};
}
__host__ __device__ __forceinline__ dcd cfma(const dcd& a, const cd& b, const dcd& c) {
  return {cfma(a.v, b, c.v), cfma(a.d, b, c.d)};
}
__host__ __device__ __forceinline__ dcd cfma(const cd& a, const dcd& b, const dcd& c) {
  return {cfma(a, b.v, c.v), cfma(a, b.d, c.d)};
}
__host__ __device__ __forceinline__ dcd cfms(const dcd& a, const cd& b, const dcd& c) {
  return {cfms(a.v, b, c.v), cfms(a.d, b, c.d)};
}
__host__ __device__ __forceinline__ dcd cfms(const cd& a, const dcd& b, const dcd& c) {
  return {cfms(a, b.v, c.v), cfms(a, b.d, c.d)};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> operator*(const Dual<T>& a, double s) {
  return {a.v * s, a.d * s};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> operator*(double s, const Dual<T>& a) {
  return {a.v * s, a.d * s};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> operator/(const Dual<T>& a, const Dual<T>& b) {
  const T q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> recip(const Dual<T>& b) {
  const T r = recip(b.v);
  return {r, -(r * r) * b.d};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> recip_det(const Dual<T>& b) {
  const T r = recip_det(b.v);
  return {r, -(r * r) * b.d};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> conj(const Dual<T>& a) {
  return {conj(a.v), conj(a.d)};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> mul_i(const Dual<T>& a) {
  return {mul_i(a.v), mul_i(a.d)};
}

__host__ __device__ __forceinline__ const cd& prim(const cd& z) { return z; }
template <class T>
__host__ __device__ __forceinline__ const cd& prim(const Dual<T>& z) {
  return prim(z.v);
}

template <class T>
__host__ __device__ __forceinline__ Dual<T> csqrt(const Dual<T>& z) {
  const T s = csqrt(z.v);
  const cd sp = prim(s);
  if (sp.r == 0.0 && sp.i == 0.0) return {s, z.d * 0.0};  // no derivative at 0
  return {s, z.d / (s * 2.0)};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> cexp(const Dual<T>& z) {
  const T e = cexp(z.v);
  return {e, e * z.d};
}

// Division by a real-valued denominator (imaginary part exactly 0): the
// primal is two real divisions, as the plain code does for real quotients.
__host__ __device__ __forceinline__ cd div_real(const cd& a, const cd& b) {
  return cd{a.r / b.r, a.i / b.r};
}
template <class T>
__host__ __device__ __forceinline__ Dual<T> div_real(const Dual<T>& a, const Dual<T>& b) {
  const T q = div_real(a.v, b.v);
  return {q, div_real(a.d - q * b.d, b.v)};
}

// Real part, keeping the dual structure (imaginary parts of every level dropped).
__host__ __device__ __forceinline__ cd re_part(const cd& z) { return cd{z.r, 0.0}; }
template <class T>
__host__ __device__ __forceinline__ Dual<T> re_part(const Dual<T>& z) {
  return {re_part(z.v), re_part(z.d)};
}

// Constants lifted into a scalar type (zero tangents).
template <class Z>
struct Lift;
template <>
struct Lift<cd> {
  __host__ __device__ static cd c(cd x) { return x; }
};
template <class T>
struct Lift<Dual<T>> {
  __host__ __device__ static Dual<T> c(cd x) { return {Lift<T>::c(x), Lift<T>::c(cd{0.0, 0.0})}; }
};
template <class Z>
__host__ __device__ __forceinline__ Z lift(double x) {
  return Lift<Z>::c(cd{x, 0.0});
}
template <class Z>
__host__ __device__ __forceinline__ Z lift_c(cd x) {
  return Lift<Z>::c(x);
}

// A real input x with first-level tangent dx (and, for nested duals,
// second-level tangent dx2 and mixed term 0).
__host__ __device__ __forceinline__ dcd seed1(double x, double dx) {
  return dcd{cd{x, 0.0}, cd{dx, 0.0}};
}
__host__ __device__ __forceinline__ ddcd seed2(double x, double d_inner, double d_outer) {
  return ddcd{dcd{cd{x, 0.0}, cd{d_inner, 0.0}}, dcd{cd{d_outer, 0.0}, cd{0.0, 0.0}}};
}

}  // namespace kh
