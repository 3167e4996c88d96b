// Bessel function J0 in double precision, for the wavenumber sum of the
// gather (the reference evaluates scipy.special.j0 on the host:
// kennett_gather.py:113-114, kennett_gather_gpu.py:152-154).
//
// Two regimes, with coefficients fitted to mpmath by gen_j0_tables.py:
//   x <  J0_XA: a degree-14 polynomial per unit interval [n, n+1) in
//               t = x - (n + 1/2);
//   x >= J0_XA: Hankel's large-argument form
//               J0(x) = sqrt(2/(pi x)) (P(x) cos(x - pi/4) - Q(x) sin(x - pi/4)),
//               P and x Q polynomials of degree 9 in z = 1/x^2.
// cos and sin of x - pi/4 come from one Cody-Waite reduction by the
// half-integer multiple (m + 1/2) of pi/2 nearest x (x - pi/4 = m pi/2 + r),
// so the shift by pi/4 costs nothing and adds no rounding.
//
// The large regime takes u ~ 1/x and amp ~ sqrt(2/(pi x)) from the caller:
// a sum over a slowness grid forms them from per-lane and per-slowness
// factors with one multiply each instead of a division and a square root.
#pragma once

#include "kh_complex.hpp"
#include "kh_j0_tables.hpp"

namespace kh {

// x in [0, J0_XA); table = J0_SMALL (or a copy of it, e.g. in LDS).
__host__ __device__ __forceinline__ double j0_small(double x, const double* __restrict__ table) {
  const int n = (int)x;
  const double t = (x - (double)n) - 0.5;
  const double* c = table + n * (J0_DEG_SMALL + 1);
  double v = c[J0_DEG_SMALL];
#pragma unroll
  for (int k = J0_DEG_SMALL - 1; k >= 0; --k) v = fma(v, t, c[k]);
  return v;
}

// cos and sin of x - pi/4 for the large regime (x >= J0_XA, finite).
__host__ __device__ __forceinline__ void j0_phase(double x, double* sp, double* cp) {
  double s, c;
  if (x < 1.0e8) {
    // x - pi/4 = m pi/2 + r with r = x - (m + 1/2) pi/2
    const double m = rint(fma(x, 0.6366197723675814, -0.5));
    const double h = m + 0.5;
    double r = fma(-h, 1.5707963267948966, x);
    r = fma(-h, 6.123233995736766e-17, r);
    r = fma(-h, -1.4973849048591698e-33, r);
    sincos_quadrant(r, (int)m, &s, &c);
  } else {
    // cos(x - pi/4) = (cos x + sin x) / sqrt 2, sin(x - pi/4) = (sin x - cos x) / sqrt 2
    double sx, cx;
    sincos(x, &sx, &cx);
    c = (cx + sx) * 0.7071067811865476;
    s = (sx - cx) * 0.7071067811865476;
  }
  *sp = s;
  *cp = c;
}

// Hankel's form given u ~ 1/x, amp ~ sqrt(2/(pi x)) and (c, s) = (cos, sin)
// of x - pi/4.
__host__ __device__ __forceinline__ double j0_large_cs(double u, double amp, double c, double s) {
  const double z = u * u;
  double P = J0_P[J0_DEG_PQ];
  double Q = J0_Q[J0_DEG_PQ];
#pragma unroll
  for (int k = J0_DEG_PQ - 1; k >= 0; --k) {
    P = fma_k(P, z, J0_P[k]);
    Q = fma_k(Q, z, J0_Q[k]);
  }
  Q *= u;
  return amp * fma(P, c, -(Q * s));
}

// x >= J0_XA (finite), u ~ 1/x, amp ~ sqrt(2/(pi x)).
__host__ __device__ __forceinline__ double j0_large(double x, double u, double amp) {
  double s, c;
  j0_phase(x, &s, &c);
  return j0_large_cs(u, amp, c, s);
}

// J0(x) for any double x (even function; J0(+-inf) = 0, NaN gives NaN).
__host__ __device__ __forceinline__ double j0_kh(double x) {
  x = fabs(x);
  if (x < (double)J0_XA) return j0_small(x, &J0_SMALL[0][0]);
  if (x == INFINITY) return 0.0;
  const double u = 1.0 / x;
  return j0_large(x, u, sqrt(0.6366197723675814 * u));
}

}  // namespace kh
