// Complex double arithmetic for the Kennett kernels (gfx950, FP64 vector ALU).
//
// A two-double struct rather than HIP's double2 (whose operators are
// element-wise) so that `a * b` is the complex product.  Division follows
// Smith's scaled algorithm (R. L. Smith, CACM 5(8), 1962), the same method
// NumPy uses for complex128 division; the square root is the principal
// branch with the sign of a signed-zero imaginary part honoured, as C99
// csqrt; exp is exp(x) * (cos y, sin y).
#pragma once

#include <hip/hip_runtime.h>

namespace kh {

struct cd {
  double r, i;
};

__host__ __device__ __forceinline__ cd mk(double r, double i) { return cd{r, i}; }
__host__ __device__ __forceinline__ bool is_inf(double x) { return fabs(x) == INFINITY; }

__host__ __device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.r + b.r, a.i + b.i}; }
__host__ __device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.r - b.r, a.i - b.i}; }
__host__ __device__ __forceinline__ cd operator-(cd a) { return cd{-a.r, -a.i}; }
__host__ __device__ __forceinline__ cd operator*(cd a, cd b) {
  return cd{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__host__ __device__ __forceinline__ cd operator*(cd a, double s) { return cd{a.r * s, a.i * s}; }
__host__ __device__ __forceinline__ cd operator*(double s, cd a) { return cd{a.r * s, a.i * s}; }
__host__ __device__ __forceinline__ cd conj(cd a) { return cd{a.r, -a.i}; }
// Stores a result that no kernel of the launch reads back, marked
// non-temporal so that the output stream does not evict from L2 the tables
// the sweep re-reads.
__device__ __forceinline__ void store_stream(cd* p, const cd& z) {
  __builtin_nontemporal_store(z.r, &p->r);
  __builtin_nontemporal_store(z.i, &p->i);
}
// acc + |a - b|^2: the one association every misfit sum of the library uses,
// so that sums over the same terms in the same order agree bit for bit.
__host__ __device__ __forceinline__ double add_sq_diff(double acc, cd a, cd b) {
  const double dr = a.r - b.r, di = a.i - b.i;
  return acc + (dr * dr + di * di);
}
// c + a*b and c - a*b as two fused multiply-adds per component.
__host__ __device__ __forceinline__ cd cfma(cd a, cd b, cd c) {
  return cd{fma(a.r, b.r, fma(-a.i, b.i, c.r)), fma(a.r, b.i, fma(a.i, b.r, c.i))};
}
__host__ __device__ __forceinline__ cd cfms(cd a, cd b, cd c) {
  return cd{fma(-a.r, b.r, fma(a.i, b.i, c.r)), fma(-a.r, b.i, fma(-a.i, b.r, c.i))};
}
// fma(a, b, c) for a constant c (a polynomial coefficient).  On the device
// the three-operand v_fma_f64 with c read from scalar registers: left to
// itself the compiler keeps such constants in vector registers and emits
// the two-operand v_fmac_f64, whose accumulator is its destination, so every
// Horner step paid a 64-bit register copy of the coefficient first.
__host__ __device__ __forceinline__ double fma_k(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
#else
  return fma(a, b, c);
#endif
}
// multiply by i and by -i: exact component swaps
__host__ __device__ __forceinline__ cd mul_i(cd a) { return cd{-a.i, a.r}; }
__host__ __device__ __forceinline__ cd mul_mi(cd a) { return cd{a.i, -a.r}; }

__host__ __device__ __forceinline__ cd operator/(cd a, cd b) {
  const double br = b.r, bi = b.i;
  if (fabs(br) >= fabs(bi)) {
    if (br == 0.0 && bi == 0.0) return cd{a.r / fabs(br), a.i / fabs(bi)};
    const double rat = bi / br;
    const double scl = 1.0 / (br + bi * rat);
    return cd{(a.r + a.i * rat) * scl, (a.i - a.r * rat) * scl};
  }
  const double rat = br / bi;
  const double scl = 1.0 / (bi + br * rat);
  return cd{(a.r * rat + a.i) * scl, (a.i * rat - a.r) * scl};
}

// 1 / b with the same scaling as the division above.
__host__ __device__ __forceinline__ cd recip(cd b) { return cd{1.0, 0.0} / b; }

// 1 / d for a positive d of order one (far from the overflow and underflow
// thresholds).  On the device: the hardware reciprocal estimate refined by
// two Newton steps r <- r + r (1 - d r), each of which squares the relative
// error, so the result is within an ulp of the rounded quotient; IEEE
// division adds the range scaling and fix-up steps this operand never
// needs.  d = 0 or inf gives NaN.
__host__ __device__ __forceinline__ double recip_pos(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / d;
#endif
}

// 1 / det for the 2x2 determinant of the recursion step, det = 1 - O(|Ru MT|):
// of order one, so |det|^2 neither overflows nor underflows and
// conj(det) / |det|^2 needs one division and no branch (Smith's scaling,
// two divisions behind a data-dependent branch, guards ranges this value
// never reaches).  A singular step (det = 0) gives non-finite values, as the
// reference's division does (NaN here where the division gives inf).
__host__ __device__ __forceinline__ cd recip_det(cd b) {
  const double inv = recip_pos(fma(b.r, b.r, b.i * b.i));
  return cd{b.r * inv, -b.i * inv};
}

__host__ __device__ __forceinline__ cd csqrt(cd z) {
  const double x = z.r, y = z.i;
  if (x == 0.0 && y == 0.0) return cd{0.0, y};
  if (is_inf(y)) return cd{INFINITY, y};
  const double t = sqrt(0.5 * (fabs(x) + hypot(x, y)));
  if (x >= 0.0) return cd{t, y / (2.0 * t)};
  return cd{fabs(y) / (2.0 * t), copysign(t, y)};
}

// (sin, cos) of k pi/2 + r for |r| <= pi/4 (slightly more is harmless): the
// Taylor series of sin r and cos r through r^17 and r^18 (the next terms are
// below 1e-19), then the quadrant q = k mod 4 rotates the pair.
__host__ __device__ __forceinline__ void sincos_quadrant(double r, int k, double* s, double* c) {
  const double z = r * r;
  double ps = 2.8114572543455206e-15;  // (-1)^n / (2n+1)!, n = 8 .. 1
  ps = fma_k(ps, z, -7.647163731819816e-13);
  ps = fma_k(ps, z, 1.6059043836821613e-10);
  ps = fma_k(ps, z, -2.505210838544172e-08);
  ps = fma_k(ps, z, 2.7557319223985893e-06);
  ps = fma_k(ps, z, -0.0001984126984126984);
  ps = fma_k(ps, z, 0.008333333333333333);
  ps = fma_k(ps, z, -0.16666666666666666);
  const double sr = fma(r * z, ps, r);
  double pc = -1.5619206968586225e-16;  // (-1)^n / (2n)!, n = 9 .. 2
  pc = fma_k(pc, z, 4.779477332387385e-14);
  pc = fma_k(pc, z, -1.1470745597729725e-11);
  pc = fma_k(pc, z, 2.08767569878681e-09);
  pc = fma_k(pc, z, -2.755731922398589e-07);
  pc = fma_k(pc, z, 2.48015873015873e-05);
  pc = fma_k(pc, z, -0.001388888888888889);
  pc = fma_k(pc, z, 0.041666666666666664);
  const double hz = 0.5 * z;
  const double w = 1.0 - hz;
  const double cr = w + (((1.0 - w) - hz) + (z * z) * pc);
  const int q = k & 3;
  double ss = (q & 1) ? cr : sr;
  double cc = (q & 1) ? sr : cr;
  if (q == 1 || q == 2) cc = -cc;
  if (q >= 2) ss = -ss;
  *s = ss;
  *c = cc;
}

// (sin x, cos x) for the unit-modulus phase factors exp(i x).
// Reduction by the nearest multiple k of pi/2, with pi/2 split into three
// doubles P1 + P2 + P3 (Cody and Waite, "Software Manual for the Elementary
// Functions", 1980) and each k * Pj removed by one FMA; then
// sincos_quadrant.  The absolute error is a few units of 2^-53 for
// |x| < 1e8, which is what a factor of modulus 1 needs (the argument's own
// rounding, ulp(x), is far larger); larger or non-finite arguments take the
// device library's sincos.
__host__ __device__ __forceinline__ void sincos_phase(double x, double* s, double* c) {
  if (!(fabs(x) < 1.0e8)) {
    sincos(x, s, c);
    return;
  }
  const double k = rint(x * 0.6366197723675814);  // 2 / pi
  double r = fma(-k, 1.5707963267948966, x);
  r = fma(-k, 6.123233995736766e-17, r);
  r = fma(-k, -1.4973849048591698e-33, r);
  double ss;
  sincos_quadrant(r, (int)k, &ss, c);
  *s = (x == 0.0) ? x : ss;  // sin(+-0) = +-0
}

__host__ __device__ __forceinline__ cd cexp(cd z) {
  double s, c;
  sincos_phase(z.i, &s, &c);
  const double e = exp(z.r);
  return cd{e * c, e * s};
}

__host__ __device__ __forceinline__ double cabs2(cd a) { return a.r * a.r + a.i * a.i; }

// A value that every lane of the wave holds, moved to scalar registers
// (v_readfirstlane); the caller guarantees it is the same in every lane.
__host__ __device__ __forceinline__ double wave_uniform(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
#else
  return x;
#endif
}

__host__ __device__ __forceinline__ cd wave_uniform(cd z) {
  return cd{wave_uniform(z.r), wave_uniform(z.i)};
}

}  // namespace kh
