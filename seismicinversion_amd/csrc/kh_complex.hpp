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

__host__ __device__ __forceinline__ cd csqrt(cd z) {
  const double x = z.r, y = z.i;
  if (x == 0.0 && y == 0.0) return cd{0.0, y};
  if (is_inf(y)) return cd{INFINITY, y};
  const double t = sqrt(0.5 * (fabs(x) + hypot(x, y)));
  if (x >= 0.0) return cd{t, y / (2.0 * t)};
  return cd{fabs(y) / (2.0 * t), copysign(t, y)};
}

__host__ __device__ __forceinline__ cd cexp(cd z) {
  double s, c;
  sincos(z.i, &s, &c);
  const double e = exp(z.r);
  return cd{e * c, e * s};
}

__host__ __device__ __forceinline__ double cabs2(cd a) { return a.r * a.r + a.i * a.i; }

}  // namespace kh
