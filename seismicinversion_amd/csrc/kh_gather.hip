// Discrete-wavenumber gather: Bessel-weighted slowness sum and the Hermitian
// inverse FFT to real traces (Kennett_Reflectivity/kennett_gather_gpu.py:56-284,
// float64 twin kennett_gather.py:195-303).
//
//   U[r, k]  = S_k * sum_j R[j, k] * taper_j * p_j * w_k * dp * J0(w_k p_j r)
//   taper_j  = (1 + cos(pi p_j / p_max)) / 2
//   trace_r[n] = Re( sum_k X_k exp(-2 pi i k n / nt) ) * exp(-gamma t_n),  nt = 2 nw,
//   X = [0, U_1..U_{nw-1}, 0, conj(U_{nw-1})..conj(U_1)]     (real frequencies w_k)
//
// The sum runs in float64 (the reference's device path sums in float32; its
// float64 twin is the parity target).  Stage 1: a one-wave workgroup owns 64
// consecutive frequencies x a group of offsets x a chunk of slownesses and
// writes one partial sum per (chunk, r, k); stage 2 adds the chunks in a fixed
// order and applies S_k, so U does not depend on scheduling.  J0 is kh_j0.hpp:
// below x = 12 an interval polynomial (its table staged in LDS), above it
// Hankel's form with 1/x and sqrt(2/(pi x)) formed as products of per-lane
// factors 1/(w r), sqrt(2/(pi w r)) and per-slowness factors 1/p, sqrt(1/p).
// Where a chunk's slownesses are uniform (p_j = p_0 + j dp to a few ulps,
// checked on the device), x = w r p advances by the fixed angle w r dp from
// one slowness to the next, so Hankel's cos and sin of x - pi/4 are rotated
// by that angle (four fused multiply-adds) instead of evaluated: directly
// only where a lane first enters the large regime in the chunk, whose 128
// rotations change the factor by a few units of 2^-53.
#include "kh_complex.hpp"
#include "kh_j0.hpp"
#include "kh_runtime.hpp"

namespace kh {

constexpr int BS_WAVE = 64;   // frequencies per stage-1 workgroup (one wave)
// offsets per stage-1 workgroup: with the rotated Hankel phase, 3 times best
// on the 100-offset gather (1.66 ms; 2: 1.71, 4: 1.79 at 142 VGPRs, three
// waves per SIMD; 5: 2.31); before it, 4 offsets at 119 VGPRs and 10 at 179
// (two waves, 1.3x slower)
constexpr int BS_RG = 3;
constexpr int BS_PCHUNK = 128;  // slownesses per stage-1 workgroup

__global__ void __launch_bounds__(BS_WAVE) bessel_partial_kernel(
    const cd* __restrict__ R, int np, int nwm, const double* __restrict__ p, double dp,
    double p_max, double k0, double dw, const double* __restrict__ offsets, int nr,
    cd* __restrict__ partial, int n_rgroups, int n_wtiles) {
  __shared__ double s_tp[BS_PCHUNK];  // taper_j * p_j
  __shared__ double s_p[BS_PCHUNK];
  __shared__ double s_ip[BS_PCHUNK];   // 1 / |p_j|
  __shared__ double s_sip[BS_PCHUNK];  // sqrt(1 / |p_j|)
  __shared__ double s_j0[J0_XA * (J0_DEG_SMALL + 1)];
  int b = blockIdx.x;
  const int wtile = b % n_wtiles;
  b /= n_wtiles;
  const int rgroup = b % n_rgroups;
  const int chunk = b / n_rgroups;
  const int j0_ = chunk * BS_PCHUNK;
  const int nj = min(BS_PCHUNK, np - j0_);
  const double p0 = p[j0_];
  int lin = dp > 0.0 && p0 >= 0.0;  // the chunk's slownesses are p0 + t dp
  for (int t = threadIdx.x; t < nj; t += blockDim.x) {
    const double pj = p[j0_ + t];
    const double taper = 0.5 * (1.0 + cos(M_PI * pj / p_max));
    const double ip = 1.0 / fabs(pj);
    s_p[t] = pj;
    s_tp[t] = taper * pj;
    s_ip[t] = ip;
    s_sip[t] = sqrt(ip);
    lin &= fabs(pj - fma((double)t, dp, p0)) <= 8.0 * 2.220446049250313e-16 * fabs(pj);
  }
  const bool uniform = __all(lin) != 0;  // one wave: the same in every lane
  for (int t = threadIdx.x; t < J0_XA * (J0_DEG_SMALL + 1); t += blockDim.x)
    s_j0[t] = (&J0_SMALL[0][0])[t];
  __syncthreads();

  const int k = wtile * BS_WAVE + threadIdx.x;
  const int kc = min(k, nwm - 1);
  const double w = (k0 + (double)kc) * dw;
  const int r0 = rgroup * BS_RG;
  const int nrg = min(BS_RG, nr - r0);
  double rr[BS_RG], cu[BS_RG], ca[BS_RG];
#pragma unroll
  for (int q = 0; q < BS_RG; ++q) {
    rr[q] = offsets[min(r0 + q, nr - 1)];
    // 1/x = cu / p and sqrt(2/(pi x)) = ca sqrt(1/p) for x = w p r; inf when
    // w r = 0, where x = 0 takes the interval polynomial
    cu[q] = 1.0 / fabs(w * rr[q]);
    ca[q] = sqrt(0.6366197723675814 * cu[q]);
  }
  cd acc[BS_RG];
#pragma unroll
  for (int q = 0; q < BS_RG; ++q) acc[q] = cd{0.0, 0.0};

  const cd* Rcol = R + (size_t)j0_ * nwm + kc;
  if (uniform) {
    // (hc, hs) = cos, sin of x - pi/4 at the current slowness once the lane
    // is in the large regime (started); (tc, ts) = cos, sin of w r dp
    double hc[BS_RG], hs[BS_RG], tc[BS_RG], ts[BS_RG];
    bool started[BS_RG];
#pragma unroll
    for (int q = 0; q < BS_RG; ++q) {
      sincos_phase(fabs(w * rr[q]) * dp, &ts[q], &tc[q]);
      hc[q] = 1.0;
      hs[q] = 0.0;
      started[q] = false;
    }
    for (int t = 0; t < nj; ++t) {
      const double pj = s_p[t];
      const double wgt = (s_tp[t] * w) * dp;
      const cd Rv = Rcol[(size_t)t * nwm];
      const cd Wv = Rv * wgt;
      const double x0 = w * pj;
      const double ip = s_ip[t], sip = s_sip[t];
#pragma unroll
      for (int q = 0; q < BS_RG; ++q) {
        const double x = fabs(x0 * rr[q]);
        double J;
        if (x < (double)J0_XA) {
          J = j0_small(x, s_j0);
        } else {
          if (!started[q]) {
            j0_phase(x, &hs[q], &hc[q]);
            started[q] = true;
          }
          J = j0_large_cs(cu[q] * ip, ca[q] * sip, hc[q], hs[q]);
          const double c2 = fma(hc[q], tc[q], -(hs[q] * ts[q]));  // advance by w r dp
          hs[q] = fma(hs[q], tc[q], hc[q] * ts[q]);
          hc[q] = c2;
        }
        acc[q].r += J * Wv.r;
        acc[q].i += J * Wv.i;
      }
    }
  } else {
    for (int t = 0; t < nj; ++t) {
      const double pj = s_p[t];
      const double wgt = (s_tp[t] * w) * dp;
      const cd Rv = Rcol[(size_t)t * nwm];
      const cd Wv = Rv * wgt;
      const double x0 = w * pj;
      const double ip = s_ip[t], sip = s_sip[t];
#pragma unroll
      for (int q = 0; q < BS_RG; ++q) {
        const double x = fabs(x0 * rr[q]);
        const double J =
            x < (double)J0_XA ? j0_small(x, s_j0) : j0_large(x, cu[q] * ip, ca[q] * sip);
        acc[q].r += J * Wv.r;
        acc[q].i += J * Wv.i;
      }
    }
  }
  if (k < nwm) {
    cd* out = partial + ((size_t)chunk * nr + r0) * nwm + k;
#pragma unroll
    for (int q = 0; q < BS_RG; ++q)
      if (q < nrg) out[(size_t)q * nwm] = acc[q];
  }
}

__global__ void zero_kernel(cd* __restrict__ x, long n) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) x[idx] = cd{0.0, 0.0};
}

// U[r, k] = S_k * sum_c partial[c, r, k] (chunks in ascending order).
__global__ void bessel_reduce_kernel(const cd* __restrict__ partial, int nchunks, int nr, int nwm,
                                     const cd* __restrict__ S, cd* __restrict__ U) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)nr * nwm;
  if (idx >= n) return;
  cd s = partial[idx];
  for (int c = 1; c < nchunks; ++c) s = s + partial[(size_t)c * n + idx];
  if (S) s = s * S[idx % nwm];
  U[idx] = s;
}

// One workgroup per trace.  The length-nt vector X is Hermitian
// (X[nt-k] = conj(X[k])), so x[n] = sum_k X[k] exp(-2 pi i k n / nt) is real
// and half-length transform gives all of it: with M = nt/2 and
// w = exp(-2 pi i / nt),
//   Z[k] = (X[k] + X[k+M]) + i (X[k] - X[k+M]) w^k,   k < M,
//   z = FFT_M(Z)  =>  z[m] = x[2m] + i x[2m+1],
// because w^(2M) = 1 and w^M = -1 split x into its even and odd samples.
// Z is built in LDS in bit-reversed order and transformed in place by
// radix-2 decimation in time; then trace[n] = x[n] * exp(-gamma n dt).
// Here X[0] = X[M] = 0, X[k] = U[k-1] and X[M+k] = conj(U[M-k-1]) (M = nw).
__global__ void __launch_bounds__(1024) hermitian_trace_kernel(const cd* __restrict__ U, int nw,
                                                               int log2m, double gamma, double dt,
                                                               double* __restrict__ traces) {
  extern __shared__ double smem_f[];
  cd* Z = reinterpret_cast<cd*>(smem_f);
  const int m = 1 << log2m;  // == nw
  const int r = blockIdx.x;
  const cd* Ur = U + (size_t)r * (nw - 1);
  for (int k = threadIdx.x; k < m; k += blockDim.x) {
    const cd a = k >= 1 ? Ur[k - 1] : cd{0.0, 0.0};
    const cd b = k >= 1 ? conj(Ur[m - k - 1]) : cd{0.0, 0.0};
    double sn, cs;
    sincospi(-(double)k / (double)m, &sn, &cs);  // w^k
    const cd v = (a + b) + mul_i((a - b) * cd{cs, sn});
    const int rev = log2m > 0 ? (int)(__brev((unsigned)k) >> (32 - log2m)) : 0;
    Z[rev] = v;
  }
  __syncthreads();
  for (int s = 1; s <= log2m; ++s) {
    const int half = 1 << (s - 1);
    for (int t = threadIdx.x; t < m / 2; t += blockDim.x) {
      const int grp = t >> (s - 1);
      const int pos = t & (half - 1);
      const int i0 = (grp << s) + pos;
      const int i1 = i0 + half;
      double sn, cs;
      sincospi(-(double)pos / (double)half, &sn, &cs);  // exp(-2 pi i pos / 2^s)
      const cd a = Z[i0];
      const cd bb = Z[i1] * cd{cs, sn};
      Z[i0] = a + bb;
      Z[i1] = a - bb;
    }
    __syncthreads();
  }
  double* tr = traces + (size_t)r * 2 * m;
  for (int q = threadIdx.x; q < m; q += blockDim.x) {
    const double t0 = (double)(2 * q) * dt, t1 = (double)(2 * q + 1) * dt;
    tr[2 * q] = Z[q].r * exp(-gamma * t0);
    tr[2 * q + 1] = Z[q].i * exp(-gamma * t1);
  }
}

static int max_lds_bytes() {
  int dev = 0, v = 65536;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  return v;
}

static int ilog2_exact(long n) {
  int l = 0;
  while ((1L << l) < n) ++l;
  return ((1L << l) == n) ? l : -1;
}

}  // namespace kh

using namespace kh;

extern "C" size_t kh_bessel_workspace_bytes(int np, int nw, int nr) {
  if (np <= 0 || nw <= 1 || nr <= 0) return 0;
  const long nchunks = (np + BS_PCHUNK - 1) / BS_PCHUNK;
  return (size_t)nchunks * nr * (nw - 1) * sizeof(cd);
}

extern "C" int kh_bessel_sum(const double* R, int np, int nw, const double* p, double dp,
                             double p_max, double k0, double dw, const double* offsets, int nr,
                             const double* S, double* U, void* workspace, size_t workspace_bytes,
                             void* stream) {
  clear_error();
  KH_REQUIRE(np >= 0 && nw >= 1 && nr >= 0, KH_ERR_INVALID_ARGUMENT, "kh_bessel_sum: bad sizes");
  const int nwm = nw - 1;
  if (nr == 0 || nwm == 0) return KH_OK;
  KH_REQUIRE(U, KH_ERR_INVALID_ARGUMENT, "kh_bessel_sum: U is NULL");
  KH_REQUIRE(p_max > 0.0, KH_ERR_INVALID_ARGUMENT, "kh_bessel_sum: p_max must be > 0");
  hipStream_t st = as_stream(stream);
  if (np == 0) {  // empty sum; a fill kernel (not a memset) keeps the call graph-capturable
    const long n = (long)nr * nwm;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<cd*>(U), n);
    return check_launch("kh_bessel_sum (zero)");
  }
  KH_REQUIRE(R && p && offsets, KH_ERR_INVALID_ARGUMENT, "kh_bessel_sum: NULL pointer");
  const size_t need = kh_bessel_workspace_bytes(np, nw, nr);
  KH_REQUIRE(workspace && workspace_bytes >= need, KH_ERR_INVALID_ARGUMENT,
             "kh_bessel_sum: workspace of %zu bytes needed, %zu given", need, workspace_bytes);
  const int nchunks = (np + BS_PCHUNK - 1) / BS_PCHUNK;
  const int n_rgroups = (nr + BS_RG - 1) / BS_RG;
  const int n_wtiles = (nwm + BS_WAVE - 1) / BS_WAVE;
  const long nblocks = (long)nchunks * n_rgroups * n_wtiles;
  KH_REQUIRE(nblocks <= 0x7fffffffL, KH_ERR_UNSUPPORTED, "kh_bessel_sum: grid too large");
  cd* partial = reinterpret_cast<cd*>(workspace);
  hipLaunchKernelGGL(bessel_partial_kernel, dim3((unsigned)nblocks), dim3(BS_WAVE), 0, st,
                     reinterpret_cast<const cd*>(R), np, nwm, p, dp, p_max, k0, dw, offsets, nr,
                     partial, n_rgroups, n_wtiles);
  KH_TRY(check_launch("kh_bessel_sum (partial)"));
  const long n = (long)nr * nwm;
  hipLaunchKernelGGL(bessel_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     partial, nchunks, nr, nwm, reinterpret_cast<const cd*>(S),
                     reinterpret_cast<cd*>(U));
  return check_launch("kh_bessel_sum (reduce)");
}

extern "C" int kh_hermitian_traces(const double* U, int nr, int nw, double gamma, double T,
                                   double* traces, void* stream) {
  clear_error();
  KH_REQUIRE(nr >= 0 && nw >= 1, KH_ERR_INVALID_ARGUMENT, "kh_hermitian_traces: bad sizes");
  if (nr == 0) return KH_OK;
  // nw = 1 has no frequencies to read (U is empty and may be NULL): zero traces
  KH_REQUIRE(traces && (U || nw == 1), KH_ERR_INVALID_ARGUMENT,
             "kh_hermitian_traces: NULL pointer");
  const long nt = 2L * nw;
  const int log2m = ilog2_exact(nw);
  KH_REQUIRE(log2m >= 0, KH_ERR_INVALID_ARGUMENT, "kh_hermitian_traces: nw=%d is not a power of 2",
             nw);
  const size_t lds = (size_t)nw * sizeof(cd);
  KH_REQUIRE(lds <= (size_t)max_lds_bytes(), KH_ERR_UNSUPPORTED,
             "kh_hermitian_traces: nt=%ld needs %zu bytes of LDS per workgroup", nt, lds);
  const int block = nw < 1024 ? (nw < 64 ? 64 : nw) : 1024;
  const double dt = T / (double)nt;
  hipLaunchKernelGGL(hermitian_trace_kernel, dim3(nr), dim3(block), lds, as_stream(stream),
                     reinterpret_cast<const cd*>(U), nw, log2m, gamma, dt, traces);
  return check_launch("kh_hermitian_traces");
}
