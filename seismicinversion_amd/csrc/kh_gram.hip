// Gram matrix G = X X^T of a real (P, K) row-major matrix on the FP64 matrix
// cores (v_mfma_f64_16x16x4f64).  With X the real view of the Jacobian
// columns J[:, k] (complex -> interleaved pairs), G = Re(J^H J): the
// Gauss-Newton part of the Levenberg-Marquardt Hessian.
//
// Tiling: rows are cut into 64-row super-tiles (four 16-row MFMA tiles) and
// K into chunks.  One workgroup of four waves owns one super-tile pair
// (A, B), A <= B (G is symmetric), and one chunk: it stages 32-column
// slabs of both 64-row panels in LDS with coalesced loads (every element of
// X is read from memory once per super-tile pair), and wave w accumulates
// the 16 x 64 strip of row tile w against the four column tiles of B.  The
// next slab is loaded into registers while the current one is multiplied,
// so the memory latency overlaps the MFMAs.  The chunk count is chosen from
// (P, K) alone, about 1024 workgroups in all, so the launch fills the
// device for thin and for tall X alike.  Per-chunk
// partial 64 x 64 blocks go to the workspace; the reduction adds the chunks
// of each entry in a fixed order (bitwise reproducible) and reads the upper
// triangle for both halves, so G is exactly symmetric.
//
// Operand layout of v_mfma_f64_16x16x4f64 (wave64, measured on gfx950 with
// single-nonzero inputs): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; result j (0..3) of lane l is D[4 * j + l / 16][l % 16].
#include <algorithm>

#include "kh_complex.hpp"
#include "kh_runtime.hpp"

namespace kh {

constexpr int GR_TILE = 16;
constexpr int GR_SUP = 64;          // rows per super-tile
constexpr int GR_KS = 32;           // K columns per LDS slab
constexpr int GR_LD = GR_KS + 1;    // LDS row stride (odd: spreads the banks)
constexpr int GR_BLOCK = 256;       // four waves, one row tile each
constexpr int GR_RED = 16;          // threads per entry in the reduction

typedef double d4 __attribute__((ext_vector_type(4)));

struct GramPlan {
  int nsup, nsp;  // super-tiles, super-tile pairs (A <= B)
  long kc, nchunks;
};

static GramPlan gram_plan(int P, long K) {
  GramPlan g;
  g.nsup = (P + GR_SUP - 1) / GR_SUP;
  g.nsp = g.nsup * (g.nsup + 1) / 2;
  const long slabs = std::max(1L, (K + GR_KS - 1) / GR_KS);
  const long want = std::max(1L, std::min(256L, 1024L / g.nsp));
  const long nch = std::min(slabs, want);
  g.kc = ((K + nch - 1) / nch + GR_KS - 1) / GR_KS * GR_KS;
  g.nchunks = K > 0 ? (K + g.kc - 1) / g.kc : 0;
  return g;
}

__device__ __forceinline__ int sp_index(int A, int B, int nsup) {
  return A * nsup - A * (A - 1) / 2 + (B - A);
}

__global__ void __launch_bounds__(GR_BLOCK) gram_partial_kernel(const double* __restrict__ X,
                                                                int P, long K, int nsup, long kc,
                                                                double* __restrict__ partial) {
  __shared__ double sA[GR_SUP * GR_LD];
  __shared__ double sB[GR_SUP * GR_LD];
  const int nsp = nsup * (nsup + 1) / 2;
  const int sp = (int)(blockIdx.x % (unsigned)nsp);
  const long chunk = blockIdx.x / (unsigned)nsp;
  int A = 0, rem = sp;
  while (rem >= nsup - A) {
    rem -= nsup - A;
    ++A;
  }
  const int B = A + rem;
  const long k0 = chunk * kc;
  const long k1 = (k0 + kc < K) ? k0 + kc : K;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = lane % 16, kk = lane / 16;
  d4 acc[4];
#pragma unroll
  for (int bt = 0; bt < 4; ++bt) acc[bt] = d4{0.0, 0.0, 0.0, 0.0};
  // a slab is GR_SUP x GR_KS elements of each panel, GR_PER per thread
  constexpr int GR_PER = GR_SUP * GR_KS / GR_BLOCK;
  static_assert(GR_PER * GR_BLOCK == GR_SUP * GR_KS, "whole slabs per thread");
  double va[GR_PER], vb[GR_PER];
  auto fetch = [&](long ks) {
#pragma unroll
    for (int q = 0; q < GR_PER; ++q) {
      const int e = (int)threadIdx.x + q * GR_BLOCK;
      const int r = e / GR_KS, c = e - (e / GR_KS) * GR_KS;  // a row's slab is contiguous
      const long k = ks + c;
      const int ra = A * GR_SUP + r, rb = B * GR_SUP + r;
      va[q] = (ra < P && k < k1) ? X[(size_t)ra * K + k] : 0.0;
      vb[q] = (rb < P && k < k1) ? X[(size_t)rb * K + k] : 0.0;
    }
  };
  if (k0 < k1) fetch(k0);
  for (long ks = k0; ks < k1; ks += GR_KS) {
#pragma unroll
    for (int q = 0; q < GR_PER; ++q) {
      const int e = (int)threadIdx.x + q * GR_BLOCK;
      const int r = e / GR_KS, c = e - (e / GR_KS) * GR_KS;
      sA[r * GR_LD + c] = va[q];
      sB[r * GR_LD + c] = vb[q];
    }
    __syncthreads();
    if (ks + GR_KS < k1) fetch(ks + GR_KS);  // in flight during the MFMAs below
#pragma unroll
    for (int s = 0; s < GR_KS; s += 4) {
      const double a = sA[(w * GR_TILE + row) * GR_LD + s + kk];
#pragma unroll
      for (int bt = 0; bt < 4; ++bt) {
        const double b = sB[(bt * GR_TILE + row) * GR_LD + s + kk];
        acc[bt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[bt], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // acc[bt][j] = sum_k X[A*64 + 16w + 4j + lane/16][k] X[B*64 + 16bt + lane%16][k]
  double* out = partial + ((size_t)chunk * nsp + sp) * (GR_SUP * GR_SUP);
#pragma unroll
  for (int bt = 0; bt < 4; ++bt)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      out[(w * GR_TILE + 4 * j + kk) * GR_SUP + bt * GR_TILE + row] = acc[bt][j];
}

// GR_RED threads per entry (i, j): thread s adds chunks s, s + GR_RED, ...
// in order, then a fixed butterfly over the GR_RED lanes; the first lane's
// total is written.  Entries below the diagonal read their mirror's partials.
__global__ void __launch_bounds__(256) gram_reduce_kernel(const double* __restrict__ partial,
                                                          int P, int nsup, long nchunks,
                                                          double scale, double* __restrict__ G) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long o = t / GR_RED;
  const int s = (int)(t - o * GR_RED);
  const bool valid = o < (long)P * P;
  const int i = valid ? (int)(o / P) : 0, j = valid ? (int)(o % P) : 0;
  const int a = i < j ? i : j, b = i < j ? j : i;
  const int A = a / GR_SUP, B = b / GR_SUP;
  const int nsp = nsup * (nsup + 1) / 2;
  const size_t off = (size_t)sp_index(A, B, nsup) * (GR_SUP * GR_SUP) +
                     (size_t)(a % GR_SUP) * GR_SUP + (b % GR_SUP);
  double acc = 0.0;
  if (valid)
    for (long c = s; c < nchunks; c += GR_RED) acc += partial[(size_t)c * nsp * (GR_SUP * GR_SUP) + off];
#pragma unroll
  for (int m = GR_RED / 2; m > 0; m >>= 1) acc += __shfl_xor(acc, m, 64);
  if (valid && s == 0) G[o] = scale * acc;
}

// chi[j] = sum_w |R[j, w] - R_obs[j, w]|^2, one workgroup per slowness row:
// each work-item sums its strided frequencies in ascending order, then a
// fixed LDS tree adds the 256 partial sums, so the value does not depend on
// timing (and a captured graph replays it exactly).
constexpr int RN_BLOCK = 256;

__global__ void __launch_bounds__(RN_BLOCK) residual_norm_kernel(const cd* __restrict__ R,
                                                                 const cd* __restrict__ R_obs,
                                                                 int nw, double* __restrict__ chi) {
  __shared__ double s[RN_BLOCK];
  const size_t row = blockIdx.x;
  const cd* a = R + row * nw;
  const cd* b = R_obs + row * nw;
  double acc = 0.0;
  for (int w = threadIdx.x; w < nw; w += RN_BLOCK) acc = add_sq_diff(acc, a[w], b[w]);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = RN_BLOCK / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) chi[row] = s[0];
}

}  // namespace kh

using namespace kh;

extern "C" int kh_residual_norms(const double* R, const double* R_obs, int np, int nw,
                                 double* chi, void* stream) {
  clear_error();
  KH_REQUIRE(np >= 0 && nw >= 0, KH_ERR_INVALID_ARGUMENT, "kh_residual_norms: negative size");
  if (np == 0) return KH_OK;
  KH_REQUIRE(R && R_obs && chi, KH_ERR_INVALID_ARGUMENT, "kh_residual_norms: NULL pointer");
  hipLaunchKernelGGL(residual_norm_kernel, dim3((unsigned)np), dim3(RN_BLOCK), 0,
                     as_stream(stream), reinterpret_cast<const cd*>(R),
                     reinterpret_cast<const cd*>(R_obs), nw, chi);
  return check_launch("kh_residual_norms");
}

extern "C" size_t kh_gram_workspace_bytes(int P, long K) {
  if (P <= 0 || K <= 0) return 0;
  const GramPlan g = gram_plan(P, K);
  return (size_t)g.nchunks * g.nsp * GR_SUP * GR_SUP * sizeof(double);
}

extern "C" int kh_gram(const double* X, int P, long K, double scale, double* G, void* workspace,
                       size_t workspace_bytes, void* stream) {
  clear_error();
  KH_REQUIRE(P >= 0 && K >= 0, KH_ERR_INVALID_ARGUMENT, "kh_gram: negative size");
  if (P == 0) return KH_OK;
  KH_REQUIRE(G, KH_ERR_INVALID_ARGUMENT, "kh_gram: G is NULL");
  hipStream_t st = as_stream(stream);
  const GramPlan g = gram_plan(P, K);
  if (K > 0) {
    KH_REQUIRE(X, KH_ERR_INVALID_ARGUMENT, "kh_gram: X is NULL");
    const size_t need = kh_gram_workspace_bytes(P, K);
    KH_REQUIRE(workspace && workspace_bytes >= need, KH_ERR_INVALID_ARGUMENT,
               "kh_gram: workspace of %zu bytes needed, %zu given", need, workspace_bytes);
    const long nblocks = g.nchunks * g.nsp;
    KH_REQUIRE(nblocks <= 0x7fffffffL, KH_ERR_UNSUPPORTED, "kh_gram: grid too large");
    hipLaunchKernelGGL(gram_partial_kernel, dim3((unsigned)nblocks), dim3(GR_BLOCK), 0, st, X, P,
                       K, g.nsup, g.kc, reinterpret_cast<double*>(workspace));
    KH_TRY(check_launch("kh_gram (partial)"));
  }
  const long nthreads = (long)P * P * GR_RED;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0,
                     st, reinterpret_cast<const double*>(workspace), P, g.nsup,
                     K > 0 ? g.nchunks : 0, scale, G);
  return check_launch("kh_gram (reduce)");
}
