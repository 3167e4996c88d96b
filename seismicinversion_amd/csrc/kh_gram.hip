// Gram matrix G = X X^T of a real (P, K) row-major matrix on the FP64 matrix
// cores (v_mfma_f64_16x16x4f64).  With X the real view of the Jacobian
// columns J[:, k] (complex -> interleaved pairs), G = Re(J^H J): the
// Gauss-Newton part of the Levenberg-Marquardt Hessian.
//
// Tiling: the output is cut into 16x16 tiles (a, b) with a <= b (G is
// symmetric); each workgroup of one wave owns one tile and one chunk of K,
// streaming 16x4 slabs of the two row panels through the MFMA.  Per-chunk
// partial tiles go to the workspace and a second kernel adds the chunks in
// a fixed order (bitwise reproducible) and mirrors the upper triangle.
//
// Operand layout of v_mfma_f64_16x16x4f64 (wave64, measured on gfx950 with
// single-nonzero inputs): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; result j (0..3) of lane l is D[4 * j + l / 16][l % 16].
#include "kh_complex.hpp"
#include "kh_runtime.hpp"

namespace kh {

constexpr int GR_TILE = 16;
constexpr int GR_KCHUNK = 4096;  // K columns per workgroup

typedef double d4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(64) gram_partial_kernel(const double* __restrict__ X, int P,
                                                          long K, int ntile,
                                                          double* __restrict__ partial) {
  // blockIdx.x -> (pair index of the upper-triangular tile list, chunk)
  const int npairs = ntile * (ntile + 1) / 2;
  const int pair = blockIdx.x % npairs;
  const long chunk = blockIdx.x / npairs;
  int ta = 0, rem = pair;
  while (rem >= ntile - ta) {
    rem -= ntile - ta;
    ++ta;
  }
  const int tb = ta + rem;
  const int lane = threadIdx.x;
  const int row = lane % 16;   // row within the 16-row panel
  const int kk = lane / 16;    // k offset within a 4-wide slab
  const int ra = ta * GR_TILE + row, rb = tb * GR_TILE + row;
  const bool va = ra < P, vb = rb < P;
  const double* Xa = X + (size_t)(va ? ra : 0) * K;
  const double* Xb = X + (size_t)(vb ? rb : 0) * K;
  const long k0 = chunk * GR_KCHUNK;
  const long k1 = (k0 + GR_KCHUNK < K) ? k0 + GR_KCHUNK : K;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  for (long k = k0; k < k1; k += 4) {
    const long kc = k + kk;
    const double a = (va && kc < k1) ? Xa[kc] : 0.0;  // A[row][kk] = X[ra][k + kk]
    const double b = (vb && kc < k1) ? Xb[kc] : 0.0;  // B[kk][row] = X[rb][k + kk]
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
  // acc[j] = D[4 j + lane/16][lane % 16] = sum_k X[ta*16 + 4j + lane/16][k] X[tb*16 + lane%16][k]
  double* out = partial + ((size_t)chunk * npairs + pair) * (GR_TILE * GR_TILE);
  for (int j = 0; j < 4; ++j) out[(4 * j + lane / 16) * GR_TILE + (lane % 16)] = acc[j];
}

__global__ void gram_reduce_kernel(const double* __restrict__ partial, int P, int ntile,
                                   long nchunks, double scale, double* __restrict__ G) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)P * P) return;
  int i = (int)(idx / P), j = (int)(idx % P);
  int a = i, b = j;
  if (a / GR_TILE > b / GR_TILE) {  // lower-triangular tile: read the mirrored entry
    a = j;
    b = i;
  }
  const int ta = a / GR_TILE, tb = b / GR_TILE;
  const int pair = ta * ntile - ta * (ta - 1) / 2 + (tb - ta);
  const int npairs = ntile * (ntile + 1) / 2;
  const size_t off = (size_t)pair * (GR_TILE * GR_TILE) + (a % GR_TILE) * GR_TILE + (b % GR_TILE);
  double s = 0.0;
  for (long c = 0; c < nchunks; ++c) s += partial[(size_t)c * npairs * (GR_TILE * GR_TILE) + off];
  G[idx] = scale * s;
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
  const long ntile = (P + GR_TILE - 1) / GR_TILE;
  const long npairs = ntile * (ntile + 1) / 2;
  const long nchunks = (K + GR_KCHUNK - 1) / GR_KCHUNK;
  return (size_t)nchunks * npairs * GR_TILE * GR_TILE * sizeof(double);
}

extern "C" int kh_gram(const double* X, int P, long K, double scale, double* G, void* workspace,
                       size_t workspace_bytes, void* stream) {
  clear_error();
  KH_REQUIRE(P >= 0 && K >= 0, KH_ERR_INVALID_ARGUMENT, "kh_gram: negative size");
  if (P == 0) return KH_OK;
  KH_REQUIRE(G, KH_ERR_INVALID_ARGUMENT, "kh_gram: G is NULL");
  hipStream_t st = as_stream(stream);
  const int ntile = (P + GR_TILE - 1) / GR_TILE;
  const long nchunks = (K + GR_KCHUNK - 1) / GR_KCHUNK;
  const int npairs = ntile * (ntile + 1) / 2;
  if (K > 0) {
    KH_REQUIRE(X, KH_ERR_INVALID_ARGUMENT, "kh_gram: X is NULL");
    const size_t need = kh_gram_workspace_bytes(P, K);
    KH_REQUIRE(workspace && workspace_bytes >= need, KH_ERR_INVALID_ARGUMENT,
               "kh_gram: workspace of %zu bytes needed, %zu given", need, workspace_bytes);
    const long nblocks = nchunks * npairs;
    KH_REQUIRE(nblocks <= 0x7fffffffL, KH_ERR_UNSUPPORTED, "kh_gram: grid too large");
    hipLaunchKernelGGL(gram_partial_kernel, dim3((unsigned)nblocks), dim3(64), 0, st, X, P, K,
                       ntile, reinterpret_cast<double*>(workspace));
    KH_TRY(check_launch("kh_gram (partial)"));
  }
  const long n = (long)P * P;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<const double*>(workspace), P, ntile, K > 0 ? nchunks : 0,
                     scale, G);
  return check_launch("kh_gram (reduce)");
}
