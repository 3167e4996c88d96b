// Global Matrix Method reflectivity on the device: one work-item per
// (slowness, W frequencies); the workgroup stages every layer's vertical
// slownesses and eigenvector matrix at its slowness in LDS, each lane then
// runs the block-Riccati sweep (kh_gmm.hpp) in registers.
#include "kh_gmm_lanes.hpp"
#include "kh_runtime.hpp"

namespace kh {

template <int W, bool GRID>
__global__ void __launch_bounds__(256) gmm_kernel(const double* __restrict__ layers, int nl,
                                                  const double* __restrict__ p,
                                                  const cd* __restrict__ omega, double k0,
                                                  double dw, double gamma, int nw,
                                                  int free_surface, cd* __restrict__ R,
                                                  int ntiles) {
  extern __shared__ double smem_g[];
  LayerPt* s_layer = reinterpret_cast<LayerPt*>(smem_g);
  EigD* s_eig = reinterpret_cast<EigD*>(s_layer + nl);
  LayerStep* s_step = reinterpret_cast<LayerStep*>(s_eig + nl);
  const int ip = blockIdx.x / ntiles;
  const int tile = blockIdx.x - ip * ntiles;
  const double pv = p[ip];
  for (int l = threadIdx.x; l < nl; l += blockDim.x) s_layer[l] = layer_at(layers + (size_t)l * NCOL, l, pv);
  __syncthreads();
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    s_eig[l] = eig_d(pv, s_layer[l]);
    if (GRID && W > 1) s_step[l] = layer_step(s_layer[l], (double)blockDim.x * dw);
  }
  __syncthreads();
  const int base = tile * blockDim.x * W + threadIdx.x;
  cd om[W];
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int w = min(base + k * (int)blockDim.x, nw - 1);
    om[k] = GRID ? cd{(k0 + (double)w) * dw, gamma} : omega[w];
  }
  cd out[W];
  gmm_lanes<W, GRID>(s_layer, s_eig, s_step, nl, om, free_surface, out);
  cd* row = R + (size_t)ip * nw;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int w = base + k * blockDim.x;
    if (w < nw) row[w] = out[k];
  }
}

static size_t gmm_lds_bytes(int nl) {
  return (size_t)nl * (sizeof(LayerPt) + sizeof(EigD) + sizeof(LayerStep));
}

template <bool GRID>
static int launch_gmm(const char* what, const double* layers, int nl, const double* p, int np,
                      const cd* omega, double k0, double dw, double gamma, int nw, int fs, cd* R,
                      hipStream_t st) {
  KH_TRY(check_model(what, layers, nl));
  KH_REQUIRE(np >= 0 && nw >= 0, KH_ERR_INVALID_ARGUMENT, "%s: negative size", what);
  if (np == 0 || nw == 0) return KH_OK;
  KH_REQUIRE(p && R && (GRID || omega), KH_ERR_INVALID_ARGUMENT, "%s: NULL pointer", what);
  // One frequency per work-item (164-170 VGPRs, 2-3 waves per SIMD): measured
  // on MI355X, 20-layer gather grid 8192 p x 1023 omega, 4.73 ms against
  // 5.48 ms with two per work-item (256 VGPRs + AGPRs, one wave per SIMD).
  const int block = 128;
  const long ntiles = (nw + block - 1) / block;
  KH_REQUIRE(ntiles * np <= 0x7fffffffL, KH_ERR_UNSUPPORTED, "%s: grid too large", what);
  const dim3 g((unsigned)(ntiles * np)), b(block);
  hipLaunchKernelGGL((gmm_kernel<1, GRID>), g, b, gmm_lds_bytes(nl), st, layers, nl, p, omega,
                     k0, dw, gamma, nw, fs, R, (int)ntiles);
  return check_launch(what);
}

}  // namespace kh

using namespace kh;

extern "C" int kh_gmm_forward(const double* layers, int nl, const double* p, int np,
                              const double* omega, int nw, int free_surface, double* R,
                              void* stream) {
  clear_error();
  return launch_gmm<false>("kh_gmm_forward", layers, nl, p, np,
                           reinterpret_cast<const cd*>(omega), 0.0, 0.0, 0.0, nw, free_surface,
                           reinterpret_cast<cd*>(R), as_stream(stream));
}

extern "C" int kh_gmm_forward_grid(const double* layers, int nl, const double* p, int np,
                                   double k0, double dw, double gamma, int nw, int free_surface,
                                   double* R, void* stream) {
  clear_error();
  return launch_gmm<true>("kh_gmm_forward_grid", layers, nl, p, np, nullptr, k0, dw, gamma, nw,
                          free_surface, reinterpret_cast<cd*>(R), as_stream(stream));
}
