// Forward Kennett reflectivity over a (slowness, frequency) grid.
//
// One work-item walks one (p, omega) lane through all interfaces, bottom to
// top (Kennett_Reflectivity/kennett_reflectivity_gpu.py:216-272):
//
//   RRd <- 0 below the deepest interface
//   for each interface i (deepest first), with b = i + 1 the layer below:
//     MT  = E_b RRd E_b,          E_b = diag(exp(i w eta_b h_b), exp(i w nu_b h_b))
//     U   = (I - Ru_i MT)^-1      (closed-form 2x2 inverse)
//     RRd = Rd_i + Tu_i MT U Td_i
//   R = exp(2 i w eta_0 h_0) RRd_00            (+ free surface: R / (1 + R))
//
// A workgroup owns one slowness and a tile of frequencies.  Its prologue
// stages the layer slownesses and the frequency-independent interface
// coefficients of that slowness in LDS (one work-item per layer, then per
// interface); every lane then reads them as wave-uniform LDS broadcasts.
// Phase factors and the 2x2 algebra stay in registers; each store writes a
// coalesced run of complex128 along omega.
#include "kh_physics.hpp"
#include "kh_runtime.hpp"
#include "kh_sweep.hpp"

namespace kh {

__global__ void vertical_slowness_kernel(const double* __restrict__ layers, int nl,
                                         const double* __restrict__ p, int np,
                                         cd* __restrict__ eta, cd* __restrict__ neta) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nl * np) return;
  const int l = (int)(idx / np);
  const int j = (int)(idx % np);
  const LayerPt L = layer_at(layers + (size_t)l * NCOL, l, p[j]);
  eta[idx] = L.eta;
  neta[idx] = L.neta;
}

__global__ void interface_kernel(const double* __restrict__ layers, int nl,
                                 const double* __restrict__ p, int np, cd* __restrict__ Rd,
                                 cd* __restrict__ Ru, cd* __restrict__ Tu, cd* __restrict__ Td) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)(nl - 1) * np) return;
  const int i = (int)(idx / np);
  const int j = (int)(idx % np);
  const double pv = p[j];
  const LayerPt up = layer_at(layers + (size_t)i * NCOL, i, pv);
  const LayerPt dn = layer_at(layers + (size_t)(i + 1) * NCOL, i + 1, pv);
  const Iface f = interface_at(i, pv, up, dn);
  cd* rd = Rd + idx * 4;
  cd* ru = Ru + idx * 4;
  cd* tu = Tu + idx * 4;
  cd* td = Td + idx * 4;
  rd[0] = f.rd00; rd[1] = f.rd01; rd[2] = f.rd01; rd[3] = f.rd11;
  ru[0] = f.ru00; ru[1] = f.ru01; ru[2] = f.ru01; ru[3] = f.ru11;
  td[0] = f.td00; td[1] = f.td01; td[2] = f.td10; td[3] = f.td11;
  tu[0] = f.td00; tu[1] = f.td10; tu[2] = f.td01; tu[3] = f.td11;
}

// Coefficients of n independent interfaces given their vertical slownesses
// (the reference's scalar API, scattering_matrices.py:42-298).
__global__ void interface_direct_kernel(int n, int ocean, const double* __restrict__ p,
                                        const cd* __restrict__ eta1, const cd* __restrict__ neta1,
                                        const double* __restrict__ rho1,
                                        const cd* __restrict__ beta1, const cd* __restrict__ eta2,
                                        const cd* __restrict__ neta2,
                                        const double* __restrict__ rho2,
                                        const cd* __restrict__ beta2, cd* __restrict__ Rd,
                                        cd* __restrict__ Ru, cd* __restrict__ Tu,
                                        cd* __restrict__ Td) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  LayerPt up, dn;
  up.eta = eta1[k];
  up.neta = ocean ? cd{0.0, 0.0} : neta1[k];
  up.beta_c = ocean ? cd{0.0, 0.0} : beta1[k];
  up.rho = rho1[k];
  dn.eta = eta2[k];
  dn.neta = neta2[k];
  dn.beta_c = beta2[k];
  dn.rho = rho2[k];
  up.h = dn.h = 0.0;
  up.halfspace = dn.halfspace = 0;
  up.pad_ = dn.pad_ = 0;
  const Iface f = ocean ? ocean_bottom(p[k], up, dn) : solid_solid(p[k], up, dn);
  cd* rd = Rd + k * 4;
  cd* ru = Ru + k * 4;
  cd* tu = Tu + k * 4;
  cd* td = Td + k * 4;
  rd[0] = f.rd00; rd[1] = f.rd01; rd[2] = f.rd01; rd[3] = f.rd11;
  ru[0] = f.ru00; ru[1] = f.ru01; ru[2] = f.ru01; ru[3] = f.ru11;
  td[0] = f.td00; td[1] = f.td01; td[2] = f.td10; td[3] = f.td11;
  tu[0] = f.td00; tu[1] = f.td10; tu[2] = f.td01; tu[3] = f.td11;
}

// grid.x = np * ntiles; block b covers slowness b / ntiles and the
// frequencies tile * blockDim * W + threadIdx.x + k * blockDim, k < W.
// GRID: omega_j = (k0 + j) * dw + i * gamma (omega unused); otherwise omega[j].
template <int W, bool GRID>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GRID ? 2 : 1))) forward_kernel(
    const double* __restrict__ layers, int nl, const double* __restrict__ p,
    const cd* __restrict__ omega, double k0, double dw, double gamma, int nw, int free_surface,
    cd* __restrict__ R, int ntiles) {
  extern __shared__ double smem_d[];
  LayerPt* s_layer = reinterpret_cast<LayerPt*>(smem_d);
  Iface* s_if = reinterpret_cast<Iface*>(s_layer + nl);
  LayerStep* s_step = reinterpret_cast<LayerStep*>(s_if + (nl - 1));

  const int ip = blockIdx.x / ntiles;
  const int tile = blockIdx.x - ip * ntiles;
  stage_slowness(layers, nl, p[ip], s_layer, s_if);
  if (GRID && W > 1) {
    const double shift = (double)blockDim.x * dw;
    for (int l = threadIdx.x; l < nl; l += blockDim.x) s_step[l] = layer_step(s_layer[l], shift);
    __syncthreads();
  }

  const int base = tile * blockDim.x * W + threadIdx.x;
  cd om[W];
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int w = min(base + k * (int)blockDim.x, nw - 1);
    om[k] = GRID ? cd{(k0 + (double)w) * dw, gamma} : omega[w];
  }
  cd out[W];
  sweep_lanes<W, GRID>(s_layer, s_if, s_step, nl, om, free_surface, out);
  cd* Rrow = R + (size_t)ip * nw;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int w = base + k * blockDim.x;
    if (w < nw) store_stream(Rrow + w, out[k]);
  }
}

size_t forward_lds_bytes(int nl) {
  return (size_t)nl * (sizeof(LayerPt) + sizeof(LayerStep)) + (size_t)(nl - 1) * sizeof(Iface);
}

// Default launch shape, measured on MI355X (clocks warmed up, interleaved
// rounds) on the 20-layer gather grid, 8192 p x 1023 omega, free surface
// (work-items per workgroup x frequencies per work-item):
//   uniform grid: 64 x 8 1.021 ms, 128 x 8 1.024, 64 x 4 1.178, 128 x 4 1.183,
//                 256 x 4 1.200, 256 x 8 2.009 ms.
// On the uniform grid eight frequencies per work-item share one phase
// exponential pair per layer and the step factors; the kernel then takes one
// wave per SIMD (276 registers) and still keeps the FP64 pipes busier than
// two waves of four frequencies.  Picks 64 x 8 on the uniform grid when that
// keeps >= 1024 workgroups and leaves less than a quarter of the last tile's
// lanes idle; otherwise 128 work-items and the most frequencies per
// work-item (4, 2, 1) that meets the same two conditions.
static void default_shape(int np, int nw, bool grid, int* block, int* wpt) {
  auto fits = [&](int blk, int c) {
    const long span = (long)c * blk;
    const long tiles = (nw + span - 1) / span;
    const long idle = tiles * span - nw;
    return tiles * (long)np >= 1024 && 4 * idle < span;
  };
  if (grid && fits(64, 8)) {
    *block = 64;
    *wpt = 8;
    return;
  }
  *block = nw <= 64 ? 64 : 128;
  *wpt = 1;
  const int cands[2] = {4, 2};
  for (int c : cands) {
    if (fits(*block, c)) {
      *wpt = c;
      break;
    }
  }
}

static bool valid_wpt(int w) { return w == 1 || w == 2 || w == 3 || w == 4 || w == 8; }

template <bool GRID>
static int launch_forward(const double* layers, int nl, const double* p, int np, const cd* omega,
                          double k0, double dw, double gamma, int nw, int free_surface, cd* R,
                          int block, int wpt, hipStream_t stream) {
  const long span = (long)block * wpt;
  const long ntiles = (nw + span - 1) / span;
  const long nblocks = ntiles * (long)np;
  KH_REQUIRE(nblocks <= 0x7fffffffL, KH_ERR_UNSUPPORTED, "kh_forward: grid too large");
  const size_t lds = forward_lds_bytes(nl);
  const dim3 g((unsigned)nblocks), b(block);
  const int nt = (int)ntiles;
#define KH_FWD_CASE(WV)                                                                       \
  case WV:                                                                                   \
    hipLaunchKernelGGL((forward_kernel<WV, GRID>), g, b, lds, stream, layers, nl, p, omega, k0, \
                       dw, gamma, nw, free_surface, R, nt);                                  \
    break;
  switch (wpt) {
    KH_FWD_CASE(1)
    KH_FWD_CASE(2)
    KH_FWD_CASE(3)
    KH_FWD_CASE(4)
    KH_FWD_CASE(8)
    default:
      break;
  }
#undef KH_FWD_CASE
  return check_launch("kh_forward");
}

static int check_forward_args(const char* what, const double* layers, int nl, const double* p,
                              int np, int nw, const void* R, int block, int wpt) {
  KH_TRY(check_model(what, layers, nl));
  KH_REQUIRE(np >= 0 && nw >= 0, KH_ERR_INVALID_ARGUMENT, "%s: negative size", what);
  if (np == 0 || nw == 0) return KH_OK;
  KH_REQUIRE(p && R, KH_ERR_INVALID_ARGUMENT, "%s: NULL pointer", what);
  KH_REQUIRE(block == 64 || block == 128 || block == 256, KH_ERR_INVALID_ARGUMENT,
             "%s: block %d not in {64, 128, 256}", what, block);
  KH_REQUIRE(valid_wpt(wpt), KH_ERR_INVALID_ARGUMENT,
             "%s: omega_per_item %d not in {1,2,3,4,8}", what, wpt);
  return KH_OK;
}

}  // namespace kh

using namespace kh;

extern "C" int kh_vertical_slowness(const double* layers, int nl, const double* p, int np,
                                    double* eta, double* neta, void* stream) {
  clear_error();
  KH_TRY(check_model("kh_vertical_slowness", layers, nl));
  KH_REQUIRE(np >= 0, KH_ERR_INVALID_ARGUMENT, "kh_vertical_slowness: np < 0");
  if (np == 0) return KH_OK;
  KH_REQUIRE(p && eta && neta, KH_ERR_INVALID_ARGUMENT, "kh_vertical_slowness: NULL pointer");
  const long n = (long)nl * np;
  const int blk = 256;
  hipLaunchKernelGGL(vertical_slowness_kernel, dim3((unsigned)((n + blk - 1) / blk)), dim3(blk),
                     0, as_stream(stream), layers, nl, p, np, reinterpret_cast<cd*>(eta),
                     reinterpret_cast<cd*>(neta));
  return check_launch("kh_vertical_slowness");
}

extern "C" int kh_interface_coefficients(const double* layers, int nl, const double* p, int np,
                                         double* Rd, double* Ru, double* Tu, double* Td,
                                         void* stream) {
  clear_error();
  KH_TRY(check_model("kh_interface_coefficients", layers, nl));
  KH_REQUIRE(np >= 0, KH_ERR_INVALID_ARGUMENT, "kh_interface_coefficients: np < 0");
  if (np == 0) return KH_OK;
  KH_REQUIRE(p && Rd && Ru && Tu && Td, KH_ERR_INVALID_ARGUMENT,
             "kh_interface_coefficients: NULL pointer");
  const long n = (long)(nl - 1) * np;
  const int blk = 256;
  hipLaunchKernelGGL(interface_kernel, dim3((unsigned)((n + blk - 1) / blk)), dim3(blk), 0,
                     as_stream(stream), layers, nl, p, np, reinterpret_cast<cd*>(Rd),
                     reinterpret_cast<cd*>(Ru), reinterpret_cast<cd*>(Tu),
                     reinterpret_cast<cd*>(Td));
  return check_launch("kh_interface_coefficients");
}

extern "C" int kh_interfaces_direct(int n, int ocean, const double* p, const double* eta1,
                                    const double* neta1, const double* rho1, const double* beta1,
                                    const double* eta2, const double* neta2, const double* rho2,
                                    const double* beta2, double* Rd, double* Ru, double* Tu,
                                    double* Td, void* stream) {
  clear_error();
  KH_REQUIRE(n >= 0, KH_ERR_INVALID_ARGUMENT, "kh_interfaces_direct: n < 0");
  if (n == 0) return KH_OK;
  KH_REQUIRE(p && eta1 && rho1 && eta2 && neta2 && rho2 && beta2 && Rd && Ru && Tu && Td &&
                 (ocean || (neta1 && beta1)),
             KH_ERR_INVALID_ARGUMENT, "kh_interfaces_direct: NULL pointer");
  const int blk = 256;
  hipLaunchKernelGGL(interface_direct_kernel, dim3((unsigned)((n + blk - 1) / blk)), dim3(blk), 0,
                     as_stream(stream), n, ocean, p, reinterpret_cast<const cd*>(eta1),
                     reinterpret_cast<const cd*>(neta1), rho1, reinterpret_cast<const cd*>(beta1),
                     reinterpret_cast<const cd*>(eta2), reinterpret_cast<const cd*>(neta2), rho2,
                     reinterpret_cast<const cd*>(beta2), reinterpret_cast<cd*>(Rd),
                     reinterpret_cast<cd*>(Ru), reinterpret_cast<cd*>(Tu),
                     reinterpret_cast<cd*>(Td));
  return check_launch("kh_interfaces_direct");
}

extern "C" int kh_forward_ex(const double* layers, int nl, const double* p, int np,
                             const double* omega, int nw, int free_surface, double* R, int block,
                             int omega_per_item, void* stream) {
  clear_error();
  KH_TRY(check_forward_args("kh_forward", layers, nl, p, np, nw, R, block, omega_per_item));
  if (np == 0 || nw == 0) return KH_OK;
  KH_REQUIRE(omega, KH_ERR_INVALID_ARGUMENT, "kh_forward: omega is NULL");
  return launch_forward<false>(layers, nl, p, np, reinterpret_cast<const cd*>(omega), 0.0, 0.0,
                               0.0, nw, free_surface, reinterpret_cast<cd*>(R), block,
                               omega_per_item, as_stream(stream));
}

extern "C" int kh_forward(const double* layers, int nl, const double* p, int np,
                          const double* omega, int nw, int free_surface, double* R,
                          void* stream) {
  int block, wpt;
  default_shape(np, nw, false, &block, &wpt);
  return kh_forward_ex(layers, nl, p, np, omega, nw, free_surface, R, block, wpt, stream);
}

extern "C" int kh_forward_grid_ex(const double* layers, int nl, const double* p, int np, double k0,
                                  double dw, double gamma, int nw, int free_surface, double* R,
                                  int block, int omega_per_item, void* stream) {
  clear_error();
  KH_TRY(check_forward_args("kh_forward_grid", layers, nl, p, np, nw, R, block, omega_per_item));
  if (np == 0 || nw == 0) return KH_OK;
  return launch_forward<true>(layers, nl, p, np, nullptr, k0, dw, gamma, nw, free_surface,
                              reinterpret_cast<cd*>(R), block, omega_per_item, as_stream(stream));
}

extern "C" int kh_forward_grid(const double* layers, int nl, const double* p, int np, double k0,
                               double dw, double gamma, int nw, int free_surface, double* R,
                               void* stream) {
  int block, wpt;
  default_shape(np, nw, true, &block, &wpt);
  return kh_forward_grid_ex(layers, nl, p, np, k0, dw, gamma, nw, free_surface, R, block, wpt,
                            stream);
}
