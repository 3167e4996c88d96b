// Many earth models of different depths, slowness counts and frequency
// grids in one persistent launch.
//
// The work is a list of items (model, slowness, run of frequency tiles).
// Items are numbered model by model (callers list the models with the
// costliest items first, so the queue drains longest-first); each workgroup
// repeatedly takes the next item number from a device-wide counter (one
// atomic per item), stages that model's layers and interfaces at that
// slowness in LDS once and sweeps the item's frequency tiles with them,
// until the counter passes the item count.  By default an item holds all of
// a slowness's tiles (one item per tile restaged the same tables ntiles
// times: the served mixed-size stream ran 5.6 % slower that way); a caller
// splits the rows of a few deep, wide models into runs of tiles_per_item
// tiles when whole rows would leave the device idle at the end of the launch.  A model whose
// frequencies fill at most about half a tile (descriptor flag half_tile,
// set by the caller) is swept with W/2 frequencies per lane instead of
// leaving half of the lanes idle.  A lane's arithmetic is that of
// forward_kernel with the same (block, frequencies per lane), so every
// model's output is bitwise identical to a kh_forward_grid_ex launch of
// that model alone with that shape.
#include "kh_physics.hpp"
#include "kh_runtime.hpp"
#include "kh_sweep.hpp"

namespace kh {

static_assert(sizeof(kh_model_desc) == 96, "kh_model_desc layout");

// Resets the work counter and, when requested, the per-model counts of
// non-finite outputs.
__global__ void zero_counter_kernel(unsigned long long* c, int* nonfinite, int nmodels) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *c = 0ULL;
  if (nonfinite != nullptr && i < nmodels) nonfinite[i] = 0;
}

__device__ __forceinline__ bool finite_cd(const cd& z) {
  return z.r == z.r && z.i == z.i && fabs(z.r) != INFINITY && fabs(z.i) != INFINITY;
}

__device__ __forceinline__ int find_model(const kh_model_desc* __restrict__ d, int n, long long item) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {  // last model with item_off <= item
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].item_off <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Frequency tiles [t0, t1) of one staged (model, slowness): each lane takes WL
// frequencies strided by the block size, as forward_kernel<WL> does.  Returns
// the lane's count of non-finite outputs.
template <int WL>
__device__ __forceinline__ int sweep_item(const kh_model_desc& md, int t0, int t1,
                                          const LayerPt* s_layer, const Iface* s_if,
                                          const LayerStep* s_step, cd* __restrict__ Rrow) {
  const int nw = md.nw;
  int bad = 0;
  for (int tile = t0; tile < t1; ++tile) {
    const int base = tile * blockDim.x * WL + threadIdx.x;
    cd om[WL];
#pragma unroll
    for (int k = 0; k < WL; ++k) {
      const int w = min(base + k * (int)blockDim.x, nw - 1);
      om[k] = cd{(md.k0 + (double)w) * md.dw, md.gamma};
    }
    cd out[WL];
    sweep_lanes<WL, true>(s_layer, s_if, s_step, md.nl, om, md.free_surface, out);
#pragma unroll
    for (int k = 0; k < WL; ++k) {
      const int w = base + k * blockDim.x;
      if (w < nw) {
        Rrow[w] = out[k];
        bad += finite_cd(out[k]) ? 0 : 1;
      }
    }
  }
  return bad;
}

template <int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) multi_kernel(const kh_model_desc* __restrict__ desc,
                                                    int nmodels, long long total,
                                                    const double* __restrict__ layers_all,
                                                    const double* __restrict__ p_all,
                                                    cd* __restrict__ R_all, int max_nl,
                                                    unsigned long long* __restrict__ counter,
                                                    int* __restrict__ nonfinite) {
  extern __shared__ double smem_m[];
  LayerPt* s_layer = reinterpret_cast<LayerPt*>(smem_m);
  Iface* s_if = reinterpret_cast<Iface*>(s_layer + max_nl);
  LayerStep* s_step = reinterpret_cast<LayerStep*>(s_if + (max_nl - 1));
  __shared__ long long s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (long long)atomicAdd(counter, 1ULL);
    __syncthreads();
    const long long item = s_item;
    __syncthreads();
    if (item >= total) break;  // every wave of every workgroup reaches this
    const int mi = find_model(desc, nmodels, item);
    const kh_model_desc& md = desc[mi];
    const long long local = item - md.item_off;
    const int tpi = md.tiles_per_item > 0 ? md.tiles_per_item : md.ntiles;
    const int nch = (md.ntiles + tpi - 1) / tpi;  // items per slowness
    const int ip = (int)(local / nch);
    const int t0 = (int)(local - (long long)ip * nch) * tpi;
    const int t1 = min(md.ntiles, t0 + tpi);
    const int nl = md.nl, nw = md.nw;
    stage_slowness(layers_all + md.layer_off * NCOL, nl, p_all[md.p_off + ip], s_layer, s_if);
    if (W > 1) {
      const double shift = (double)blockDim.x * md.dw;
      for (int l = threadIdx.x; l < nl; l += blockDim.x) s_step[l] = layer_step(s_layer[l], shift);
      __syncthreads();
    }
    cd* Rrow = R_all + md.r_off + (long long)ip * nw;
    const int bad = (W > 1 && md.half_tile != 0)
                        ? sweep_item<(W > 1 ? W / 2 : 1)>(md, t0, t1, s_layer, s_if, s_step, Rrow)
                        : sweep_item<W>(md, t0, t1, s_layer, s_if, s_step, Rrow);
    if (nonfinite != nullptr && bad != 0) atomicAdd(nonfinite + mi, bad);
    __syncthreads();  // LDS is restaged by the next item
  }
}

static size_t multi_lds_bytes(int max_nl) {
  return (size_t)max_nl * (sizeof(LayerPt) + sizeof(LayerStep)) +
         (size_t)(max_nl - 1) * sizeof(Iface);
}

template <int W>
static int launch_multi(const kh_model_desc* desc, int nmodels, long long total,
                        const double* layers_all, const double* p_all, cd* R_all, int block,
                        int max_nl, unsigned long long* counter, int* nonfinite, int wgs,
                        hipStream_t st) {
  hipLaunchKernelGGL((multi_kernel<W>), dim3(wgs), dim3(block), multi_lds_bytes(max_nl), st, desc,
                     nmodels, total, layers_all, p_all, R_all, max_nl, counter, nonfinite);
  return check_launch("kh_forward_multi");
}

}  // namespace kh

using namespace kh;

extern "C" int kh_multi_tiles(int nw, int block, int omega_per_item, int half_tile) {
  if (nw <= 0 || block <= 0 || omega_per_item <= 0) return 0;
  const int f = (half_tile != 0 && omega_per_item >= 2) ? omega_per_item / 2 : omega_per_item;
  const long long span = (long long)block * f;
  return (int)((nw + span - 1) / span);
}

extern "C" long long kh_multi_items(int np, int ntiles, int tiles_per_item) {
  if (np <= 0 || ntiles <= 0) return 0;
  const int tpi = tiles_per_item > 0 ? tiles_per_item : ntiles;
  return (long long)np * ((ntiles + tpi - 1) / tpi);
}

static int multi_resident(int block, int omega_per_item, int max_nl) {
  int dev = 0, cus = 256, per_cu = 1;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t lds = multi_lds_bytes(max_nl);
  hipError_t e = hipSuccess;
  switch (omega_per_item) {
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, multi_kernel<1>, block, lds); break;
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, multi_kernel<2>, block, lds); break;
    case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, multi_kernel<4>, block, lds); break;
    case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, multi_kernel<8>, block, lds); break;
    default: break;
  }
  if (e != hipSuccess || per_cu < 1) per_cu = 1;
  return cus * per_cu;
}

extern "C" int kh_multi_workgroups(int block, int omega_per_item, int max_nl, int* workgroups) {
  clear_error();
  KH_REQUIRE(workgroups != nullptr, KH_ERR_INVALID_ARGUMENT, "kh_multi_workgroups: NULL pointer");
  KH_REQUIRE(block == 64 || block == 128 || block == 256, KH_ERR_INVALID_ARGUMENT,
             "kh_multi_workgroups: block %d not in {64, 128, 256}", block);
  KH_REQUIRE(omega_per_item == 1 || omega_per_item == 2 || omega_per_item == 4 ||
                 omega_per_item == 8,
             KH_ERR_INVALID_ARGUMENT, "kh_multi_workgroups: omega_per_item %d not in {1, 2, 4, 8}",
             omega_per_item);
  KH_REQUIRE(max_nl >= 2 && max_nl <= KH_MAX_LAYERS, KH_ERR_UNSUPPORTED,
             "kh_multi_workgroups: max_nl %d outside [2, %d]", max_nl, KH_MAX_LAYERS);
  *workgroups = multi_resident(block, omega_per_item, max_nl);
  return KH_OK;
}

extern "C" int kh_forward_multi(const kh_model_desc* desc, int nmodels, long long total_items,
                                int max_nl, const double* layers_all, const double* p_all,
                                double* R_all, int block, int omega_per_item, int workgroups,
                                void* counter, int* nonfinite, void* stream) {
  clear_error();
  KH_REQUIRE(nmodels >= 0 && total_items >= 0, KH_ERR_INVALID_ARGUMENT,
             "kh_forward_multi: negative size");
  if (nmodels == 0 || total_items == 0) return KH_OK;
  KH_REQUIRE(desc && layers_all && p_all && R_all && counter, KH_ERR_INVALID_ARGUMENT,
             "kh_forward_multi: NULL pointer");
  KH_REQUIRE(max_nl >= 2 && max_nl <= KH_MAX_LAYERS, KH_ERR_UNSUPPORTED,
             "kh_forward_multi: max_nl %d outside [2, %d]", max_nl, KH_MAX_LAYERS);
  KH_REQUIRE(block == 64 || block == 128 || block == 256, KH_ERR_INVALID_ARGUMENT,
             "kh_forward_multi: block %d not in {64, 128, 256}", block);
  hipStream_t st = as_stream(stream);
  int wgs = workgroups;
  if (wgs <= 0) wgs = multi_resident(block, omega_per_item, max_nl);  // fill the device
  if ((long long)wgs > total_items) wgs = (int)total_items;
  auto* c = reinterpret_cast<unsigned long long*>(counter);
  const int zero_blocks = nonfinite != nullptr ? (nmodels + 255) / 256 : 1;
  hipLaunchKernelGGL(zero_counter_kernel, dim3(zero_blocks), dim3(256), 0, st, c, nonfinite, nmodels);
  KH_TRY(check_launch("kh_forward_multi (counter)"));
  cd* R = reinterpret_cast<cd*>(R_all);
  switch (omega_per_item) {
    case 1: return launch_multi<1>(desc, nmodels, total_items, layers_all, p_all, R, block, max_nl, c, nonfinite, wgs, st);
    case 2: return launch_multi<2>(desc, nmodels, total_items, layers_all, p_all, R, block, max_nl, c, nonfinite, wgs, st);
    case 4: return launch_multi<4>(desc, nmodels, total_items, layers_all, p_all, R, block, max_nl, c, nonfinite, wgs, st);
    case 8: return launch_multi<8>(desc, nmodels, total_items, layers_all, p_all, R, block, max_nl, c, nonfinite, wgs, st);
    default:
      set_error("kh_forward_multi: omega_per_item %d not in {1, 2, 4, 8}", omega_per_item);
      return KH_ERR_INVALID_ARGUMENT;
  }
}
