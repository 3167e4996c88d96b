// Many earth models of different depths, slowness counts and frequency
// grids in one persistent launch.
//
// The work is a list of items (model, slowness, run of frequency tiles).
// Items are numbered model by model (callers list the models with the
// costliest items first, so the queue drains longest-first); each workgroup
// repeatedly takes the next item number from a device-wide counter (one
// atomic per item) and sweeps the item's tiles with that slowness's layer,
// interface and phase-step tables, until the counter passes the item count.
// The tables come one of two ways:
//   - staged tables (a caller workspace is given): stage_tables_kernel first
//     computes every (model, slowness, layer) record of the batch, one thread
//     per record; the sweep (multi_kernel) reads them with scalar loads, as
//     the addresses are the same for every lane of a wave.  The sweep uses
//     no LDS, so occupancy is set by its registers alone (two waves per SIMD
//     at 8 frequencies per lane, whatever the depth), and an item costs no
//     restaging, so rows are split into runs of tiles wherever that balances
//     the drain at the end of the launch.  Each tile re-reads its tables
//     from L2 or memory.
//   - in-LDS tables (no workspace): each item's workgroup stages its
//     slowness's tables in LDS (multi_lds_kernel), one kernel in all.  LDS
//     bounds occupancy for deep models, and every item restages; a batch is
//     one kernel, which waits for free compute units once rather than twice
//     when other work shares the device.
// A model whose frequencies fill at most about half a tile (descriptor flag
// half_tile, set by the caller) is swept with W/2 frequencies per lane
// instead of leaving half of the lanes idle.  A lane's arithmetic is that
// of forward_kernel with the same (block, frequencies per lane), so every
// model's output is bitwise identical to a kh_forward_grid_ex launch of that
// model alone with that shape, on either path.
#include <mutex>

#include "kh_physics.hpp"
#include "kh_runtime.hpp"
#include "kh_sweep.hpp"

namespace kh {

static_assert(sizeof(kh_model_desc) == 96, "kh_model_desc layout");

typedef const KH_CONST kh_model_desc* CDesc;

// One staged record per (model, slowness, layer), 256 bytes (four 64-byte
// lines that the sweep reads together): the interface below the layer
// (unused for the last layer), what the sweep reads of the layer (vertical
// slownesses, thickness, half-space flag) and its phase step.
struct TableRec {
  Iface f;
  cd eta, neta;
  double h;
  int halfspace, pad_;
  LayerStep s;
  double pad2_[2];
};
static_assert(sizeof(TableRec) == 256, "TableRec layout");
typedef const KH_CONST TableRec* CRec;

// The record array seen as the sweep's three tables.
struct RecLayers {
  CRec r;
  __device__ __forceinline__ const KH_CONST TableRec& operator[](int i) const { return r[i]; }
};
struct RecIfaces {
  CRec r;
  __device__ __forceinline__ const KH_CONST Iface& operator[](int i) const { return r[i].f; }
};
struct RecSteps {
  CRec r;
  __device__ __forceinline__ const KH_CONST LayerStep& operator[](int i) const { return r[i].s; }
};

static size_t table_bytes(long long records) { return (size_t)records * sizeof(TableRec); }

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

// Last model whose `field` offset is <= x (descriptors are in queue order,
// both offsets increase along it).
template <class D, class F>
__device__ __forceinline__ int find_model(D d, int n, long long x, F field) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (field(d[mid]) <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ long long wave_uniform_i64(long long x) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)x);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(x >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// One thread per table record r: layer l of slowness ip of the model whose
// records start at table_off <= r.  The arithmetic is stage_slowness's:
// layer_at for layers l and l + 1, interface_at between them, layer_step.
__global__ void __launch_bounds__(256) stage_tables_kernel(
    const kh_model_desc* __restrict__ desc, int nmodels, long long records,
    const double* __restrict__ layers_all, const double* __restrict__ p_all, int block,
    TableRec* __restrict__ T) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= records) return;
  const int m = find_model(desc, nmodels, r, [](const kh_model_desc& d) { return d.table_off; });
  const kh_model_desc& md = desc[m];
  const long long local = r - md.table_off;
  const int ip = (int)(local / md.nl);
  const int l = (int)(local - (long long)ip * md.nl);
  const double pv = p_all[md.p_off + ip];
  const double* lay = layers_all + md.layer_off * NCOL;
  const LayerPt a = layer_at(lay + (size_t)l * NCOL, l, pv);
  TableRec& t = T[r];
  t.eta = a.eta;
  t.neta = a.neta;
  t.h = a.h;
  t.halfspace = a.halfspace;
  t.pad_ = 0;
  t.s = layer_step(a, (double)block * md.dw);
  if (l < md.nl - 1) t.f = interface_at(l, pv, a, layer_at(lay + (size_t)(l + 1) * NCOL, l + 1, pv));
}

// Frequency tiles [t0, t1) of one (model, slowness): each lane takes WL
// frequencies strided by the block size, as forward_kernel<WL> does.  Returns
// the lane's count of non-finite outputs.
template <int WL, bool UNIFORM, class LP, class FP, class SP>
__device__ __forceinline__ int sweep_item(const kh_model_desc& md, int t0, int t1, LP L, FP F,
                                          SP S, cd* __restrict__ Rrow) {
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
    sweep_lanes<WL, true, UNIFORM>(L, F, S, md.nl, om, md.free_surface, out);
#pragma unroll
    for (int k = 0; k < WL; ++k) {
      const int w = base + k * blockDim.x;
      if (w < nw) {
        store_stream(Rrow + w, out[k]);
        bad += finite_cd(out[k]) ? 0 : 1;
      }
    }
  }
  return bad;
}

template <int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) multi_kernel(
    const kh_model_desc* __restrict__ desc_g, int nmodels, long long total,
    const TableRec* __restrict__ T, cd* __restrict__ R_all,
    unsigned long long* __restrict__ counter, int* __restrict__ nonfinite) {
  const CDesc desc = (CDesc)desc_g;
  __shared__ long long s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (long long)atomicAdd(counter, 1ULL);
    __syncthreads();
    const long long item = wave_uniform_i64(s_item);  // same for the whole workgroup
    __syncthreads();
    if (item >= total) break;  // every wave of every workgroup reaches this
    const int mi = find_model(desc, nmodels, item,
                              [](const KH_CONST kh_model_desc& d) { return d.item_off; });
    const kh_model_desc md = desc[mi];
    const long long local = item - md.item_off;
    const int tpi = md.tiles_per_item > 0 ? md.tiles_per_item : md.ntiles;
    const int nch = (md.ntiles + tpi - 1) / tpi;  // items per slowness
    const int ip = (int)(local / nch);
    const int t0 = (int)(local - (long long)ip * nch) * tpi;
    const int t1 = min(md.ntiles, t0 + tpi);
    const CRec rec = (CRec)T + md.table_off + (long long)ip * md.nl;
    const RecLayers L{rec};
    const RecIfaces F{rec};
    const RecSteps S{rec};
    cd* Rrow = R_all + md.r_off + (long long)ip * md.nw;
    const int bad = (W > 1 && md.half_tile != 0)
                        ? sweep_item<(W > 1 ? W / 2 : 1), false>(md, t0, t1, L, F, S, Rrow)
                        : sweep_item<W, false>(md, t0, t1, L, F, S, Rrow);
    if (nonfinite != nullptr && bad != 0) atomicAdd(nonfinite + mi, bad);
  }
}

// The same items with the tables staged per item in LDS (interface values
// moved to scalar registers by the sweep).
template <int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) multi_lds_kernel(
    const kh_model_desc* __restrict__ desc, int nmodels, long long total,
    const double* __restrict__ layers_all, const double* __restrict__ p_all,
    cd* __restrict__ R_all, int max_nl, unsigned long long* __restrict__ counter,
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
    const int mi = find_model(desc, nmodels, item,
                              [](const kh_model_desc& d) { return d.item_off; });
    const kh_model_desc& md = desc[mi];
    const long long local = item - md.item_off;
    const int tpi = md.tiles_per_item > 0 ? md.tiles_per_item : md.ntiles;
    const int nch = (md.ntiles + tpi - 1) / tpi;  // items per slowness
    const int ip = (int)(local / nch);
    const int t0 = (int)(local - (long long)ip * nch) * tpi;
    const int t1 = min(md.ntiles, t0 + tpi);
    const int nl = md.nl;
    stage_slowness(layers_all + md.layer_off * NCOL, nl, p_all[md.p_off + ip], s_layer, s_if);
    if (W > 1) {
      const double shift = (double)blockDim.x * md.dw;
      for (int l = threadIdx.x; l < nl; l += blockDim.x) s_step[l] = layer_step(s_layer[l], shift);
      __syncthreads();
    }
    cd* Rrow = R_all + md.r_off + (long long)ip * md.nw;
    const LayerPt* L = s_layer;
    const Iface* F = s_if;
    const LayerStep* S = s_step;
    const int bad = (W > 1 && md.half_tile != 0)
                        ? sweep_item<(W > 1 ? W / 2 : 1), true>(md, t0, t1, L, F, S, Rrow)
                        : sweep_item<W, true>(md, t0, t1, L, F, S, Rrow);
    if (nonfinite != nullptr && bad != 0) atomicAdd(nonfinite + mi, bad);
    __syncthreads();  // LDS is restaged by the next item
  }
}

static size_t multi_lds_bytes(int max_nl) {
  return (size_t)max_nl * (sizeof(LayerPt) + sizeof(LayerStep)) +
         (size_t)(max_nl - 1) * sizeof(Iface);
}

// Launch arguments shared by both paths.
struct MultiArgs {
  const kh_model_desc* desc;
  int nmodels;
  long long total;
  const double* layers_all;
  const double* p_all;
  cd* R;
  int block, max_nl;
  unsigned long long* counter;
  int* nonfinite;
  int wgs;
  hipStream_t st;
};

template <int W>
static int launch_multi(const MultiArgs& a, const TableRec* t) {
  if (t != nullptr) {
    hipLaunchKernelGGL((multi_kernel<W>), dim3(a.wgs), dim3(a.block), 0, a.st, a.desc, a.nmodels,
                       a.total, t, a.R, a.counter, a.nonfinite);
  } else {
    hipLaunchKernelGGL((multi_lds_kernel<W>), dim3(a.wgs), dim3(a.block),
                       multi_lds_bytes(a.max_nl), a.st, a.desc, a.nmodels, a.total, a.layers_all,
                       a.p_all, a.R, a.max_nl, a.counter, a.nonfinite);
  }
  return check_launch("kh_forward_multi");
}

template <class K>
static int per_cu(K kernel, int block, size_t lds) {
  int n = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, block, lds) != hipSuccess || n < 1)
    n = 1;
  return n;
}

// CUs x resident workgroups of the path's kernel (max_nl sizes the LDS of
// the in-LDS path).
static int multi_resident(int block, int omega_per_item, int max_nl, bool tables) {
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t lds = multi_lds_bytes(max_nl);
  int n = 1;
  switch (omega_per_item) {
    case 1: n = tables ? per_cu(multi_kernel<1>, block, 0) : per_cu(multi_lds_kernel<1>, block, lds); break;
    case 2: n = tables ? per_cu(multi_kernel<2>, block, 0) : per_cu(multi_lds_kernel<2>, block, lds); break;
    case 4: n = tables ? per_cu(multi_kernel<4>, block, 0) : per_cu(multi_lds_kernel<4>, block, lds); break;
    default: n = tables ? per_cu(multi_kernel<8>, block, 0) : per_cu(multi_lds_kernel<8>, block, lds); break;
  }
  return cus * n;
}

// Argument checks and launches shared by kh_forward_multi and
// kh_forward_multi_static (tables = NULL: the in-LDS path).
static int forward_multi_impl(const char* what, const kh_model_desc* desc, int nmodels,
                              long long total_items, int max_nl, const double* layers_all,
                              const double* p_all, double* R_all, int block, int omega_per_item,
                              int workgroups, long long records, void* tables,
                              size_t table_bytes_given, void* counter, int* nonfinite,
                              hipStream_t st) {
  KH_REQUIRE(nmodels >= 0 && total_items >= 0 && records >= 0, KH_ERR_INVALID_ARGUMENT,
             "%s: negative size", what);
  if (nmodels == 0 || total_items == 0) return KH_OK;
  KH_REQUIRE(desc && layers_all && p_all && R_all && counter, KH_ERR_INVALID_ARGUMENT,
             "%s: NULL pointer", what);
  if (tables != nullptr) {
    KH_REQUIRE(records > 0, KH_ERR_INVALID_ARGUMENT, "%s: no table records", what);
    KH_REQUIRE((reinterpret_cast<uintptr_t>(tables) & 255) == 0, KH_ERR_INVALID_ARGUMENT,
               "%s: table workspace not aligned to 256 bytes", what);
    KH_REQUIRE(table_bytes_given >= table_bytes(records), KH_ERR_OUT_OF_MEMORY,
               "%s: table workspace of %zu bytes, %zu needed", what, table_bytes_given,
               table_bytes(records));
  }
  KH_REQUIRE(max_nl >= 2 && max_nl <= KH_MAX_LAYERS, KH_ERR_UNSUPPORTED,
             "%s: max_nl %d outside [2, %d]", what, max_nl, KH_MAX_LAYERS);
  KH_REQUIRE(block == 64 || block == 128 || block == 256, KH_ERR_INVALID_ARGUMENT,
             "%s: block %d not in {64, 128, 256}", what, block);
  KH_REQUIRE(omega_per_item == 1 || omega_per_item == 2 || omega_per_item == 4 ||
                 omega_per_item == 8,
             KH_ERR_INVALID_ARGUMENT, "%s: omega_per_item %d not in {1, 2, 4, 8}", what,
             omega_per_item);
  const bool use_tables = tables != nullptr;
  int wgs = workgroups > 0 ? workgroups : multi_resident(block, omega_per_item, max_nl, use_tables);
  if ((long long)wgs > total_items) wgs = (int)total_items;
  auto* c = reinterpret_cast<unsigned long long*>(counter);
  hipLaunchKernelGGL(zero_counter_kernel, dim3((nmodels + 255) / 256), dim3(256), 0, st, c,
                     nonfinite, nmodels);
  KH_TRY(check_launch(what));
  TableRec* t = use_tables ? static_cast<TableRec*>(tables) : nullptr;
  if (use_tables) {
    hipLaunchKernelGGL(stage_tables_kernel, dim3((unsigned)((records + 255) / 256)), dim3(256), 0,
                       st, desc, nmodels, records, layers_all, p_all, block, t);
    KH_TRY(check_launch(what));
  }
  const MultiArgs a{desc, nmodels, total_items, layers_all, p_all, reinterpret_cast<cd*>(R_all),
                    block, max_nl, c, nonfinite, wgs, st};
  switch (omega_per_item) {
    case 1: return launch_multi<1>(a, t);
    case 2: return launch_multi<2>(a, t);
    case 4: return launch_multi<4>(a, t);
    default: return launch_multi<8>(a, t);
  }
}

// Library-owned scratch of kh_forward_multi_static, one copy per device:
// the work counter in the first 256 bytes, then up to KH_STATIC_MAX_RECORDS
// table records.  Host code never passes this name to a kernel (on the host
// it is a stand-in, not a device address); it uses the device address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[KH_STATIC_SCRATCH_BYTES];

constexpr int KH_STATIC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[KH_STATIC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per
// device (the first call on a device makes the lookup; later calls read the
// cached address).
static int static_scratch(const char* what, unsigned char** out) {
  int dev = -1;
  KH_REQUIRE(hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < KH_STATIC_MAX_DEVICES,
             KH_ERR_LAUNCH, "%s: no current HIP device (or device index >= %d)", what,
             KH_STATIC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    KH_REQUIRE(e == hipSuccess && ptr != nullptr, KH_ERR_LAUNCH,
               "%s: hipGetSymbolAddress(scratch) failed on device %d: %s", what, dev,
               hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return KH_OK;
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

extern "C" size_t kh_multi_table_bytes(long long records) {
  return records > 0 ? table_bytes(records) : 0;
}

extern "C" int kh_multi_workgroups(int block, int omega_per_item, int max_nl, int tables,
                                   int* workgroups) {
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
  *workgroups = multi_resident(block, omega_per_item, max_nl, tables != 0);
  return KH_OK;
}

extern "C" int kh_forward_multi(const kh_model_desc* desc, int nmodels, long long total_items,
                                int max_nl, const double* layers_all, const double* p_all,
                                double* R_all, int block, int omega_per_item, int workgroups,
                                long long records, void* tables, size_t table_bytes_given,
                                void* counter, int* nonfinite, void* stream) {
  clear_error();
  return forward_multi_impl("kh_forward_multi", desc, nmodels, total_items, max_nl, layers_all,
                            p_all, R_all, block, omega_per_item, workgroups, records, tables,
                            table_bytes_given, counter, nonfinite, as_stream(stream));
}

extern "C" int kh_forward_multi_static(const kh_model_desc* desc, int nmodels,
                                       long long total_items, int max_nl,
                                       const double* layers_all, const double* p_all,
                                       double* R_all, int block, int omega_per_item,
                                       int workgroups, long long records, int use_tables,
                                       int* nonfinite, void* stream) {
  const char* what = "kh_forward_multi_static";
  clear_error();
  unsigned char* base = nullptr;
  KH_TRY(static_scratch(what, &base));  // resolved even when there is nothing to launch
  KH_REQUIRE(records >= 0, KH_ERR_INVALID_ARGUMENT, "%s: negative size", what);
  KH_REQUIRE(use_tables == 0 || records <= (long long)KH_STATIC_MAX_RECORDS,
             KH_ERR_INVALID_ARGUMENT,
             "%s: %lld table records exceed the static scratch maximum of %lld "
             "(KH_STATIC_MAX_RECORDS; use_tables = 0 takes any batch)",
             what, records, (long long)KH_STATIC_MAX_RECORDS);
  void* tables = use_tables != 0 ? static_cast<void*>(base + KH_STATIC_COUNTER_BYTES) : nullptr;
  const size_t tbytes = use_tables != 0 ? (size_t)KH_STATIC_SCRATCH_BYTES - KH_STATIC_COUNTER_BYTES
                                        : 0;
  return forward_multi_impl(what, desc, nmodels, total_items, max_nl, layers_all, p_all, R_all,
                            block, omega_per_item, workgroups, records, tables, tbytes,
                            static_cast<void*>(base), nonfinite, as_stream(stream));
}
