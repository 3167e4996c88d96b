// Derivative kernels of the Kennett reflectivity (the differentiable path of
// Kennett_Reflectivity/kennett_torch.py:312-480, its Jacobian/Hessian API
// :611-684, the analytical Frechet derivatives of frechet_analytical.py:583-911
// and the misfit/gradient/Hessian of taup_inversion.py:112-187).
//
//   jvp_kernel            tangent-linear sweep (Z = Dual<cd>): dR = J dm for
//                         one parameter direction per grid slice
//   adjoint_kernel<Z, M>  one wave per (direction, p, 64-frequency tile):
//                         forward keeping the whole tape (or sqrt(N)
//                         checkpoints, when the workspace holds only those),
//                         reverse sweep, per-level cotangents summed over the
//                         wave in registers (a fixed exchange tree) and
//                         written per tile.
//                         Z = cd: vector-Jacobian product (M = 0, given
//                         cotangent) or misfit gradient (M = 1);
//                         Z = Dual<cd>, M = 1: tangent-over-adjoint, i.e. the
//                         misfit gradient and the Hessian-vector product H dm
//   chain_kernel<Z>       per (direction, p): sums the tiles in order, then
//                         one thread per (layer, parameter) chains coefficient
//                         and slowness cotangents to that layer parameter
#include <type_traits>

#include "kh_deriv.hpp"
#include "kh_runtime.hpp"
#include "kh_wave.hpp"

namespace kh {

constexpr int AD_WAVE = 64;
// Waves per adjoint workgroup (frequency tiles sharing one staged table):
// two when a (direction, slowness) has at least two tiles.
constexpr int AD_WAVES_MAX = 2;
constexpr int AD_BLOCK_MAX = AD_WAVE * AD_WAVES_MAX;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = AD_WAVE / 2; o >= 1; o >>= 1) x += __shfl_xor(x, o, AD_WAVE);
  return x;
}
__device__ __forceinline__ cd wave_sum(const cd& z) { return cd{wave_sum(z.r), wave_sum(z.i)}; }
template <class T>
__device__ __forceinline__ Dual<T> wave_sum(const Dual<T>& z) {
  return {wave_sum(z.v), wave_sum(z.d)};
}
// Flat access to the doubles of a scalar (compile-time component index).
template <class Z>
struct Flat;
template <>
struct Flat<cd> {
  static constexpr int N = 2;
  __device__ static double get(const cd& z, int k) { return k == 0 ? z.r : z.i; }
};
template <>
struct Flat<dcd> {
  static constexpr int N = 4;
  __device__ static double get(const dcd& z, int k) {
    return k == 0 ? z.v.r : (k == 1 ? z.v.i : (k == 2 ? z.d.r : z.d.i));
  }
};

// Per-wave sums of the per-lane cotangent records, written per tile: the
// doubles of a record are summed over the 64 lanes in registers by
// wave_reduce_scatter (a fixed exchange tree, the same order on every run),
// after which each lane stores the sum it holds.  No LDS, and no wait on an
// LDS round trip per summand.
template <class Z>
struct WaveSink {
  IfaceZ<Z>* cif;
  LayerCot<Z>* cly;
  bool store;  // false for a wave with no tile

  // the sums of the doubles of v: the first n0 to d0, the rest to d1
  template <int NF>
  __device__ void reduce(const Z (&v)[NF], double* __restrict__ d0, int n0,
                         double* __restrict__ d1) {
    constexpr int N = Flat<Z>::N;
    double x[NF * N];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
#pragma unroll
      for (int k = 0; k < N; ++k) x[f * N + k] = Flat<Z>::get(v[f], k);
    }
    const double s = wave_reduce_scatter<NF * N>(x);
    const int e = wave_rs_index(NF * N, (int)(threadIdx.x % AD_WAVE));
    if (store && e >= 0) *(e < n0 ? d0 + e : d1 + (e - n0)) = s;
  }
  __device__ void iface(int i, const IfaceZ<Z>& f) {
    const Z v[10] = {f.rd00, f.rd01, f.rd11, f.ru00, f.ru01,
                     f.ru11, f.td00, f.td01, f.td10, f.td11};
    reduce(v, reinterpret_cast<double*>(cif + i), 1 << 30, nullptr);
  }
  __device__ void layer(int l, const LayerCot<Z>& c) {
    const Z v[3] = {c.eta, c.neta, c.h};
    reduce(v, reinterpret_cast<double*>(cly + l), 1 << 30, nullptr);
  }
  // interface i and layer i + 1 in one reduction
  __device__ void level(int i, const IfaceZ<Z>& f, const LayerCot<Z>& c) {
    const Z v[13] = {f.rd00, f.rd01, f.rd11, f.ru00, f.ru01, f.ru11, f.td00,
                     f.td01, f.td10, f.td11, c.eta,  c.neta, c.h};
    reduce(v, reinterpret_cast<double*>(cif + i), 10 * Flat<Z>::N,
           reinterpret_cast<double*>(cly + i + 1));
  }
};

// Partial modes: PARTIAL_OFF runs every direction's reverse sweep to the
// bottom; PARTIAL_REST stops it one interface below each direction's deepest
// moved layer except for direction 0 of the call, whose cotangents also give
// the gradient; PARTIAL_ALL stops it for every direction (no gradient asked).
enum { PARTIAL_OFF = 0, PARTIAL_REST = 1, PARTIAL_ALL = 2 };
__device__ __forceinline__ bool stops_early(int partial, int dir) {
  return partial == PARTIAL_ALL || (partial == PARTIAL_REST && dir != 0);
}

// The deepest layer in which direction row `drow` ((nl, 6)) is nonzero,
// -1 for a zero direction, nl - 1 when there is no direction (every
// tangent may be nonzero).  The same for every thread of the workgroup.
__device__ __forceinline__ int deepest_moved_layer(const double* __restrict__ drow, int nl) {
  if (drow == nullptr) return nl - 1;
  for (int l = nl - 1; l >= 0; --l) {
    for (int c = 0; c < NCOL; ++c)
      if (drow[(size_t)l * NCOL + c] != 0.0) return l;
  }
  return -1;
}

// The shallowest layer in which direction row `drow` ((nl, 6)) is nonzero,
// nl for a zero direction, 0 when there is no direction.
__device__ __forceinline__ int shallowest_moved_layer(const double* __restrict__ drow, int nl) {
  if (drow == nullptr) return 0;
  for (int l = 0; l < nl; ++l) {
    for (int c = 0; c < NCOL; ++c)
      if (drow[(size_t)l * NCOL + c] != 0.0) return l;
  }
  return nl;
}

template <class Z>
struct GlobalStore {
  Z* base;
  long stride;
  __device__ Z& operator()(int e) { return base[(long)e * stride]; }
};

// A lane's tape: elements e >= from_e (the states below the deepest layer a
// direction moves) in the tape shared by every direction, which is only
// read; the rest in the (direction, p, tile)'s own storage.  The element's
// address is selected, not its value, so a load issued a level ahead is not
// waited for at the selection.
template <class Z>
struct TapeStore {
  Z* own;
  long own_stride;
  const Z* sh;
  long sh_stride;
  int from_e;
  __device__ Z& operator()(int e) {
    Z* const a = own + (long)e * own_stride;
    Z* const b = const_cast<Z*>(sh) + (long)e * sh_stride;
    return *(e >= from_e ? b : a);
  }
};

template <class Z>
__device__ __forceinline__ void stage_z(const double* __restrict__ layers,
                                        const double* __restrict__ drow_base, int nl, double pv,
                                        LayerZ<Z>* s_layer, IfaceZ<Z>* s_if) {
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    Z x[NCOL];
    row_z(layers + (size_t)l * NCOL, drow_base ? drow_base + (size_t)l * NCOL : nullptr, x);
    s_layer[l] = layer_t<Z>(x, l, pv);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nl - 1; i += blockDim.x)
    s_if[i] = interface_t<Z>(i, pv, s_layer[i], s_layer[i + 1]);
  __syncthreads();
}

// grid.x = ndir * np * ntiles, blockDim = 256: dR[dir, p, w] = J dm_dir.
__global__ void __launch_bounds__(256) jvp_kernel(const double* __restrict__ layers,
                                                  const double* __restrict__ dlayers, int nl,
                                                  const double* __restrict__ p, int np,
                                                  const cd* __restrict__ omega, int nw,
                                                  int free_surface, cd* __restrict__ R,
                                                  cd* __restrict__ dR, int ntiles) {
  extern __shared__ double smem_j[];
  LayerZ<dcd>* s_layer = reinterpret_cast<LayerZ<dcd>*>(smem_j);
  IfaceZ<dcd>* s_if = reinterpret_cast<IfaceZ<dcd>*>(s_layer + nl);
  int b = blockIdx.x;
  const int tile = b % ntiles;
  b /= ntiles;
  const int ip = b % np;
  const int dir = b / np;
  stage_z<dcd>(layers, dlayers + (size_t)dir * nl * NCOL, nl, p[ip], s_layer, s_if);
  const int w = tile * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  const cd om = omega[w];
  const IfaceZ<dcd>& fb = s_if[nl - 2];
  M2<dcd> S{fb.rd00, fb.rd01, fb.rd01, fb.rd11};
  for (int i = nl - 3; i >= 0; --i) {
    dcd ea, eb;
    phases_of(s_layer[i + 1], om, ea, eb);
    S = step_fwd(S, ea, eb, s_if[i]);
  }
  const LayerZ<dcd>& water = s_layer[0];
  dcd ea0 = lift<dcd>(1.0);
  if (!water.halfspace) ea0 = phase_t(water.eta, om, water.h);
  const dcd top = (ea0 * ea0) * S.a00;
  const dcd r = free_surface ? top / (lift<dcd>(1.0) + top) : top;
  const size_t o = (size_t)ip * nw + w;
  dR[(size_t)dir * np * nw + o] = r.d;
  if (R && dir == 0) R[o] = r.v;
}

// The staged tables of every (direction, slowness), computed once for all
// the frequency tiles that read them: one thread per (direction, p, layer)
// writes layer l's (eta, nu, h) and, below it, interface l's coefficients
// (layer l + 1 computed again, so no table needs the whole LayerZ record).
// Block (dir, p) holds nl LayerPh<Z> records and then nl - 1 IfaceZ<Z>
// records, the layout the adjoint kernel copies into its LDS.
template <class Z>
__host__ __device__ __forceinline__ size_t stage_block_bytes(int nl) {
  return (size_t)nl * sizeof(LayerPh<Z>) + (size_t)(nl - 1) * sizeof(IfaceZ<Z>);
}

template <class Z>
__global__ void __launch_bounds__(128) stage_deriv_tables_kernel(
    const double* __restrict__ layers, const double* __restrict__ dlayers, int nl,
    const double* __restrict__ p, int np, int ndir, char* __restrict__ tab) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ndir * np * nl) return;
  const int l = (int)(t % nl);
  const long dp = t / nl;
  const int ip = (int)(dp % np), dir = (int)(dp / np);
  const double* drow = dlayers ? dlayers + (size_t)dir * nl * NCOL : nullptr;
  const double pv = p[ip];
  auto lay = [&](int ll) {
    Z x[NCOL];
    row_z(layers + (size_t)ll * NCOL, drow ? drow + (size_t)ll * NCOL : nullptr, x);
    return layer_t<Z>(x, ll, pv);
  };
  char* blk = tab + (size_t)dp * stage_block_bytes<Z>(nl);
  LayerPh<Z>* ph = reinterpret_cast<LayerPh<Z>*>(blk);
  IfaceZ<Z>* fz = reinterpret_cast<IfaceZ<Z>*>(blk + (size_t)nl * sizeof(LayerPh<Z>));
  const LayerZ<Z> a = lay(l);
  ph[l] = layer_ph(a);
  if (l < nl - 1) fz[l] = interface_t<Z>(l, pv, a, lay(l + 1));
}

// The primal phase factors of every (slowness, layer, frequency), entry
// (ip * nl + l) * nw + w, one thread each, from direction 0's staged layer
// records (the primal parts are the same for every direction): the dual
// sweep reads them instead of evaluating them again for every direction.
template <class Z>
__global__ void __launch_bounds__(256) phase_table_kernel(const char* __restrict__ tab, int nl,
                                                          int np, const cd* __restrict__ omega,
                                                          int nw, PhPair* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)np * nl * nw) return;
  const int w = (int)(t % nw);
  const long pl = t / nw;
  const int l = (int)(pl % nl), ip = (int)(pl / nl);
  const LayerPh<Z>& L =
      reinterpret_cast<const LayerPh<Z>*>(tab + (size_t)ip * stage_block_bytes<Z>(nl))[l];
  PhPair q{cd{1.0, 0.0}, cd{1.0, 0.0}};
  if (!L.halfspace) {
    q.a = phase_t(prim(L.eta), omega[w], prim(L.h));
    q.b = phase_t(prim(L.neta), omega[w], prim(L.h));
  }
  out[t] = q;
}

// The forward states S_j (j = nl-2 .. 0) of the primal sweep of every
// (slowness, frequency), one thread each, as Z with zero tangents: element
// (j * 4 + q) of lane (ip, w) at ((ip * (nl - 1)) * 4 + j * 4 + q) * nw + w.
// The steps are step_fwd_prim's (direction 0's staged interfaces, the
// phase table): the states every direction's sweep would compute below
// the deepest layer it moves.
template <class Z>
__global__ void __launch_bounds__(256) primal_tape_kernel(const char* __restrict__ tab, int nl,
                                                          int np, int nw,
                                                          const PhPair* __restrict__ ph,
                                                          Z* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)np * nw) return;
  const int w = (int)(t % nw), ip = (int)(t / nw);
  const IfaceZ<Z>* F = reinterpret_cast<const IfaceZ<Z>*>(
      tab + (size_t)ip * stage_block_bytes<Z>(nl) + (size_t)nl * sizeof(LayerPh<Z>));
  const PhPair* q = ph + (size_t)ip * nl * nw + w;
  Z* o = out + (size_t)ip * (nl - 1) * 4 * nw + w;
  auto put = [&](int j, const M2<Z>& S) {
    o[(size_t)(j * 4 + 0) * nw] = S.a00;
    o[(size_t)(j * 4 + 1) * nw] = S.a01;
    o[(size_t)(j * 4 + 2) * nw] = S.a10;
    o[(size_t)(j * 4 + 3) * nw] = S.a11;
  };
  const IfaceZ<cd> fb = prim_iface(F[nl - 2]);
  M2<Z> S{lift_c<Z>(fb.rd00), lift_c<Z>(fb.rd01), lift_c<Z>(fb.rd01), lift_c<Z>(fb.rd11)};
  put(nl - 2, S);
  for (int i = nl - 3; i >= 0; --i) {
    const PhPair e = q[(size_t)(i + 1) * nw];
    S = step_fwd_prim(S, e.a, e.b, F[i]);
    put(i, S);
  }
}

// grid.x = ndir * np * ngroups, blockDim = 64 * wpb: wave k of group g
// sweeps frequency tile g * wpb + k of (dir, p), all waves sharing one staged
// table.  ntp = ngroups * wpb tiles have workspace and cotangent slots; a
// wave whose tile is >= ntiles (padding) helps stage the table and returns.
template <class Z, bool MISFIT>
__global__ void __launch_bounds__(AD_BLOCK_MAX) adjoint_kernel(
    const double* __restrict__ layers, const double* __restrict__ dlayers, int nl,
    const double* __restrict__ p, int np, const cd* __restrict__ omega, int nw, int free_surface,
    const cd* __restrict__ Gin, const cd* __restrict__ dGin, int second_only, Z* __restrict__ ws,
    long nlanes,
    IfaceZ<Z>* __restrict__ cot_if, LayerCot<Z>* __restrict__ cot_ly, Z* __restrict__ chi_part,
    cd* __restrict__ R, cd* __restrict__ dR, int ntiles, int ntp, int partial,
    int kstride, const char* __restrict__ tables, const PhPair* __restrict__ ph,
    const Z* __restrict__ pts) {
  extern __shared__ double smem_a[];
  LayerPh<Z>* s_layer = reinterpret_cast<LayerPh<Z>*>(smem_a);
  IfaceZ<Z>* s_if = reinterpret_cast<IfaceZ<Z>*>(s_layer + nl);
  const int wpb = (int)(blockDim.x / AD_WAVE);
  const int wave = (int)(threadIdx.x / AD_WAVE), lane = (int)(threadIdx.x % AD_WAVE);
  const int ngroups = ntp / wpb;
  int b = blockIdx.x;
  const int grp = b % ngroups;
  b /= ngroups;
  const int ip = b % np;
  const int dir = b / np;
  {  // this (direction, slowness)'s staged tables: LayerPh records, then IfaceZ records
    const size_t nb = stage_block_bytes<Z>(nl);
    const double* src = reinterpret_cast<const double*>(tables + ((size_t)dir * np + ip) * nb);
    double* dst = reinterpret_cast<double*>(s_layer);
    for (int e = threadIdx.x; e < (int)(nb / sizeof(double)); e += blockDim.x) dst[e] = src[e];
    __syncthreads();
  }
  const int tile = grp * wpb + wave;
  const bool wave_ok = tile < ntiles;
  if (!wave_ok) return;  // no workgroup barrier follows the staging
  const int w = tile * AD_WAVE + lane;
  const bool valid = wave_ok && w < nw;
  const int wc = valid ? w : nw - 1;
  const cd om = omega[wc];
  const size_t slot = ((size_t)dir * np + ip) * ntp + tile;  // (dir, p, tile)
  WaveSink<Z> sink{cot_if + slot * (nl - 1), cot_ly + slot * nl, wave_ok};
  const cd gin = Gin[(size_t)ip * nw + wc];
  const cd dgin = dGin ? dGin[((size_t)dir * np + ip) * nw + wc] : cd{0.0, 0.0};
  Z chi = lift<Z>(0.0);
  // tangents exist only down to the deepest layer the direction moves; with
  // `partial` the reverse sweep stops one interface below it, except for
  // direction 0, whose cotangents also give the gradient (all layers)
  const int ltan = std::is_same<Z, cd>::value
                       ? nl : deepest_moved_layer(dlayers ? dlayers + (size_t)dir * nl * NCOL
                                                          : nullptr, nl);
  const int lrev = stops_early(partial, dir) ? ltan + 1 : nl;
  // and down from the shallowest moved layer: levels above it take the
  // coefficients and phases without tangent
  const int lsh = std::is_same<Z, cd>::value
                      ? 0 : shallowest_moved_layer(dlayers ? dlayers + (size_t)dir * nl * NCOL
                                                           : nullptr, nl);
  // with the whole tape, the states below the deepest moved layer come from
  // the tape every direction shares (pts, when given)
  PhaseTab pt{ph ? ph + (size_t)ip * nl * nw + wc : nullptr, nw};
  const bool shared = pts != nullptr && kstride == 1 && ltan <= nl - 3;
  if (shared) pt.tape_from = ltan + 1;
  Z* const own = ws + (long)slot * AD_WAVE + lane;
  TapeStore<Z> st{own, nlanes, shared ? pts + (size_t)ip * (nl - 1) * 4 * nw + wc : own,
                  shared ? (long)nw : nlanes, shared ? 4 * (ltan + 1) : 1 << 30};
  // dual Z reads the primal phase factors from the table ph (always given)
  const Z r = lane_adjoint<Z, !std::is_same<Z, cd>::value>(s_layer, s_if, nl, om, free_surface, st, sink, [&](const Z& Rv) {
    if (!valid) return lift<Z>(0.0);
    if (MISFIT) {
      const Z res = Rv - lift_c<Z>(gin);
      chi = re_part(conj(res) * res);
      Z g = res * 2.0;
      if (second_only) g = lift_c<Z>(prim(g));
      return g;
    }
    Z g = lift_c<Z>(gin);
    if constexpr (!std::is_same<Z, cd>::value) g.d = dgin;
    return g;
  }, ltan, lrev, kstride, lsh, pt);
  if (wave_ok && lane == 0) {
    const Z zero = lift<Z>(0.0);
    sink.cly[nl - 1] = LayerCot<Z>{zero, zero, zero};
  }
  if (MISFIT) {
    const Z c = wave_sum(chi);
    if (wave_ok && lane == 0) chi_part[slot] = c;
  }
  if (valid) {
    const size_t o = (size_t)ip * nw + w;
    if (R && dir == 0) R[o] = prim(r);
    if constexpr (!std::is_same<Z, cd>::value) {
      if (dR) dR[(size_t)dir * np * nw + o] = r.d;
    }
  }
}

__device__ __forceinline__ double tangent_re(const cd&) { return 0.0; }
__device__ __forceinline__ double tangent_re(const dcd& z) { return z.d.r; }

// One workgroup per (dir, p).  grad (np, nl, 6) from dir 0; hvp (ndir, np,
// nl, 6) for dual Z; chi (ndir, np, 2) [value, tangent] when chi_part is
// given.  The workgroup first adds the per-tile cotangents of every interface
// and layer in tile order (one double per thread at a time) into LDS, then
// each thread chains one (layer, parameter column) pair: the chain rule
// through the interface formulas runs on dual-of-dual values and takes a
// wave's whole register file, so spreading the columns over threads is what
// fills the device.
constexpr int CH_BLOCK = 128;

// g[0] + g[stride] + ... + g[(ntiles - 1) stride], added in that order from
// 0.0, with eight loads in flight at a time.
__device__ __forceinline__ double tile_sum(const double* g, size_t stride, int ntiles) {
  double acc = 0.0;
  for (int t0 = 0; t0 < ntiles; t0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = t0 + u < ntiles ? g[(size_t)(t0 + u) * stride] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (t0 + u < ntiles) acc += v[u];
  }
  return acc;
}

template <class Z>
__global__ void __launch_bounds__(CH_BLOCK) chain_kernel(const double* __restrict__ layers, const double* __restrict__ dlayers,
                             int nl, const double* __restrict__ p, int np, int ndir,
                             const IfaceZ<Z>* __restrict__ cot_if,
                             const LayerCot<Z>* __restrict__ cot_ly,
                             const Z* __restrict__ chi_part, int ntiles, int ntp,
                             double* __restrict__ grad, double* __restrict__ hvp,
                             double* __restrict__ chi, int partial) {
  extern __shared__ double smem_c[];
  IfaceZ<Z>* s_if = reinterpret_cast<IfaceZ<Z>*>(smem_c);
  LayerCot<Z>* s_ly = reinterpret_cast<LayerCot<Z>*>(s_if + (nl - 1));
  const int ip = (int)(blockIdx.x % (unsigned)np);
  const int dir = (int)(blockIdx.x / (unsigned)np);
  const size_t base = ((size_t)dir * np + ip) * ntp;  // tiles >= ntiles are padding
  constexpr int NI = (int)(sizeof(IfaceZ<Z>) / sizeof(double));
  constexpr int NLY = (int)(sizeof(LayerCot<Z>) / sizeof(double));
  const int nif = (nl - 1) * NI;
  const int nall = nif + nl * NLY;
  const double* gif = reinterpret_cast<const double*>(cot_if + base * (nl - 1));
  const double* gly = reinterpret_cast<const double*>(cot_ly + base * nl);
  double* sif = reinterpret_cast<double*>(s_if);
  double* sly = reinterpret_cast<double*>(s_ly);
  for (int e = threadIdx.x; e < nall; e += blockDim.x) {
    if (e < nif) sif[e] = tile_sum(gif + e, (size_t)nif, ntiles);
    else sly[e - nif] = tile_sum(gly + (e - nif), (size_t)(nall - nif), ntiles);
  }
  __syncthreads();
  const Z zero = lift<Z>(0.0);
  const double* dl = dlayers ? dlayers + (size_t)dir * nl * NCOL : nullptr;
  // with `partial`, the reverse sweep stopped one interface below the
  // deepest moved layer (direction 0 excepted): only the layers above it
  // have whole cotangents
  const int lmax = stops_early(partial, dir) ? deepest_moved_layer(dl, nl) + 1 : nl - 1;
  for (int k = threadIdx.x; k < nl * NCOL; k += blockDim.x) {
    const int l = k / NCOL, col = k - (k / NCOL) * NCOL;
    const size_t g = ((size_t)ip * nl + l) * NCOL + col;
    if (l > lmax) {
      if (hvp) hvp[(size_t)dir * np * nl * NCOL + g] = 0.0;
      continue;
    }
    // chain_column ignores the interface above layer 0 and below layer nl-1:
    // any table entry stands in (a placeholder value would live in scratch)
    const Z v = chain_column<Z>(layers, dl, nl, l, p[ip], s_if[l >= 1 ? l - 1 : 0],
                                s_if[l <= nl - 2 ? l : nl - 2], s_ly[l], col);
    if (grad && dir == 0) grad[g] = prim(v).r;
    if (hvp) hvp[(size_t)dir * np * nl * NCOL + g] = tangent_re(v);
  }
  if (chi && chi_part && threadIdx.x == 0) {
    Z s = zero;
    for (int t = 0; t < ntiles; ++t) s = s + chi_part[base + t];
    chi[((size_t)dir * np + ip) * 2] = prim(s).r;
    chi[((size_t)dir * np + ip) * 2 + 1] = tangent_re(s);
  }
}

// ---- the chain of a partial Hessian, split by what a direction reaches ----
//
// The second-order part of a pair (layer l, column c) of direction d comes
// from the curvature of the coefficient formulas along d at the interfaces
// around layer l, which read layers l - 1, l and l + 1 only.  Where d does
// not move any of them, that curvature is exactly zero and the pair's
// Hessian-vector product is the first-order chain of the cotangents'
// tangent parts (one dual level instead of two).  The few pairs d reaches
// run the full dual-of-dual chain, packed across (direction, slowness) pairs
// so that their waves are full.

// Layers [lo, hi] whose pairs direction d reaches at second order (empty
// when lo > hi), and the deepest layer with a whole cotangent.
struct Reach {
  int lo, hi, lmax;
};
__device__ __forceinline__ Reach reach_of(const double* __restrict__ dl, int nl, int dir,
                                          int partial) {
  const int deep = deepest_moved_layer(dl, nl);
  const int shallow = shallowest_moved_layer(dl, nl);
  Reach r;
  r.lmax = stops_early(partial, dir) ? deep + 1 : nl - 1;
  r.lo = shallow - 1 > 0 ? shallow - 1 : 0;
  const int top = r.lmax < nl - 1 ? r.lmax : nl - 1;
  r.hi = deep + 1 < top ? deep + 1 : top;
  return r;
}

__device__ __forceinline__ IfaceZ<cd> tangent_iface(const IfaceZ<dcd>& f) {
  return IfaceZ<cd>{f.rd00.d, f.rd01.d, f.rd11.d, f.ru00.d, f.ru01.d,
                    f.ru11.d, f.td00.d, f.td01.d, f.td10.d, f.td11.d};
}

// Per (dir, p): the tiles' cotangents summed in tile order over tile 0's
// slot (each thread reads every tile of its double before writing it), and
// chi from the tile partials.
template <class Z>
__global__ void __launch_bounds__(CH_BLOCK) cot_sum_kernel(int nl, int np,
                                                           IfaceZ<Z>* __restrict__ cot_if,
                                                           LayerCot<Z>* __restrict__ cot_ly,
                                                           const Z* __restrict__ chi_part,
                                                           int ntiles, int ntp,
                                                           double* __restrict__ chi) {
  const int ip = (int)(blockIdx.x % (unsigned)np);
  const int dir = (int)(blockIdx.x / (unsigned)np);
  const size_t base = ((size_t)dir * np + ip) * ntp;
  constexpr int NI = (int)(sizeof(IfaceZ<Z>) / sizeof(double));
  constexpr int NLY = (int)(sizeof(LayerCot<Z>) / sizeof(double));
  const int nif = (nl - 1) * NI;
  const int nall = nif + nl * NLY;
  double* gif = reinterpret_cast<double*>(cot_if + base * (nl - 1));
  double* gly = reinterpret_cast<double*>(cot_ly + base * nl);
  for (int e = threadIdx.x; e < nall; e += blockDim.x) {
    double* const g = e < nif ? gif + e : gly + (e - nif);
    const size_t stride = e < nif ? (size_t)nif : (size_t)(nall - nif);
    *g = tile_sum(g, stride, ntiles);
  }
  if (chi && chi_part && threadIdx.x == 0) {
    Z sz = lift<Z>(0.0);
    for (int t = 0; t < ntiles; ++t) sz = sz + chi_part[base + t];
    chi[((size_t)dir * np + ip) * 2] = prim(sz).r;
    chi[((size_t)dir * np + ip) * 2 + 1] = tangent_re(sz);
  }
}

// The first-order factors of the far pairs, one thread per (p, layer,
// column other than h): J[(ip * nl + l) * NJCOL + ci].
constexpr int NJCOL = NCOL - 1;

__global__ void __launch_bounds__(CH_BLOCK) local_jac_kernel(const double* __restrict__ layers,
                                                             int nl, const double* __restrict__ p,
                                                             int np, LocalJac* __restrict__ J) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)np * nl * NJCOL) return;
  const int ci = (int)(t % NJCOL);
  const long r = t / NJCOL;
  const int l = (int)(r % nl), ip = (int)(r / nl);
  J[t] = local_jac(layers, nl, l, p[ip], ci < COL_H ? ci : ci + 1);
}

// Per (dir, p): the pairs the direction does not reach (first-order chain of
// the tangent parts; for direction 0 also the gradient, from the primal
// parts) and zeros beyond lmax.  Reads the tile sums at tile 0's slot and
// the factors J of local_jac_kernel, shared by all directions.
__global__ void __launch_bounds__(CH_BLOCK) chain_far_kernel(
    const double* __restrict__ dlayers, int nl, int np, const IfaceZ<dcd>* __restrict__ cot_if,
    const LayerCot<dcd>* __restrict__ cot_ly, const LocalJac* __restrict__ J, int ntp,
    double* __restrict__ grad, double* __restrict__ hvp, int partial) {
  extern __shared__ double smem_f[];
  IfaceZ<dcd>* s_if = reinterpret_cast<IfaceZ<dcd>*>(smem_f);
  LayerCot<dcd>* s_ly = reinterpret_cast<LayerCot<dcd>*>(s_if + (nl - 1));
  const int ip = (int)(blockIdx.x % (unsigned)np);
  const int dir = (int)(blockIdx.x / (unsigned)np);
  const size_t base = ((size_t)dir * np + ip) * ntp;
  constexpr int NI = (int)(sizeof(IfaceZ<dcd>) / sizeof(double));
  constexpr int NLY = (int)(sizeof(LayerCot<dcd>) / sizeof(double));
  const int nif = (nl - 1) * NI;
  const int nall = nif + nl * NLY;
  const double* gif = reinterpret_cast<const double*>(cot_if + base * (nl - 1));
  const double* gly = reinterpret_cast<const double*>(cot_ly + base * nl);
  double* sif = reinterpret_cast<double*>(s_if);
  double* sly = reinterpret_cast<double*>(s_ly);
  for (int e = threadIdx.x; e < nall; e += blockDim.x) {
    if (e < nif) sif[e] = gif[e];
    else sly[e - nif] = gly[e - nif];
  }
  __syncthreads();
  const double* dl = dlayers + (size_t)dir * nl * NCOL;
  const Reach rc = reach_of(dl, nl, dir, partial);
  const cd zc{0.0, 0.0};
  const IfaceZ<cd> none{zc, zc, zc, zc, zc, zc, zc, zc, zc, zc};
  const LocalJac* Jp = J + (size_t)ip * nl * NJCOL;
  for (int k = threadIdx.x; k < nl * NCOL; k += blockDim.x) {
    const int l = k / NCOL, col = k - (k / NCOL) * NCOL;
    const size_t g = ((size_t)ip * nl + l) * NCOL + col;
    double* hv = hvp + (size_t)dir * np * nl * NCOL + g;
    if (l > rc.lmax) {
      *hv = 0.0;
      continue;
    }
    if (l >= rc.lo && l <= rc.hi) continue;  // chain_near_kernel's
    const LayerCot<dcd>& lc = s_ly[l];
    const LocalJac jc = col == COL_H ? LocalJac{} : Jp[l * NJCOL + (col < COL_H ? col : col - 1)];
    *hv = chain_from_jac(jc, l, nl, l >= 1 ? tangent_iface(s_if[l - 1]) : none,
                         l <= nl - 2 ? tangent_iface(s_if[l]) : none,
                         LayerCot<cd>{lc.eta.d, lc.neta.d, lc.h.d}, col).r;
    if (grad && dir == 0)
      grad[g] = chain_from_jac(jc, l, nl, l >= 1 ? prim_iface(s_if[l - 1]) : none,
                               l <= nl - 2 ? prim_iface(s_if[l]) : none,
                               LayerCot<cd>{lc.eta.v, lc.neta.v, lc.h.v}, col).r;
  }
}

// The pairs each direction reaches, NEAR_THREADS threads per (dir, p)
// (three layers' columns: a basis direction's reach), packed into full
// waves; a wider reach is covered by the same threads in turn.
constexpr int NEAR_THREADS = 3 * NCOL;

__global__ void __launch_bounds__(CH_BLOCK) chain_near_kernel(
    const double* __restrict__ layers, const double* __restrict__ dlayers, int nl,
    const double* __restrict__ p, int np, int ndir, const IfaceZ<dcd>* __restrict__ cot_if,
    const LayerCot<dcd>* __restrict__ cot_ly, int ntp, double* __restrict__ grad,
    double* __restrict__ hvp, int partial) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long pr = t / NEAR_THREADS;
  if (pr >= (long)ndir * np) return;
  const int j0 = (int)(t - pr * NEAR_THREADS);
  const int ip = (int)(pr % np), dir = (int)(pr / np);
  const size_t base = ((size_t)dir * np + ip) * ntp;
  const IfaceZ<dcd>* cif = cot_if + base * (nl - 1);
  const LayerCot<dcd>* cly = cot_ly + base * nl;
  const double* dl = dlayers + (size_t)dir * nl * NCOL;
  const Reach rc = reach_of(dl, nl, dir, partial);
  for (int j = j0; rc.lo + j / NCOL <= rc.hi; j += NEAR_THREADS) {
    const int l = rc.lo + j / NCOL, col = j - (j / NCOL) * NCOL;
    const size_t g = ((size_t)ip * nl + l) * NCOL + col;
    const dcd v = chain_column<dcd>(layers, dl, nl, l, p[ip], cif[l >= 1 ? l - 1 : 0],
                                    cif[l <= nl - 2 ? l : nl - 2], cly[l], col);
    if (grad && dir == 0) grad[g] = prim(v).r;
    hvp[(size_t)dir * np * nl * NCOL + g] = tangent_re(v);
  }
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct DerivLayout {
  size_t ws, cif, cly, chi, jac, tab, ph, pts, total;
  long nlanes;
  int ntiles;  // frequency tiles of AD_WAVE lanes
  int wpb;     // waves (tiles) per workgroup
  int ntp;     // tiles with storage: ntiles rounded up to a multiple of wpb
  int kstride;  // levels between checkpoints: 1 (whole tape) or ckpt_sqrt(nl)
};

// K = 1: the whole forward tape; K = ckpt_sqrt(nl): sqrt(N) checkpoints.
template <class Z>
static DerivLayout deriv_layout(int nl, int np, int nw, int ndir, int K) {
  DerivLayout d;
  d.kstride = K;
  d.ntiles = (nw + AD_WAVE - 1) / AD_WAVE;
  d.wpb = d.ntiles >= 2 ? AD_WAVES_MAX : 1;
  d.ntp = (d.ntiles + d.wpb - 1) / d.wpb * d.wpb;
  const long nblocks = (long)ndir * np * d.ntp;  // (dir, p, tile) slots
  d.nlanes = nblocks * AD_WAVE;
  d.ws = 0;
  d.cif = align256(d.ws + (size_t)d.nlanes * lane_slots_k(nl, K) * sizeof(Z));
  d.cly = align256(d.cif + (size_t)nblocks * (nl - 1) * sizeof(IfaceZ<Z>));
  d.chi = align256(d.cly + (size_t)nblocks * nl * sizeof(LayerCot<Z>));
  d.jac = align256(d.chi + (size_t)nblocks * sizeof(Z));
  // the far-chain factors of the partial Hessian (dual calls only)
  const size_t jac = std::is_same<Z, dcd>::value ? (size_t)np * nl * NJCOL * sizeof(LocalJac) : 0;
  d.tab = align256(d.jac + jac);
  d.ph = align256(d.tab + (size_t)ndir * np * stage_block_bytes<Z>(nl));
  // the primal phase table of the dual sweep
  const size_t phb = std::is_same<Z, dcd>::value ? (size_t)np * nl * nw * sizeof(PhPair) : 0;
  // with the whole tape, the primal states every direction shares
  d.pts = align256(d.ph + phb);
  const size_t ptsb =
      (std::is_same<Z, dcd>::value && K == 1) ? (size_t)np * (nl - 1) * 4 * nw * sizeof(Z) : 0;
  d.total = align256(d.pts + ptsb);
  return d;
}

// The layout a call's workspace holds: the whole tape when it fits, else
// sqrt(N) checkpoints (ok = false when neither fits).
template <class Z>
static DerivLayout layout_for(int nl, int np, int nw, int ndir, size_t ws_bytes, bool& ok) {
  DerivLayout d = deriv_layout<Z>(nl, np, nw, ndir, 1);
  ok = ws_bytes >= d.total;
  if (!ok) {
    d = deriv_layout<Z>(nl, np, nw, ndir, ckpt_sqrt(nl));
    ok = ws_bytes >= d.total;
  }
  return d;
}

static int check_deriv_args(const char* what, const double* layers, int nl, int np, int nw,
                            int ndir) {
  KH_TRY(check_model(what, layers, nl));
  KH_REQUIRE(nl <= KH_MAX_LAYERS_DERIV, KH_ERR_UNSUPPORTED,
             "%s: %d layers exceeds KH_MAX_LAYERS_DERIV=%d", what, nl, KH_MAX_LAYERS_DERIV);
  KH_REQUIRE(np >= 0 && nw >= 0 && ndir >= 1, KH_ERR_INVALID_ARGUMENT, "%s: bad sizes", what);
  return KH_OK;
}

template <class Z, bool MISFIT>
static int run_adjoint(const char* what, const double* layers, const double* dlayers, int ndir,
                       int nl, const double* p, int np, const double* omega, int nw, int fs,
                       const double* Gin, const double* dGin, int second_only, double* chi,
                       double* grad,
                       double* hvp, double* R, double* dR, void* workspace, size_t ws_bytes,
                       hipStream_t st, int partial = 0) {
  // direction 0 keeps its whole sweep only when the call asks for the gradient
  if (partial) partial = grad != nullptr ? PARTIAL_REST : PARTIAL_ALL;
  bool fits = false;
  const DerivLayout d = layout_for<Z>(nl, np, nw, ndir, ws_bytes, fits);
  KH_REQUIRE(workspace && fits, KH_ERR_INVALID_ARGUMENT,
             "%s: workspace of %zu bytes (sqrt(N) checkpoints) or %zu (whole tape) needed, "
             "%zu given", what, d.total, deriv_layout<Z>(nl, np, nw, ndir, 1).total, ws_bytes);
  char* wsb = reinterpret_cast<char*>(workspace);
  const long nblocks = (long)ndir * np * (d.ntp / d.wpb);
  KH_REQUIRE(nblocks <= 0x7fffffffL, KH_ERR_UNSUPPORTED, "%s: grid too large", what);
  // LDS: the (direction, slowness)'s staged tables
  const size_t lds = (size_t)nl * sizeof(LayerPh<Z>) + (size_t)(nl - 1) * sizeof(IfaceZ<Z>);
  auto* cif = reinterpret_cast<IfaceZ<Z>*>(wsb + d.cif);
  auto* cly = reinterpret_cast<LayerCot<Z>*>(wsb + d.cly);
  auto* chip = reinterpret_cast<Z*>(wsb + d.chi);
  const long nstage = (long)ndir * np * nl;
  hipLaunchKernelGGL((stage_deriv_tables_kernel<Z>), dim3((unsigned)((nstage + 127) / 128)),
                     dim3(128), 0, st, layers, dlayers, nl, p, np, ndir, wsb + d.tab);
  KH_TRY(check_launch(what));
  PhPair* ph = nullptr;
  Z* pts = nullptr;
  if constexpr (std::is_same<Z, dcd>::value) {
    ph = reinterpret_cast<PhPair*>(wsb + d.ph);
    const long nph = (long)np * nl * nw;
    hipLaunchKernelGGL((phase_table_kernel<Z>), dim3((unsigned)((nph + 255) / 256)), dim3(256), 0,
                       st, wsb + d.tab, nl, np, reinterpret_cast<const cd*>(omega), nw, ph);
    KH_TRY(check_launch(what));
    if (d.kstride == 1) {
      pts = reinterpret_cast<Z*>(wsb + d.pts);
      const long npw = (long)np * nw;
      hipLaunchKernelGGL((primal_tape_kernel<Z>), dim3((unsigned)((npw + 255) / 256)), dim3(256), 0,
                         st, wsb + d.tab, nl, np, nw, ph, pts);
      KH_TRY(check_launch(what));
    }
  }
  hipLaunchKernelGGL((adjoint_kernel<Z, MISFIT>), dim3((unsigned)nblocks), dim3(AD_WAVE * d.wpb), lds, st,
                     layers, dlayers, nl, p, np, reinterpret_cast<const cd*>(omega), nw, fs,
                     reinterpret_cast<const cd*>(Gin), reinterpret_cast<const cd*>(dGin), second_only,
                     reinterpret_cast<Z*>(wsb + d.ws), d.nlanes, cif, cly, chip,
                     reinterpret_cast<cd*>(R), reinterpret_cast<cd*>(dR), d.ntiles, d.ntp,
                     partial, d.kstride, wsb + d.tab, ph, pts);
  KH_TRY(check_launch(what));
  const size_t lds_c = (size_t)(nl - 1) * sizeof(IfaceZ<Z>) + (size_t)nl * sizeof(LayerCot<Z>);
  if constexpr (std::is_same<Z, dcd>::value) {
    if (partial && dlayers && hvp) {
      hipLaunchKernelGGL((cot_sum_kernel<Z>), dim3((unsigned)((long)ndir * np)), dim3(CH_BLOCK),
                         0, st, nl, np, cif, cly, MISFIT ? chip : nullptr, d.ntiles, d.ntp, chi);
      KH_TRY(check_launch(what));
      auto* jac = reinterpret_cast<LocalJac*>(wsb + d.jac);
      const long njac = (long)np * nl * NJCOL;
      hipLaunchKernelGGL(local_jac_kernel, dim3((unsigned)((njac + CH_BLOCK - 1) / CH_BLOCK)),
                         dim3(CH_BLOCK), 0, st, layers, nl, p, np, jac);
      KH_TRY(check_launch(what));
      hipLaunchKernelGGL(chain_far_kernel, dim3((unsigned)((long)ndir * np)), dim3(CH_BLOCK),
                         lds_c, st, dlayers, nl, np, cif, cly, jac, d.ntp, grad, hvp, partial);
      KH_TRY(check_launch(what));
      const long near_threads = (long)ndir * np * NEAR_THREADS;
      hipLaunchKernelGGL(chain_near_kernel, dim3((unsigned)((near_threads + CH_BLOCK - 1) / CH_BLOCK)),
                         dim3(CH_BLOCK), 0, st, layers, dlayers, nl, p, np, ndir, cif, cly, d.ntp,
                         grad, hvp, partial);
      return check_launch(what);
    }
  }
  hipLaunchKernelGGL((chain_kernel<Z>), dim3((unsigned)((long)ndir * np)), dim3(CH_BLOCK), lds_c, st,
                     layers, dlayers, nl, p, np, ndir, cif, cly, MISFIT ? chip : nullptr,
                     d.ntiles, d.ntp, grad, hvp, chi, partial);
  return check_launch(what);
}

}  // namespace kh

using namespace kh;

extern "C" size_t kh_deriv_workspace_bytes(int nl, int np, int nw, int ndir, int dual) {
  if (nl < 2 || np <= 0 || nw <= 0 || ndir <= 0) return 0;
  return dual ? deriv_layout<dcd>(nl, np, nw, ndir, 1).total
              : deriv_layout<cd>(nl, np, nw, ndir, 1).total;
}

extern "C" size_t kh_deriv_workspace_bytes_min(int nl, int np, int nw, int ndir, int dual) {
  if (nl < 2 || np <= 0 || nw <= 0 || ndir <= 0) return 0;
  const int K = ckpt_sqrt(nl);
  const size_t sq = dual ? deriv_layout<dcd>(nl, np, nw, ndir, K).total
                         : deriv_layout<cd>(nl, np, nw, ndir, K).total;
  const size_t tape = kh_deriv_workspace_bytes(nl, np, nw, ndir, dual);
  return sq < tape ? sq : tape;  // very shallow models: the tape is the smaller
}

extern "C" int kh_jvp(const double* layers, const double* dlayers, int ndir, int nl,
                      const double* p, int np, const double* omega, int nw, int free_surface,
                      double* R, double* dR, void* stream) {
  clear_error();
  KH_TRY(check_deriv_args("kh_jvp", layers, nl, np, nw, ndir));
  if (np == 0 || nw == 0) return KH_OK;
  KH_REQUIRE(dlayers && p && omega && dR, KH_ERR_INVALID_ARGUMENT, "kh_jvp: NULL pointer");
  const int blk = 256;
  const int ntiles = (nw + blk - 1) / blk;
  const long nblocks = (long)ndir * np * ntiles;
  KH_REQUIRE(nblocks <= 0x7fffffffL, KH_ERR_UNSUPPORTED, "kh_jvp: grid too large");
  const size_t lds = (size_t)nl * sizeof(LayerZ<dcd>) + (size_t)(nl - 1) * sizeof(IfaceZ<dcd>);
  hipLaunchKernelGGL(jvp_kernel, dim3((unsigned)nblocks), dim3(blk), lds, as_stream(stream),
                     layers, dlayers, nl, p, np, reinterpret_cast<const cd*>(omega), nw,
                     free_surface, reinterpret_cast<cd*>(R), reinterpret_cast<cd*>(dR), ntiles);
  return check_launch("kh_jvp");
}

extern "C" int kh_vjp(const double* layers, int nl, const double* p, int np, const double* omega,
                      int nw, int free_surface, const double* G, double* grad, double* R,
                      void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  KH_TRY(check_deriv_args("kh_vjp", layers, nl, np, nw, 1));
  if (np == 0) return KH_OK;
  KH_REQUIRE(p && omega && G && grad, KH_ERR_INVALID_ARGUMENT, "kh_vjp: NULL pointer");
  KH_REQUIRE(nw > 0, KH_ERR_INVALID_ARGUMENT, "kh_vjp: nw must be > 0");
  return run_adjoint<cd, false>("kh_vjp", layers, nullptr, 1, nl, p, np, omega, nw, free_surface,
                                G, nullptr, 0, nullptr, grad, nullptr, R, nullptr, workspace,
                                workspace_bytes, as_stream(stream));
}

extern "C" int kh_misfit_grad(const double* layers, int nl, const double* p, int np,
                              const double* omega, int nw, int free_surface, const double* R_obs,
                              double* chi, double* grad, double* R, void* workspace,
                              size_t workspace_bytes, void* stream) {
  clear_error();
  KH_TRY(check_deriv_args("kh_misfit_grad", layers, nl, np, nw, 1));
  if (np == 0) return KH_OK;
  KH_REQUIRE(p && omega && R_obs && grad && chi, KH_ERR_INVALID_ARGUMENT,
             "kh_misfit_grad: NULL pointer");
  KH_REQUIRE(nw > 0, KH_ERR_INVALID_ARGUMENT, "kh_misfit_grad: nw must be > 0");
  return run_adjoint<cd, true>("kh_misfit_grad", layers, nullptr, 1, nl, p, np, omega, nw,
                               free_surface, R_obs, nullptr, 0, chi, grad, nullptr, R, nullptr,
                               workspace,
                               workspace_bytes, as_stream(stream));
}

extern "C" int kh_misfit_hvp(const double* layers, const double* dlayers, int ndir, int nl,
                             const double* p, int np, const double* omega, int nw,
                             int free_surface, const double* R_obs, int second_only, double* chi,
                             double* grad, double* hvp, double* dR, void* workspace,
                             size_t workspace_bytes, void* stream) {
  clear_error();
  KH_TRY(check_deriv_args("kh_misfit_hvp", layers, nl, np, nw, ndir));
  if (np == 0) return KH_OK;
  KH_REQUIRE(dlayers && p && omega && R_obs && hvp, KH_ERR_INVALID_ARGUMENT,
             "kh_misfit_hvp: NULL pointer");
  KH_REQUIRE(nw > 0, KH_ERR_INVALID_ARGUMENT, "kh_misfit_hvp: nw must be > 0");
  return run_adjoint<dcd, true>("kh_misfit_hvp", layers, dlayers, ndir, nl, p, np, omega, nw,
                                free_surface, R_obs, nullptr, second_only, chi, grad, hvp, nullptr,
                                dR,
                                workspace, workspace_bytes, as_stream(stream));
}

extern "C" int kh_misfit_hvp_partial(const double* layers, const double* dlayers, int ndir,
                                     int nl, const double* p, int np, const double* omega,
                                     int nw, int free_surface, const double* R_obs,
                                     int second_only, double* chi, double* grad, double* hvp,
                                     double* dR, void* workspace, size_t workspace_bytes,
                                     void* stream) {
  clear_error();
  KH_TRY(check_deriv_args("kh_misfit_hvp_partial", layers, nl, np, nw, ndir));
  if (np == 0) return KH_OK;
  KH_REQUIRE(dlayers && p && omega && R_obs && hvp, KH_ERR_INVALID_ARGUMENT,
             "kh_misfit_hvp_partial: NULL pointer");
  KH_REQUIRE(nw > 0, KH_ERR_INVALID_ARGUMENT, "kh_misfit_hvp_partial: nw must be > 0");
  return run_adjoint<dcd, true>("kh_misfit_hvp_partial", layers, dlayers, ndir, nl, p, np, omega,
                                nw, free_surface, R_obs, nullptr, second_only, chi, grad, hvp,
                                nullptr, dR, workspace, workspace_bytes, as_stream(stream), 1);
}

extern "C" int kh_vjp_hvp(const double* layers, const double* dlayers, int ndir, int nl,
                          const double* p, int np, const double* omega, int nw, int free_surface,
                          const double* G, const double* dG, double* grad, double* dgrad,
                          double* dR, void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  KH_TRY(check_deriv_args("kh_vjp_hvp", layers, nl, np, nw, ndir));
  if (np == 0) return KH_OK;
  KH_REQUIRE(dlayers && p && omega && G && dgrad, KH_ERR_INVALID_ARGUMENT,
             "kh_vjp_hvp: NULL pointer");
  KH_REQUIRE(nw > 0, KH_ERR_INVALID_ARGUMENT, "kh_vjp_hvp: nw must be > 0");
  return run_adjoint<dcd, false>("kh_vjp_hvp", layers, dlayers, ndir, nl, p, np, omega, nw,
                                 free_surface, G, dG, 0, nullptr, grad, dgrad, nullptr, dR,
                                 workspace, workspace_bytes, as_stream(stream));
}
