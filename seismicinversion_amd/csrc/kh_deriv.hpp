// Per-lane derivative algorithms of the Kennett recursion, shared by the
// device kernels (kh_deriv.hip) and the host build used by the CPU tests.
//
// lane_adjoint<Z>: one (p, omega) lane.
//   1. forward sweep from the deepest interface, saving the state S_j every
//      K levels (j % K == 0) in per-lane storage: K = 1 (the whole tape)
//      where the workspace holds it, else K = ceil(sqrt(nl-1));
//   2. R at the surface and its cotangent G (given, or 2 (R - R_obs) for the
//      misfit |R - R_obs|^2);
//   3. reverse sweep top-down: with the whole tape each level reads its
//      state and recomputes its phase factors; with K > 1, in segments of K
//      interfaces, each segment's states are recomputed from the checkpoint
//      below it (with the phase factors) into a K-entry buffer, then walked
//      back with step_rev.
//   The cotangents of every interface's coefficients and of every layer's
//   (eta, nu, h) are handed to a sink (the kernel reduces them over the
//   wavefront; the host harness sums them).
// With Z = Dual<cd> seeded along a parameter direction dm every quantity
// also carries its derivative along dm: the surface tangent is J dm and the
// cotangents' tangents, chained by chain_layer, give the Hessian-vector
// product (tangent-over-adjoint).
//
// chain_layer<Z>: cotangents of the six parameters [alpha, beta, rho, h,
// Q_alpha, Q_beta] of layer l from the (omega-summed) cotangents of the two
// interfaces around it and of its own (eta, nu, h), using the derivatives of
// the coefficient formulas along each parameter (one dual level above Z).
#pragma once

#include <type_traits>

#include "kh_adjoint.hpp"
#include "kennett_hip.h"

namespace kh {

template <class Z>
struct LayerCot {
  Z eta, neta, h;
};

// Levels between checkpoints: K = 1 keeps the whole forward tape (every
// level's state; the reverse sweep recomputes nothing and reads one state
// per level), K = ckpt_sqrt(nl) = ceil(sqrt(nl - 1)) keeps sqrt(N)
// checkpoints and recomputes each segment of K levels once.
__host__ __device__ __forceinline__ int ckpt_sqrt(int nl) {
  int k = 1;
  while (k * k < nl - 1) ++k;
  return k;
}
__host__ __device__ __forceinline__ int ckpt_count(int nl, int K) { return (nl - 2) / K + 1; }
// Z-slots of per-lane storage at stride K: the checkpoints (4 each), then,
// for K > 1, the K-level buffer of the segment being walked back (6 each).
__host__ __device__ __forceinline__ int lane_slots_k(int nl, int K) {
  return ckpt_count(nl, K) * 4 + (K > 1 ? K * 6 : 0);
}

template <class Z>
__host__ __device__ __forceinline__ Z zero_z() {
  return lift<Z>(0.0);
}

// What the sweep reads of a layer: its vertical slownesses and thickness
// (the interface coefficients are precomputed).  Staging this instead of
// the whole LayerZ keeps deep models within the LDS that four workgroups
// per CU can share.
template <class Z>
struct LayerPh {
  Z eta, neta, h;
  int halfspace;
};

template <class Z>
__host__ __device__ __forceinline__ LayerPh<Z> layer_ph(const LayerZ<Z>& L) {
  return LayerPh<Z>{L.eta, L.neta, L.h, L.halfspace};
}

template <class Z, class LZ>
__host__ __device__ __forceinline__ void phases_of(const LZ& lb, cd om, Z& ea, Z& eb) {
  if (lb.halfspace) {
    ea = lift<Z>(1.0);
    eb = lift<Z>(1.0);
  } else {
    ea = phase_t(lb.eta, om, lb.h);
    eb = phase_t(lb.neta, om, lb.h);
  }
}

// The primal parts of an interface's coefficients, which are the same for
// every lane of the wave (one (direction, slowness) per workgroup), moved
// to scalar registers; tangent parts stay per lane.
__host__ __device__ __forceinline__ cd primal_uniform(const cd& z) { return wave_uniform(z); }
__host__ __device__ __forceinline__ dcd primal_uniform(const dcd& z) {
  return dcd{wave_uniform(z.v), z.d};
}
template <class Z>
__host__ __device__ __forceinline__ IfaceZ<Z> primal_uniform(const IfaceZ<Z>& f) {
  return IfaceZ<Z>{primal_uniform(f.rd00), primal_uniform(f.rd01), primal_uniform(f.rd11),
                   primal_uniform(f.ru00), primal_uniform(f.ru01), primal_uniform(f.ru11),
                   primal_uniform(f.td00), primal_uniform(f.td01), primal_uniform(f.td10),
                   primal_uniform(f.td11)};
}

// Primal parts of an interface's coefficients.
__host__ __device__ __forceinline__ const IfaceZ<cd>& prim_iface(const IfaceZ<cd>& f) { return f; }
template <class T>
__host__ __device__ __forceinline__ IfaceZ<cd> prim_iface(const IfaceZ<Dual<T>>& f) {
  return IfaceZ<cd>{prim(f.rd00), prim(f.rd01), prim(f.rd11), prim(f.ru00), prim(f.ru01),
                    prim(f.ru11), prim(f.td00), prim(f.td01), prim(f.td10), prim(f.td11)};
}

// A forward step whose inputs carry no tangent: every layer it reads lies
// below the deepest layer the direction moves, so the state below, the
// interface and the phases (a, b: the primal phase factors of the layer
// below the interface) all have zero tangents.  The plain step on the
// primal parts, lifted back with zero tangents: bitwise the primal of the
// full step, and the tangents it leaves out are exactly zero.
template <class Z>
__host__ __device__ __forceinline__ M2<Z> step_fwd_prim(const M2<Z>& S, const cd& a, const cd& b,
                                                        const IfaceZ<Z>& f) {
  if constexpr (std::is_same<Z, cd>::value) {
    return step_fwd(S, a, b, f);
  } else {
    const M2<cd> s{prim(S.a00), prim(S.a01), prim(S.a10), prim(S.a11)};
    const M2<cd> r = step_fwd(s, a, b, prim_iface(f));
    return M2<Z>{lift_c<Z>(r.a00), lift_c<Z>(r.a01), lift_c<Z>(r.a10), lift_c<Z>(r.a11)};
  }
}

// Primal phase factors (ea, eb) of one layer at one frequency, and a lane's
// view of a table of them: entry l (layer l) at p[l * stride].  The table
// holds what the sweep would compute from the same primal eta, neta and h
// ((1, 0) for the half-space), so reading it changes no result.
struct PhPair {
  cd a, b;
};
struct PhaseTab {
  const PhPair* p;  // nullptr: the sweep computes the phases itself
  long stride;
  // With the whole tape (K = 1): the first level whose state the tape
  // storage already holds, shared by every direction (a primal state with
  // zero tangent); the forward sweep starts from it.  Default: none.
  int tape_from = 1 << 30;
};

// Storage: st(e) returns a reference to Z-slot e of this lane.
// Sink: sink.iface(i, IfaceZ<Z>), sink.layer(l, LayerCot<Z>) and
// sink.level(i, IfaceZ<Z>, LayerCot<Z>) for interface i and layer i + 1 together.
// G_of(R) -> cotangent of R.  Returns R.
// L: per-layer records with eta, neta, h and halfspace (LayerZ<Z> or
// LayerPh<Z>).
// ltan: the deepest layer whose parameters carry a tangent (dual Z): forward
// steps at interfaces i > ltan read only tangent-free layers and run on the
// primal parts alone (step_fwd_prim; results unchanged).  lsh: the
// shallowest such layer (0 when unknown); steps at interfaces whose two
// layers lie outside [lsh, ltan] take the coefficients as plain cd, and
// every layer outside [lsh, ltan] gives its phases as plain cd (results
// unchanged).  pt: the primal phase factors of this lane's frequency per
// layer, read when TAB (else computed in place).  lrev: the reverse
// sweep stops after interface lrev, so the sink receives interfaces
// 0..min(lrev, nl-2) and layers 0..min(lrev + 1, nl-1) only.
// Checkpoints and buffer in separate storage: st holds the states S_j at
// every K-th level (slots 0 .. ckpt_count(nl, K) * 4), bst the K-level
// buffer of a segment being walked back (K * 6 slots).  K = ckpt_stride(nl)
// balances the two; a smaller K keeps the buffer small enough for on-chip
// storage (the reverse reads it right after the recompute writes it).
template <class Z, bool TAB = false, class LZ, class Store, class BufStore, class Sink, class GFn>
__host__ __device__ inline Z lane_adjoint_k(const LZ* __restrict__ L,
                                            const IfaceZ<Z>* __restrict__ F, int nl, cd om,
                                            int free_surface, Store& st, BufStore& bst, int K,
                                            Sink& sink, GFn&& G_of, int ltan, int lrev,
                                            int lsh = 0, PhaseTab pt = PhaseTab{nullptr, 0}) {
  auto save = [&](int slot, const M2<Z>& S) {
    st(slot * 4 + 0) = S.a00;
    st(slot * 4 + 1) = S.a01;
    st(slot * 4 + 2) = S.a10;
    st(slot * 4 + 3) = S.a11;
  };
  auto load = [&](int slot) {
    return M2<Z>{st(slot * 4 + 0), st(slot * 4 + 1), st(slot * 4 + 2), st(slot * 4 + 3)};
  };
  auto buf = [&](int k, int q) -> Z& { return bst(k * 6 + q); };
  // the reverse sweep ends at interface ilast: checkpoints below the one
  // its last segment starts from are never read
  const int ilast = (lrev < nl - 2) ? lrev : nl - 2;
  const int ck_max = ilast / K + 1;
  // dual Z: interface i and the phases of layer i + 1 carry no tangent when
  // neither layer i nor i + 1 lies in [lsh, ltan], the layers the direction
  // moves; the step then takes them as plain cd (mixed operations, same
  // numbers, a third fewer products with them)
  constexpr bool DUAL = !std::is_same<Z, cd>::value;
  auto flat = [&](int i) { return DUAL && (i + 1 < lsh || i > ltan); };
  // layer l's phases carry a tangent (dual Z only)
  auto moved = [&](int l) { return DUAL && l >= lsh && l <= ltan; };
  constexpr bool tab = TAB;
  auto phases_prim = [&](int l, cd& ea, cd& eb) {
    if (tab) {
      const PhPair q = pt.p[(long)l * pt.stride];
      ea = q.a;
      eb = q.b;
      return;
    }
    ea = cd{1.0, 0.0};
    eb = cd{1.0, 0.0};
    if (!L[l].halfspace) {
      ea = phase_t(prim(L[l].eta), om, prim(L[l].h));
      eb = phase_t(prim(L[l].neta), om, prim(L[l].h));
    }
  };
  // the forward step at interface i, given the phases of layer i + 1: the
  // primal ones (a, b) when it does not move, else the dual ones (ea, eb)
  auto fwd_step = [&](const M2<Z>& T, int i, const cd& a, const cd& b, const Z& ea,
                      const Z& eb) -> M2<Z> {
    if constexpr (DUAL) {
      if (i > ltan) return step_fwd_prim(T, a, b, F[i]);
      if (flat(i)) return step_fwd(T, a, b, primal_uniform(prim_iface(F[i])));
      if (!moved(i + 1)) return step_fwd(T, a, b, F[i]);
      return step_fwd(T, ea, eb, F[i]);
    } else {
      return step_fwd(T, a, b, F[i]);
    }
  };

  // 1. forward with checkpoints (with a table, each layer's phases are
  // loaded one step ahead)
  const IfaceZ<Z>& fb = F[nl - 2];
  M2<Z> S{fb.rd00, fb.rd01, fb.rd01, fb.rd11};
  PhPair ahead{cd{1.0, 0.0}, cd{1.0, 0.0}};
  auto fwd_level = [&](int i) {
    cd a{1.0, 0.0}, b{1.0, 0.0};
    Z ea = lift<Z>(1.0), eb = lift<Z>(1.0);
    const bool mv = moved(i + 1);
    if constexpr (tab) {
      a = ahead.a;
      b = ahead.b;
      ahead = pt.p[(long)(i >= 1 ? i : 0) * pt.stride];  // unconditional: no merge, no early wait
    } else if (!mv) {
      phases_prim(i + 1, a, b);
    }
    if (mv) phases_of(L[i + 1], om, ea, eb);
    S = fwd_step(S, i, a, b, ea, eb);
  };
  if (K == 1 && pt.tape_from <= nl - 2) {
    // the storage holds the states from tape_from down: start there and
    // save every state computed (unconditional stores)
    S = load(pt.tape_from);
    const int i0 = pt.tape_from - 1;
    if (tab && i0 >= 0) ahead = pt.p[(long)(i0 + 1) * pt.stride];
    for (int i = i0; i >= 0; --i) {
      fwd_level(i);
      save(i, S);
    }
  } else {
    if ((nl - 2) % K == 0 && (nl - 2) / K <= ck_max) save((nl - 2) / K, S);
    if (tab && nl >= 3) ahead = pt.p[(long)(nl - 2) * pt.stride];
    for (int i = nl - 3; i >= 0; --i) {
      fwd_level(i);
      if (i % K == 0 && i / K <= ck_max) save(i / K, S);
    }
  }

  // 2. surface value and its cotangent
  const LZ& water = L[0];
  Z ea0 = lift<Z>(1.0);
  if (!water.halfspace) {
    if (moved(0)) {
      ea0 = phase_t(water.eta, om, water.h);
    } else {
      cd a, b;
      phases_prim(0, a, b);
      ea0 = lift_c<Z>(a);
    }
  }
  const Z e2 = ea0 * ea0;
  const Z top = e2 * S.a00;
  const Z one = lift<Z>(1.0);
  const Z R = free_surface ? top / (one + top) : top;
  const Z G = G_of(R);
  const Z opt = one + top;
  const Z tb = free_surface ? conj(recip(opt * opt)) * G : G;
  M2<Z> sb{conj(e2) * tb, zero_z<Z>(), zero_z<Z>(), zero_z<Z>()};
  {
    LayerCot<Z> c0{zero_z<Z>(), zero_z<Z>(), zero_z<Z>()};
    if (!water.halfspace) {
      const Z ea0b = conj(ea0 * 2.0 * S.a00) * tb;
      phase_rev(water.eta, om, water.h, ea0, ea0b, c0.eta, c0.h);
    }
    sink.layer(0, c0);
  }

  // 3. reverse, top-down to interface ilast
  // the bottom interface: RRd = Rd, no step
  auto bottom = [&](int i) {
    IfaceZ<Z> fbar;
    fbar.ru00 = fbar.ru01 = fbar.ru11 = zero_z<Z>();
    fbar.td00 = fbar.td01 = fbar.td10 = fbar.td11 = zero_z<Z>();
    fbar.rd00 = sb.a00;
    fbar.rd01 = sb.a01 + sb.a10;
    fbar.rd11 = sb.a11;
    sink.iface(i, fbar);
  };
  // interface i, from the state below it and the phases of layer i + 1;
  // then() runs once the step's arithmetic is done (the whole-tape sweep
  // issues the next level's state load there, so that its latency overlaps
  // this level's cotangent reduction)
  // (phases of type Z, or plain cd where layer i + 1 does not move;
  // flat_iface: std::true_type where interface i carries no tangent)
  auto level = [&](int i, const M2<Z>& Sin, const auto& ea, const auto& eb, auto flat_iface,
                   auto&& then) {
    using EZ = std::decay_t<decltype(ea)>;
    IfaceZ<Z> fbar;  // step_rev stores the Rd and Td cotangents, adds the Ru ones
    fbar.ru00 = fbar.ru01 = fbar.ru11 = zero_z<Z>();
    M2<Z> rb;
    Z eab, ebb;
    if constexpr (decltype(flat_iface)::value)
      step_rev(Sin, ea, eb, primal_uniform(prim_iface(F[i])), sb, rb, fbar, eab, ebb);
    else
      step_rev(Sin, ea, eb, primal_uniform(F[i]), sb, rb, fbar, eab, ebb);
    then();
    LayerCot<Z> c{zero_z<Z>(), zero_z<Z>(), zero_z<Z>()};
    const LZ& lb = L[i + 1];
    if (!lb.halfspace) {
      Z hb2 = zero_z<Z>();
      if constexpr (std::is_same<EZ, Z>::value) {
        phase_rev(lb.eta, om, lb.h, ea, eab, c.eta, c.h);
        phase_rev(lb.neta, om, lb.h, eb, ebb, c.neta, hb2);
      } else {
        phase_rev(prim(lb.eta), om, prim(lb.h), ea, eab, c.eta, c.h);
        phase_rev(prim(lb.neta), om, prim(lb.h), eb, ebb, c.neta, hb2);
      }
      c.h = c.h + hb2;
    }
    sink.level(i, fbar, c);  // interface i and layer i + 1
    sb = rb;
  };
  if (K == 1) {  // the whole tape: S_{i+1} is checkpoint i + 1
    M2<Z> next;
    PhPair nph{cd{1.0, 0.0}, cd{1.0, 0.0}};  // with a table: layer i + 1's phases
    if (ilast >= 0 && nl - 2 > 0) {
      next = load(1);
      if (tab) nph = pt.p[pt.stride];
    }
    for (int i = 0; i <= ilast; ++i) {
      if (i == nl - 2) {
        bottom(i);
        continue;
      }
      const M2<Z> cur = next;
      const PhPair q = nph;
      auto prefetch = [&] {
        if (i + 1 <= ilast && i + 1 != nl - 2) {
          next = load(i + 2);
          if (tab) nph = pt.p[(long)(i + 2) * pt.stride];
        }
      };
      cd a = q.a, b = q.b;
      if (!tab && !moved(i + 1)) phases_prim(i + 1, a, b);
      if constexpr (DUAL) {
        if (flat(i)) {
          level(i, cur, a, b, std::true_type{}, prefetch);
        } else if (!moved(i + 1)) {
          level(i, cur, a, b, std::false_type{}, prefetch);
        } else {
          Z ea, eb;
          phases_of(L[i + 1], om, ea, eb);
          level(i, cur, ea, eb, std::false_type{}, prefetch);
        }
      } else {
        level(i, cur, a, b, std::false_type{}, prefetch);
      }
    }
    return R;
  }
  // segment by segment of K levels
  for (int i0 = 0; i0 <= ilast; i0 += K) {
    const int i1 = (i0 + K - 1 < nl - 2) ? i0 + K - 1 : nl - 2;
    // recompute S_{i+1} (and the phases of layer i+1) for i = i1 .. i0
    M2<Z> T;
    if (i1 == nl - 2) {
      T = M2<Z>{fb.rd00, fb.rd01, fb.rd01, fb.rd11};  // S_{nl-2}
    } else {
      T = load((i1 + 1) / K);  // (i1 + 1) is a multiple of K
    }
    for (int i = i1; i >= i0; --i) {
      if (i == nl - 2) continue;  // bottom interface: no step (S_{nl-2} = Rd)
      const int k = i - i0;
      buf(k, 0) = T.a00;
      buf(k, 1) = T.a01;
      buf(k, 2) = T.a10;
      buf(k, 3) = T.a11;
      cd a{1.0, 0.0}, b{1.0, 0.0};
      Z ea, eb;
      if (moved(i + 1)) {
        phases_of(L[i + 1], om, ea, eb);
      } else {
        phases_prim(i + 1, a, b);
        ea = lift_c<Z>(a);
        eb = lift_c<Z>(b);
      }
      // the state the step at i0 would give lies above the segment
      if (i != i0) T = fwd_step(T, i, a, b, ea, eb);
      buf(k, 4) = ea;
      buf(k, 5) = eb;
    }
    const int iend = (i1 < ilast) ? i1 : ilast;
    for (int i = i0; i <= iend; ++i) {
      if (i == nl - 2) {
        bottom(i);
        continue;
      }
      const int k = i - i0;
      const M2<Z> Sin{buf(k, 0), buf(k, 1), buf(k, 2), buf(k, 3)};
      if constexpr (DUAL) {
        if (flat(i)) {  // the buffered phases have no tangent here
          level(i, Sin, prim(buf(k, 4)), prim(buf(k, 5)), std::true_type{}, [] {});
          continue;
        }
        if (!moved(i + 1)) {  // nor here (the interface has one)
          level(i, Sin, prim(buf(k, 4)), prim(buf(k, 5)), std::false_type{}, [] {});
          continue;
        }
      }
      level(i, Sin, buf(k, 4), buf(k, 5), std::false_type{}, [] {});
    }
  }
  return R;
}

// One storage for both: the checkpoints, then the buffer (lane_slots_k(nl, K)
// slots).
template <class Z, class Store>
struct BufAfter {
  Store& st;
  int off;
  __host__ __device__ Z& operator()(int e) { return st(off + e); }
};

template <class Z, bool TAB = false, class LZ, class Store, class Sink, class GFn>
__host__ __device__ inline Z lane_adjoint(const LZ* __restrict__ L,
                                          const IfaceZ<Z>* __restrict__ F, int nl, cd om,
                                          int free_surface, Store& st, Sink& sink, GFn&& G_of,
                                          int ltan = 1 << 30, int lrev = 1 << 30, int K = 1,
                                          int lsh = 0, PhaseTab pt = PhaseTab{nullptr, 0}) {
  BufAfter<Z, Store> bst{st, ckpt_count(nl, K) * 4};
  return lane_adjoint_k<Z, TAB>(L, F, nl, om, free_surface, st, bst, K, sink, G_of, ltan, lrev,
                                lsh, pt);
}

// One dual level above a scalar type, with the seeding of its inputs.
template <class Z>
struct Up;
template <>
struct Up<cd> {
  using W = dcd;
  // value x, direction dm component (unused), outer seed e
  __host__ __device__ static W seed(double x, double /*dm*/, double e) { return seed1(x, e); }
  __host__ __device__ static cd outer(const W& w) { return w.d; }
};
template <>
struct Up<dcd> {
  using W = ddcd;
  __host__ __device__ static W seed(double x, double dm, double e) { return seed2(x, dm, e); }
  __host__ __device__ static dcd outer(const W& w) { return w.d; }
};

// Z-valued layer from a parameter row (+ direction row when Z is dual).
__host__ __device__ __forceinline__ void row_z(const double* row, const double*, cd (&x)[NCOL]) {
  for (int c = 0; c < NCOL; ++c) x[c] = cd{row[c], 0.0};
}
__host__ __device__ __forceinline__ void row_z(const double* row, const double* drow,
                                               dcd (&x)[NCOL]) {
  for (int c = 0; c < NCOL; ++c) x[c] = seed1(row[c], drow ? drow[c] : 0.0);
}

template <class Z>
__host__ __device__ __forceinline__ Z re_dot(const Z& a, const Z& b) {
  return re_part(conj(a) * b);
}

template <class Z>
__host__ __device__ __forceinline__ Z iface_dot(const IfaceZ<Z>& dc, const IfaceZ<Z>& cb) {
  return re_dot(dc.rd00, cb.rd00) + re_dot(dc.rd01, cb.rd01) + re_dot(dc.rd11, cb.rd11) +
         re_dot(dc.ru00, cb.ru00) + re_dot(dc.ru01, cb.ru01) + re_dot(dc.ru11, cb.ru11) +
         re_dot(dc.td00, cb.td00) + re_dot(dc.td01, cb.td01) + re_dot(dc.td10, cb.td10) +
         re_dot(dc.td11, cb.td11);
}

template <class W, class Z>
__host__ __device__ __forceinline__ IfaceZ<Z> outer_iface(const IfaceZ<W>& f) {
  IfaceZ<Z> o;
  o.rd00 = f.rd00.d; o.rd01 = f.rd01.d; o.rd11 = f.rd11.d;
  o.ru00 = f.ru00.d; o.ru01 = f.ru01.d; o.ru11 = f.ru11.d;
  o.td00 = f.td00.d; o.td01 = f.td01.d; o.td10 = f.td10.d; o.td11 = f.td11.d;
  return o;
}

// Cotangent of parameter column col (a COL_* index) of layer l, given the
// summed cotangents: cb_up = interface l-1 (ignored for l = 0), cb_dn =
// interface l (ignored for l = nl-1), lc = layer l.  `layers`/`dlayers` are
// the row-major (nl, 6) parameter and direction tables (dlayers only for
// dual Z).  The thickness enters only through the phase, whose cotangent
// is lc.h.
template <class Z>
__host__ __device__ __forceinline__ Z chain_column(const double* layers, const double* dlayers, int nl,
                                          int l, double p, const IfaceZ<Z>& cb_up,
                                          const IfaceZ<Z>& cb_dn, const LayerCot<Z>& lc,
                                          int col) {
  using W = typename Up<Z>::W;
  if (col == COL_H) return lc.h;
  // Layer ll with the outer seed on (l, col); one neighbour at a time, so
  // that two dual-of-dual layers and one interface are live at once.
  auto layer = [&](int ll) {
    W x[NCOL];
    for (int c = 0; c < NCOL; ++c) {
      const double dm = dlayers ? dlayers[(size_t)ll * NCOL + c] : 0.0;
      x[c] = Up<Z>::seed(layers[(size_t)ll * NCOL + c], dm, (ll == l && c == col) ? 1.0 : 0.0);
    }
    return layer_t<W>(x, ll, p);
  };
  Z acc;
  {
    const LayerZ<W> mid = layer(l);
    acc = re_dot(Up<Z>::outer(mid.eta), lc.eta) + re_dot(Up<Z>::outer(mid.neta), lc.neta);
    if (l >= 1) {
      const IfaceZ<W> f = interface_t<W>(l - 1, p, layer(l - 1), mid);
      acc = acc + iface_dot(outer_iface<W, Z>(f), cb_up);
    }
  }
  if (l <= nl - 2) {
    // layer l again rather than kept live across the interface above (the
    // same numbers; the opaque copy of l stops the compiler from merging
    // the two evaluations, which would keep 40 doubles live)
    int lo = l;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(lo));
#endif
    const IfaceZ<W> f = interface_t<W>(l, p, layer(lo), layer(l + 1));
    acc = acc + iface_dot(outer_iface<W, Z>(f), cb_dn);
  }
  return acc;
}

// The first-order factors of chain_column<cd> for parameter column col
// (not COL_H) of layer l at slowness p: d eta, d neta and the derivatives of
// the interfaces above (up, zero for l = 0) and below (dn, zero for
// l = nl-1).  They depend on neither the direction nor the cotangents, so
// one evaluation serves every direction of a launch.
struct LocalJac {
  cd eta, neta;
  IfaceZ<cd> up, dn;
};

__host__ __device__ inline LocalJac local_jac(const double* layers, int nl, int l, double p,
                                              int col) {
  using W = dcd;
  auto layer = [&](int ll) {  // as chain_column<cd> seeds it
    W x[NCOL];
    for (int c = 0; c < NCOL; ++c)
      x[c] = Up<cd>::seed(layers[(size_t)ll * NCOL + c], 0.0, (ll == l && c == col) ? 1.0 : 0.0);
    return layer_t<W>(x, ll, p);
  };
  const cd z{0.0, 0.0};
  const IfaceZ<cd> none{z, z, z, z, z, z, z, z, z, z};
  LocalJac j;
  const LayerZ<W> mid = layer(l);
  j.eta = Up<cd>::outer(mid.eta);
  j.neta = Up<cd>::outer(mid.neta);
  j.up = l >= 1 ? outer_iface<W, cd>(interface_t<W>(l - 1, p, layer(l - 1), mid)) : none;
  j.dn = l <= nl - 2 ? outer_iface<W, cd>(interface_t<W>(l, p, mid, layer(l + 1))) : none;
  return j;
}

// chain_column<cd> from its factors: the same products summed in the same
// order, so the two agree bit for bit.
__host__ __device__ inline cd chain_from_jac(const LocalJac& j, int l, int nl,
                                             const IfaceZ<cd>& cb_up, const IfaceZ<cd>& cb_dn,
                                             const LayerCot<cd>& lc, int col) {
  if (col == COL_H) return lc.h;
  cd acc = re_dot(j.eta, lc.eta) + re_dot(j.neta, lc.neta);
  if (l >= 1) acc = acc + iface_dot(j.up, cb_up);
  if (l <= nl - 2) acc = acc + iface_dot(j.dn, cb_dn);
  return acc;
}

// All six parameter cotangents of layer l (out[c] for c in COL_*).
template <class Z>
__host__ __device__ inline void chain_layer(const double* layers, const double* dlayers, int nl,
                                            int l, double p, const IfaceZ<Z>& cb_up,
                                            const IfaceZ<Z>& cb_dn, const LayerCot<Z>& lc,
                                            Z (&out)[NCOL]) {
  for (int col = 0; col < NCOL; ++col)
    out[col] = chain_column<Z>(layers, dlayers, nl, l, p, cb_up, cb_dn, lc, col);
}

}  // namespace kh
