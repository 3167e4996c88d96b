// Host build of the derivative algorithms (csrc/kh_deriv.hpp): the per-lane
// forward/checkpoint/reverse sweep and the parameter chain rule, run over the
// frequencies sequentially (the kernels reduce the same per-lane cotangents
// over a wavefront).  Compared on the CPU with torch autograd of the oracle.
#include "kh_deriv.hpp"

#include <vector>

using namespace kh;

template <class Z>
static IfaceZ<Z> add(const IfaceZ<Z>& a, const IfaceZ<Z>& b) {
  IfaceZ<Z> o;
  o.rd00 = a.rd00 + b.rd00; o.rd01 = a.rd01 + b.rd01; o.rd11 = a.rd11 + b.rd11;
  o.ru00 = a.ru00 + b.ru00; o.ru01 = a.ru01 + b.ru01; o.ru11 = a.ru11 + b.ru11;
  o.td00 = a.td00 + b.td00; o.td01 = a.td01 + b.td01; o.td10 = a.td10 + b.td10;
  o.td11 = a.td11 + b.td11;
  return o;
}

template <class Z>
struct Sink {
  std::vector<IfaceZ<Z>> cif;
  std::vector<LayerCot<Z>> cly;
  explicit Sink(int nl) : cif(nl), cly(nl) {
    const Z z = lift<Z>(0.0);
    for (auto& f : cif) f = IfaceZ<Z>{z, z, z, z, z, z, z, z, z, z};
    for (auto& c : cly) c = LayerCot<Z>{z, z, z};
  }
  void iface(int i, const IfaceZ<Z>& f) { cif[i] = add(cif[i], f); }
  void layer(int l, const LayerCot<Z>& c) {
    cly[l].eta = cly[l].eta + c.eta;
    cly[l].neta = cly[l].neta + c.neta;
    cly[l].h = cly[l].h + c.h;
  }
  void level(int i, const IfaceZ<Z>& f, const LayerCot<Z>& c) {
    iface(i, f);
    layer(i + 1, c);
  }
};

template <class Z>
struct Store {
  std::vector<Z> v;
  explicit Store(int n) : v(n) {}
  Z& operator()(int e) { return v[e]; }
};

template <class Z>
static void stage(const double* layers, const double* dlayers, int nl, double p,
                  std::vector<LayerZ<Z>>& L, std::vector<IfaceZ<Z>>& F) {
  for (int l = 0; l < nl; ++l) {
    Z x[NCOL];
    row_z(layers + (size_t)l * NCOL, dlayers ? dlayers + (size_t)l * NCOL : nullptr, x);
    L[l] = layer_t<Z>(x, l, p);
  }
  for (int i = 0; i < nl - 1; ++i) F[i] = interface_t<Z>(i, p, L[i], L[i + 1]);
}

// grad (nl, 6) of Re(sum_w conj(G_w) R_w); R (nw,) complex.
extern "C" int host_vjp(const double* layers, int nl, double p, const double* omega, int nw,
                        int fs, const double* G, double* grad, double* R) {
  std::vector<LayerZ<cd>> L(nl);
  std::vector<IfaceZ<cd>> F(nl - 1);
  stage<cd>(layers, nullptr, nl, p, L, F);
  Sink<cd> sink(nl);
  Store<cd> st(lane_slots_k(nl, 1));
  for (int w = 0; w < nw; ++w) {
    const cd om{omega[2 * w], omega[2 * w + 1]};
    const cd g{G[2 * w], G[2 * w + 1]};
    const cd r = lane_adjoint<cd>(L.data(), F.data(), nl, om, fs, st, sink,
                                  [&](const cd&) { return g; });
    R[2 * w] = r.r;
    R[2 * w + 1] = r.i;
  }
  const cd z{0.0, 0.0};
  const IfaceZ<cd> none{z, z, z, z, z, z, z, z, z, z};
  for (int l = 0; l < nl; ++l) {
    cd out[NCOL];
    chain_layer<cd>(layers, nullptr, nl, l, p, l >= 1 ? sink.cif[l - 1] : none,
                    l <= nl - 2 ? sink.cif[l] : none, sink.cly[l], out);
    for (int c = 0; c < NCOL; ++c) grad[l * NCOL + c] = out[c].r;
  }
  return 0;
}

// chi = sum |R - Robs|^2, its gradient, and (H dm) along the direction
// dlayers (nl, 6); dR = J dm.  second_only drops the Gauss-Newton part.
extern "C" int host_misfit_hvp(const double* layers, const double* dlayers, int nl, double p,
                               const double* omega, int nw, int fs, const double* Robs,
                               int second_only, double* chi, double* grad, double* hvp,
                               double* dR) {
  std::vector<LayerZ<dcd>> L(nl);
  std::vector<IfaceZ<dcd>> F(nl - 1);
  stage<dcd>(layers, dlayers, nl, p, L, F);
  Sink<dcd> sink(nl);
  Store<dcd> st(lane_slots_k(nl, 1));
  dcd chis = lift<dcd>(0.0);
  for (int w = 0; w < nw; ++w) {
    const cd om{omega[2 * w], omega[2 * w + 1]};
    const dcd ro = lift_c<dcd>(cd{Robs[2 * w], Robs[2 * w + 1]});
    const dcd r = lane_adjoint<dcd>(L.data(), F.data(), nl, om, fs, st, sink, [&](const dcd& R) {
      const dcd res = R - ro;
      chis = chis + re_part(conj(res) * res);
      dcd g = res * 2.0;
      if (second_only) g.d = cd{0.0, 0.0};
      return g;
    });
    dR[2 * w] = r.d.r;
    dR[2 * w + 1] = r.d.i;
  }
  chi[0] = chis.v.r;
  chi[1] = chis.d.r;
  const dcd z = lift<dcd>(0.0);
  const IfaceZ<dcd> none{z, z, z, z, z, z, z, z, z, z};
  for (int l = 0; l < nl; ++l) {
    dcd out[NCOL];
    chain_layer<dcd>(layers, dlayers, nl, l, p, l >= 1 ? sink.cif[l - 1] : none,
                     l <= nl - 2 ? sink.cif[l] : none, sink.cly[l], out);
    for (int c = 0; c < NCOL; ++c) {
      grad[l * NCOL + c] = out[c].v.r;
      hvp[l * NCOL + c] = out[c].d.r;
    }
  }
  return 0;
}

// host_misfit_hvp with the kernels' shortcuts: mode 1 runs the forward steps
// below the deepest moved layer on primal parts only (step_fwd_prim), mode 2
// also stops the reverse sweep one interface below it and leaves the deeper
// layers' entries of hvp at 0 (kh_misfit_hvp_partial).  Mode 0 is
// host_misfit_hvp.  Modes 5, 6, 7 are modes 2, 3, 4 with the steps above the
// shallowest moved layer on plain-cd coefficients and phases.
extern "C" int host_misfit_hvp_mode(const double* layers, const double* dlayers, int nl,
                                    double p, const double* omega, int nw, int fs,
                                    const double* Robs, int second_only, int mode, double* chi,
                                    double* grad, double* hvp, double* dR) {
  int ltan = nl - 1;
  for (; ltan >= 0; --ltan) {
    bool moved = false;
    for (int c = 0; c < NCOL; ++c) moved = moved || dlayers[ltan * NCOL + c] != 0.0;
    if (moved) break;
  }
  int lsh = 0;  // modes 5-7: modes 2-4 with the plain-cd steps above the shallowest moved layer
  for (; lsh < nl; ++lsh) {
    bool moved = false;
    for (int c = 0; c < NCOL; ++c) moved = moved || dlayers[lsh * NCOL + c] != 0.0;
    if (moved) break;
  }
  const int lsh_used = mode >= 5 ? lsh : 0;
  if (mode >= 5) mode -= 3;
  const int lt = mode >= 1 ? ltan : (1 << 30);
  const int lrev = mode >= 2 ? ltan + 1 : (1 << 30);
  // mode 3 is mode 2 with checkpoints every 2 levels and the buffer in a
  // store of its own; mode 4 is mode 2 with sqrt(N) checkpoints and the
  // buffer after them (the layout of a workspace too small for the tape)
  std::vector<LayerZ<dcd>> L(nl);
  std::vector<IfaceZ<dcd>> F(nl - 1);
  stage<dcd>(layers, dlayers, nl, p, L, F);
  Sink<dcd> sink(nl);
  Store<dcd> st(lane_slots_k(nl, 1) + lane_slots_k(nl, ckpt_sqrt(nl)) + 64);
  Store<dcd> bst(2 * 6);  // mode 3: two-level segments, buffer apart (the LDS layout)
  dcd chis = lift<dcd>(0.0);
  for (int w = 0; w < nw; ++w) {
    const cd om{omega[2 * w], omega[2 * w + 1]};
    const dcd ro = lift_c<dcd>(cd{Robs[2 * w], Robs[2 * w + 1]});
    auto G = [&](const dcd& R) {
      const dcd res = R - ro;
      chis = chis + re_part(conj(res) * res);
      dcd g = res * 2.0;
      if (second_only) g.d = cd{0.0, 0.0};
      return g;
    };
    dcd r;
    if (mode == 3) {
      r = lane_adjoint_k<dcd>(L.data(), F.data(), nl, om, fs, st, bst, 2, sink, G, ltan,
                              ltan + 1, lsh_used);
    } else if (mode == 4) {
      const int K = ckpt_sqrt(nl);
      BufAfter<dcd, Store<dcd>> after{st, ckpt_count(nl, K) * 4};
      r = lane_adjoint_k<dcd>(L.data(), F.data(), nl, om, fs, st, after, K, sink, G, ltan,
                              ltan + 1, lsh_used);
    } else {
      r = lane_adjoint<dcd>(L.data(), F.data(), nl, om, fs, st, sink, G, lt, lrev, 1, lsh_used);
    }
    dR[2 * w] = r.d.r;
    dR[2 * w + 1] = r.d.i;
  }
  chi[0] = chis.v.r;
  chi[1] = chis.d.r;
  const dcd z = lift<dcd>(0.0);
  const IfaceZ<dcd> none{z, z, z, z, z, z, z, z, z, z};
  const int lmax = mode >= 2 ? ltan + 1 : nl - 1;
  for (int l = 0; l < nl; ++l) {
    dcd out[NCOL];
    if (l > lmax) {
      for (int c = 0; c < NCOL; ++c) hvp[l * NCOL + c] = 0.0;
      continue;
    }
    chain_layer<dcd>(layers, dlayers, nl, l, p, l >= 1 ? sink.cif[l - 1] : none,
                     l <= nl - 2 ? sink.cif[l] : none, sink.cly[l], out);
    for (int c = 0; c < NCOL; ++c) {
      grad[l * NCOL + c] = out[c].v.r;
      hvp[l * NCOL + c] = out[c].d.r;
    }
  }
  return 0;
}

// Every (layer, column) of chain_column<cd> directly and from the shared
// first-order factors (local_jac, then chain_from_jac), for the cotangents
// cif ((nl - 1) interfaces x 10 complex) and cly (nl layers x 3 complex);
// outputs (nl, 6) complex each.
extern "C" int host_chain_factored(const double* layers, int nl, double p, const double* cif,
                                   const double* cly, double* direct, double* factored) {
  const cd* ci = reinterpret_cast<const cd*>(cif);
  const cd* cl = reinterpret_cast<const cd*>(cly);
  auto iface = [&](int i) {
    const cd* c = ci + (size_t)i * 10;
    return IfaceZ<cd>{c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9]};
  };
  const cd z{0.0, 0.0};
  const IfaceZ<cd> none{z, z, z, z, z, z, z, z, z, z};
  for (int l = 0; l < nl; ++l) {
    const IfaceZ<cd> up = l >= 1 ? iface(l - 1) : none, dn = l <= nl - 2 ? iface(l) : none;
    const LayerCot<cd> lc{cl[3 * l], cl[3 * l + 1], cl[3 * l + 2]};
    for (int c = 0; c < NCOL; ++c) {
      const cd a = chain_column<cd>(layers, nullptr, nl, l, p, up, dn, lc, c);
      const cd b = chain_from_jac(c == COL_H ? LocalJac{} : local_jac(layers, nl, l, p, c), l, nl,
                                  up, dn, lc, c);
      direct[2 * (l * NCOL + c)] = a.r;
      direct[2 * (l * NCOL + c) + 1] = a.i;
      factored[2 * (l * NCOL + c)] = b.r;
      factored[2 * (l * NCOL + c) + 1] = b.i;
    }
  }
  return 0;
}
