// Host build of the device arithmetic (csrc/kh_physics.hpp, csrc/kh_sweep.hpp):
// the same source lines the kernels run, compiled for the CPU so the CPU test
// suite can compare them with the NumPy oracle without a GPU.
#include "kh_gmm_lanes.hpp"
#include "kh_j0.hpp"
#include "kh_sweep.hpp"
#include "kh_wave.hpp"

#include <vector>

using namespace kh;

static void stage(const double* layers, int nl, double p, std::vector<LayerPt>& L,
                  std::vector<Iface>& F) {
  for (int l = 0; l < nl; ++l) L[l] = layer_at(layers + (size_t)l * NCOL, l, p);
  for (int i = 0; i < nl - 1; ++i) F[i] = interface_at(i, p, L[i], L[i + 1]);
}

extern "C" int host_forward(const double* layers, int nl, const double* p, int np,
                            const double* omega, int nw, int free_surface, double* R) {
  std::vector<LayerPt> L(nl);
  std::vector<Iface> F(nl - 1);
  for (int j = 0; j < np; ++j) {
    stage(layers, nl, p[j], L, F);
    for (int w = 0; w < nw; ++w) {
      const cd om{omega[2 * w], omega[2 * w + 1]};
      const cd r = sweep_lane(L.data(), F.data(), nl, om, free_surface);
      R[2 * ((size_t)j * nw + w)] = r.r;
      R[2 * ((size_t)j * nw + w) + 1] = r.i;
    }
  }
  return 0;
}

// The uniform-grid path of the kernel: frequencies omega_j = (k0 + j) dw + i gamma,
// four per lane, `block` apart (as forward_kernel<4, true> with blockDim = block).
extern "C" int host_forward_grid(const double* layers, int nl, const double* p, int np, double k0,
                                 double dw, double gamma, int nw, int free_surface, int block,
                                 double* R) {
  constexpr int W = 4;
  std::vector<LayerPt> L(nl);
  std::vector<Iface> F(nl - 1);
  std::vector<LayerStep> S(nl);
  for (int j = 0; j < np; ++j) {
    stage(layers, nl, p[j], L, F);
    for (int l = 0; l < nl; ++l) S[l] = layer_step(L[l], (double)block * dw);
    const int span = block * W;
    for (int tile = 0; tile * span < nw; ++tile) {
      for (int t = 0; t < block; ++t) {
        const int base = tile * span + t;
        cd om[W];
        for (int k = 0; k < W; ++k) {
          int w = base + k * block;
          if (w > nw - 1) w = nw - 1;
          om[k] = cd{(k0 + (double)w) * dw, gamma};
        }
        cd out[W];
        sweep_lanes<W, true>(L.data(), F.data(), S.data(), nl, om, free_surface, out);
        for (int k = 0; k < W; ++k) {
          const int w = base + k * block;
          if (w < nw) {
            R[2 * ((size_t)j * nw + w)] = out[k].r;
            R[2 * ((size_t)j * nw + w) + 1] = out[k].i;
          }
        }
      }
    }
  }
  return 0;
}

// The phase factors' sine and cosine (csrc/kh_complex.hpp) for n arguments.
extern "C" void host_sincos_phase(const double* x, int n, double* s, double* c) {
  for (int i = 0; i < n; ++i) kh::sincos_phase(x[i], s + i, c + i);
}

// J0 (csrc/kh_j0.hpp) for n arguments: the stand-alone form, and the large
// regime with u = 1/x and amp = sqrt(2/(pi x)) passed in as the gather
// kernel passes its per-lane and per-slowness products.
extern "C" void host_j0(const double* x, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = kh::j0_kh(x[i]);
}

extern "C" void host_j0_split(const double* x, const double* u, const double* amp, int n,
                              double* out) {
  for (int i = 0; i < n; ++i) {
    const double a = fabs(x[i]);
    out[i] = a < kh::J0_XA ? kh::j0_small(a, &kh::J0_SMALL[0][0]) : kh::j0_large(a, u[i], amp[i]);
  }
}

// The Global Matrix Method lane (csrc/kh_gmm_lanes.hpp), one frequency per
// lane, explicit frequencies.
extern "C" int host_gmm_forward(const double* layers, int nl, const double* p, int np,
                                const double* omega, int nw, int free_surface, double* R) {
  std::vector<LayerPt> L(nl);
  std::vector<EigD> E(nl);
  for (int j = 0; j < np; ++j) {
    for (int l = 0; l < nl; ++l) L[l] = layer_at(layers + (size_t)l * NCOL, l, p[j]);
    for (int l = 0; l < nl; ++l) E[l] = eig_d(p[j], L[l]);
    for (int w = 0; w < nw; ++w) {
      const cd om[1] = {cd{omega[2 * w], omega[2 * w + 1]}};
      cd out[1];
      gmm_lanes<1, false>(L.data(), E.data(), nullptr, nl, om, free_surface, out);
      R[2 * ((size_t)j * nw + w)] = out[0].r;
      R[2 * ((size_t)j * nw + w) + 1] = out[0].i;
    }
  }
  return 0;
}

// The lane -> value map of the device's wave reduce-scatter (kh_wave.hpp).
extern "C" int host_wave_rs_index(int n, int lane) { return wave_rs_index(n, lane); }
