// Host build of the device arithmetic (csrc/kh_physics.hpp, csrc/kh_sweep.hpp):
// the same source lines the kernels run, compiled for the CPU so the CPU test
// suite can compare them with the NumPy oracle without a GPU.
#include "kh_sweep.hpp"

#include <vector>

using namespace kh;

extern "C" int host_forward(const double* layers, int nl, const double* p, int np,
                            const double* omega, int nw, int free_surface, double* R) {
  std::vector<LayerPt> L(nl);
  std::vector<Iface> F(nl - 1);
  for (int j = 0; j < np; ++j) {
    for (int l = 0; l < nl; ++l) L[l] = layer_at(layers + (size_t)l * NCOL, l, p[j]);
    for (int i = 0; i < nl - 1; ++i) F[i] = interface_at(i, p[j], L[i], L[i + 1]);
    for (int w = 0; w < nw; ++w) {
      const cd om{omega[2 * w], omega[2 * w + 1]};
      const cd r = sweep_lane(L.data(), F.data(), nl, om, free_surface);
      R[2 * ((size_t)j * nw + w)] = r.r;
      R[2 * ((size_t)j * nw + w) + 1] = r.i;
    }
  }
  return 0;
}
