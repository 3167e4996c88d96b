// Lane driver of the block-Riccati reflectivity (kh_gmm.hpp) for W
// frequencies of one slowness, shared by the device kernel (kh_gmm.hip) and
// the host build used by the CPU tests.
#pragma once

#include "kh_gmm.hpp"
#include "kh_sweep.hpp"

namespace kh {

// Phase pairs (exp(i w eta h), exp(i w nu h)) of layer L at W frequencies;
// with GRID, om[k] = om[0] + k * shift and `st` holds the factors for that shift.
template <int W, bool GRID>
__host__ __device__ __forceinline__ void layer_phases(const LayerPt& L, const LayerStep st,
                                                      const cd (&om)[W], cd (&ph)[W][2]) {
  if (L.halfspace) {
#pragma unroll
    for (int k = 0; k < W; ++k) ph[k][0] = ph[k][1] = cd{1.0, 0.0};
  } else if (GRID) {
    ph[0][0] = phase(L.eta, om[0], L.h);
    ph[0][1] = phase(L.neta, om[0], L.h);
#pragma unroll
    for (int k = 1; k < W; ++k) {
      ph[k][0] = ph[k - 1][0] * st.a;
      ph[k][1] = ph[k - 1][1] * st.b;
    }
  } else {
#pragma unroll
    for (int k = 0; k < W; ++k) {
      ph[k][0] = phase(L.eta, om[k], L.h);
      ph[k][1] = phase(L.neta, om[k], L.h);
    }
  }
}

template <int W, bool GRID>
__host__ __device__ __forceinline__ void gmm_lanes(const LayerPt* __restrict__ L,
                                                   const EigD* __restrict__ E,
                                                   const LayerStep* __restrict__ steps, int nl,
                                                   const cd (&om)[W], int free_surface,
                                                   cd (&out)[W]) {
  const LayerStep unit{{1.0, 0.0}, {1.0, 0.0}};
  cd Y[W][2][2];
  cd phb[W][2];
#pragma unroll
  for (int k = 0; k < W; ++k) {
    Y[k][0][0] = Y[k][0][1] = Y[k][1][0] = Y[k][1][1] = cd{0.0, 0.0};
    phb[k][0] = phb[k][1] = cd{1.0, 0.0};
  }
  for (int j = nl - 2; j >= 1; --j) {  // finite solid layers, deepest first
    cd phk[W][2];
    LayerStep st = unit;  // a value, never a reference into LDS or the stack
    if (GRID) st = steps[j];
    layer_phases<W, GRID>(L[j], st, om, phk);
    const bool below_half = (j + 1 == nl - 1);
#pragma unroll
    for (int k = 0; k < W; ++k) {
      riccati_step(Y[k], E[j], phk[k], E[j + 1], phb[k], below_half);
      phb[k][0] = phk[k][0];
      phb[k][1] = phk[k][1];
    }
  }
  cd e0[W][2];
  LayerStep st0 = unit;
  if (GRID) st0 = steps[0];
  layer_phases<W, GRID>(L[0], st0, om, e0);
#pragma unroll
  for (int k = 0; k < W; ++k)
    out[k] = riccati_surface(Y[k], E[1], phb[k], nl == 2, L[0].eta, L[0].rho, e0[k][0],
                             free_surface);
}

}  // namespace kh
