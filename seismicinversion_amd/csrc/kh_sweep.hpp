// The per-lane Kennett upward sweep and the workgroup prologue that stages
// one slowness's layer and interface tables in LDS.  Shared by the forward,
// multi-model and gather kernels so that every caller runs the identical
// arithmetic for a lane.
#pragma once

#include "kh_physics.hpp"

namespace kh {

// Workgroup-cooperative: layer slownesses of every layer at slowness pv,
// then the coefficients of every interface.  Ends with a barrier.
__device__ __forceinline__ void stage_slowness(const double* __restrict__ layers, int nl,
                                               double pv, LayerPt* s_layer, Iface* s_if) {
  for (int l = threadIdx.x; l < nl; l += blockDim.x)
    s_layer[l] = layer_at(layers + (size_t)l * NCOL, l, pv);
  __syncthreads();
  for (int i = threadIdx.x; i < nl - 1; i += blockDim.x)
    s_if[i] = interface_at(i, pv, s_layer[i], s_layer[i + 1]);
  __syncthreads();
}

struct Mat2 {
  cd a00, a01, a10, a11;
};

// RRd at the top of the stack (just below the water layer) for one omega.
__host__ __device__ __forceinline__ Mat2 sweep_rrd(const LayerPt* __restrict__ L,
                                         const Iface* __restrict__ F, int nl, cd om) {
  // Deepest interface: RRd below it is 0, so RRd = Rd.
  const Iface& fb = F[nl - 2];
  Mat2 r{fb.rd00, fb.rd01, fb.rd01, fb.rd11};
  for (int i = nl - 3; i >= 0; --i) {
    const LayerPt& lb = L[i + 1];
    cd ea{1.0, 0.0}, eb{1.0, 0.0};
    if (!lb.halfspace) {
      ea = phase(lb.eta, om, lb.h);
      eb = phase(lb.neta, om, lb.h);
    }
    const cd eaea = ea * ea, ebeb = eb * eb, eaeb = ea * eb;
    const cd m00 = eaea * r.a00, m01 = eaeb * r.a01, m10 = eaeb * r.a10, m11 = ebeb * r.a11;
    const Iface& f = F[i];
    // A = I - Ru MT
    const cd one{1.0, 0.0};
    const cd a00 = one - (f.ru00 * m00 + f.ru01 * m10);
    const cd a01 = -(f.ru00 * m01 + f.ru01 * m11);
    const cd a10 = -(f.ru01 * m00 + f.ru11 * m10);
    const cd a11 = one - (f.ru01 * m01 + f.ru11 * m11);
    const cd rdet = recip(a00 * a11 - a01 * a10);
    const cd u00 = a11 * rdet, u01 = -a01 * rdet, u10 = -a10 * rdet, u11 = a00 * rdet;
    // X = MT U
    const cd x00 = m00 * u00 + m01 * u10, x01 = m00 * u01 + m01 * u11;
    const cd x10 = m10 * u00 + m11 * u10, x11 = m10 * u01 + m11 * u11;
    // Y = Tu X, with Tu = Td^T
    const cd y00 = f.td00 * x00 + f.td10 * x10, y01 = f.td00 * x01 + f.td10 * x11;
    const cd y10 = f.td01 * x00 + f.td11 * x10, y11 = f.td01 * x01 + f.td11 * x11;
    // RRd = Rd + Y Td
    r.a00 = f.rd00 + (y00 * f.td00 + y01 * f.td10);
    r.a01 = f.rd01 + (y00 * f.td01 + y01 * f.td11);
    r.a10 = f.rd01 + (y10 * f.td00 + y11 * f.td10);
    r.a11 = f.rd11 + (y10 * f.td01 + y11 * f.td11);
  }
  return r;
}

// PP reflectivity seen from the sea surface (kennett_reflectivity_gpu.py:259-272).
__host__ __device__ __forceinline__ cd surface_pp(const LayerPt& water, cd rrd00, cd om, int free_surface) {
  cd ea0{1.0, 0.0};
  if (!water.halfspace) ea0 = phase(water.eta, om, water.h);
  const cd top = (ea0 * ea0) * rrd00;
  if (free_surface) return top / (cd{1.0, 0.0} + top);
  return top;
}

__host__ __device__ __forceinline__ cd sweep_lane(const LayerPt* __restrict__ L, const Iface* __restrict__ F,
                                         int nl, cd om, int free_surface) {
  const Mat2 r = sweep_rrd(L, F, nl, om);
  return surface_pp(L[0], r.a00, om, free_surface);
}

}  // namespace kh
