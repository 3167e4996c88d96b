// The per-lane Kennett upward sweep and the workgroup prologue that stages
// one slowness's layer and interface tables in LDS.  Shared by the forward,
// multi-model and gather kernels so every caller runs identical arithmetic
// for a lane.
//
// A lane carries the full 2x2 RRd and applies the reference's update
// (kennett_reflectivity_gpu.py:235-257) in the same association:
//     MT = E RRd E,  U = (I - Ru MT)^-1,  RRd = Rd + (Tu (MT U)) Td,
// with Tu read as Td^T and U formed as adj(I - Ru MT) times one reciprocal
// of the determinant.  (Dropping the lower triangle of the symmetric RRd
// saves ~10% of the FLOPs but changes the rounding enough to leave the
// reference's result at large slowness, so the full form is kept.)
//
// A lane may carry W frequencies at once (the interface coefficients are
// then read once for all W).  On a uniform grid omega_j = (k0 + j) dw + i gamma
// the W frequencies of a lane are B apart, and the phase factors of the
// 2nd..W-th come from the first by the exact identity
//     exp(i (w + B dw) eta h) = exp(i w eta h) * exp(i B dw eta h),
// with exp(i B dw eta h) computed once per layer in the prologue.
#pragma once

#include "kh_physics.hpp"

// Read-only global memory: loads at wave-uniform addresses through such a
// pointer become scalar loads into scalar registers.
#if defined(__HIP_DEVICE_COMPILE__)
#define KH_CONST __attribute__((address_space(4)))
#else
#define KH_CONST
#endif

namespace kh {

struct LayerStep {  // per-layer phase step for a frequency shift of B*dw
  cd a, b;
};

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

__host__ __device__ __forceinline__ LayerStep layer_step(const LayerPt& L, double shift) {
  LayerStep s{{1.0, 0.0}, {1.0, 0.0}};
  if (!L.halfspace) {
    const cd w{shift, 0.0};
    s.a = phase(L.eta, w, L.h);
    s.b = phase(L.neta, w, L.h);
  }
  return s;
}

// One interface step for one frequency: RRd <- Rd + Tu MT U Td, with
// MT = E RRd E, U = (I - Ru MT)^-1, Tu = Td^T, given the phase products
// eaea = ea^2, ebeb = eb^2, eaeb = ea eb of E = diag(ea, eb).  Every sum of
// products is one chain of fused multiply-adds seeded with its first product
// or its constant term (1 - ..., Rd + ...), which keeps the FP64 instruction
// count near one per two FLOPs.
__host__ __device__ __forceinline__ void kennett_step_e2(cd& r00, cd& r01, cd& r10, cd& r11,
                                                         cd eaea, cd ebeb, cd eaeb,
                                                         const Iface& f) {
  const cd m00 = eaea * r00, m01 = eaeb * r01, m10 = eaeb * r10, m11 = ebeb * r11;
  const cd one{1.0, 0.0};
  // A = I - Ru MT (Ru symmetric)
  const cd a00 = cfms(f.ru01, m10, cfms(f.ru00, m00, one));
  const cd a01 = cfms(f.ru01, m11, -(f.ru00 * m01));
  const cd a10 = cfms(f.ru11, m10, -(f.ru01 * m00));
  const cd a11 = cfms(f.ru11, m11, cfms(f.ru01, m01, one));
  const cd rdet = recip_det(cfms(a01, a10, a00 * a11));
  const cd u00 = a11 * rdet, u01 = -(a01 * rdet), u10 = -(a10 * rdet), u11 = a00 * rdet;
  // X = MT U
  const cd x00 = cfma(m01, u10, m00 * u00), x01 = cfma(m01, u11, m00 * u01);
  const cd x10 = cfma(m11, u10, m10 * u00), x11 = cfma(m11, u11, m10 * u01);
  // Y = Tu X with Tu = Td^T
  const cd y00 = cfma(f.td10, x10, f.td00 * x00), y01 = cfma(f.td10, x11, f.td00 * x01);
  const cd y10 = cfma(f.td11, x10, f.td01 * x00), y11 = cfma(f.td11, x11, f.td01 * x01);
  // RRd = Rd + Y Td
  r00 = cfma(y01, f.td10, cfma(y00, f.td00, f.rd00));
  r01 = cfma(y01, f.td11, cfma(y00, f.td01, f.rd01));
  r10 = cfma(y11, f.td10, cfma(y10, f.td00, f.rd01));
  r11 = cfma(y11, f.td11, cfma(y10, f.td01, f.rd11));
}

// The same step from the phase factors themselves.
__host__ __device__ __forceinline__ void kennett_step(cd& r00, cd& r01, cd& r10, cd& r11, cd ea,
                                                      cd eb, const Iface& f) {
  kennett_step_e2(r00, r01, r10, r11, ea * ea, eb * eb, ea * eb, f);
}

// The interface coefficients of a layer in scalar registers (wave_uniform):
// the step's FMAs then take them as their one scalar operand instead of
// holding 20 doubles per layer in vector registers (which is what lets the
// 8-frequency sweep run two waves per SIMD).
__host__ __device__ __forceinline__ Iface wave_uniform(const Iface& f) {
  return Iface{wave_uniform(f.rd00), wave_uniform(f.rd01), wave_uniform(f.rd11),
               wave_uniform(f.ru00), wave_uniform(f.ru01), wave_uniform(f.ru11),
               wave_uniform(f.td00), wave_uniform(f.td01), wave_uniform(f.td10),
               wave_uniform(f.td11)};
}

// RRd00 just below the water layer for W frequencies of one lane.  With
// GRID, om[k] = om[0] + k * shift (real shift) and `steps` holds the
// per-layer factors for that shift: the phase products of frequency k are
// those of frequency k-1 times the products of the factors (ea^2 steps by
// a^2, eb^2 by b^2, ea eb by a b), evaluated once per layer and lane.
// Otherwise every phase is evaluated.  With UNIFORM the interface tables
// must be the same for every lane of the wave (one slowness per workgroup)
// and are moved to scalar registers.  LP, FP and SP are pointers to the
// tables, in LDS or in read-only global memory (KH_CONST, read with scalar
// loads at wave-uniform addresses).
template <int W, bool GRID, bool UNIFORM = true, class LP, class FP, class SP>
__host__ __device__ __forceinline__ void sweep_rrd00(LP L, FP F, SP steps, int nl,
                                                     const cd (&om)[W], cd (&out)[W]) {
  cd q00[W], q01[W], q10[W], q11[W];
  {
    const Iface fb = F[nl - 2];  // below it RRd = 0, so RRd = Rd
#pragma unroll
    for (int k = 0; k < W; ++k) {
      q00[k] = fb.rd00;
      q01[k] = fb.rd01;
      q10[k] = fb.rd01;
      q11[k] = fb.rd11;
    }
  }
  const cd one{1.0, 0.0};
  for (int i = nl - 3; i >= 0; --i) {
    auto&& lb = L[i + 1];
    Iface f;
    if constexpr (UNIFORM) f = wave_uniform(F[i]);
    else f = F[i];
    const bool hs = lb.halfspace;  // unit phase factors
    cd aa = one, bb = one, ab = one, sa2 = one, sb2 = one, sab = one;
    if (GRID && !hs) {
      const cd ea = phase(lb.eta, om[0], lb.h);
      const cd eb = phase(lb.neta, om[0], lb.h);
      aa = ea * ea;
      bb = eb * eb;
      ab = ea * eb;
      if (W > 1) {
        const LayerStep s = steps[i + 1];
        sa2 = s.a * s.a;
        sb2 = s.b * s.b;
        sab = s.a * s.b;
      }
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
      if (GRID) {
        if (k > 0) {
          aa = aa * sa2;
          bb = bb * sb2;
          ab = ab * sab;
        }
      } else if (!hs) {
        const cd ea = phase(lb.eta, om[k], lb.h);
        const cd eb = phase(lb.neta, om[k], lb.h);
        aa = ea * ea;
        bb = eb * eb;
        ab = ea * eb;
      }
      kennett_step_e2(q00[k], q01[k], q10[k], q11[k], aa, bb, ab, f);
    }
  }
#pragma unroll
  for (int k = 0; k < W; ++k) out[k] = q00[k];
}

// Water-layer phase and optional free surface (kennett_reflectivity_gpu.py:259-272).
template <int W, bool GRID, class LT>
__host__ __device__ __forceinline__ void surface_pp(const LT& water, const LayerStep& step,
                                                    const cd (&om)[W], cd (&r)[W],
                                                    int free_surface) {
  cd ea0[W];
  if (water.halfspace) {
#pragma unroll
    for (int k = 0; k < W; ++k) ea0[k] = cd{1.0, 0.0};
  } else if (GRID) {
    ea0[0] = phase(water.eta, om[0], water.h);
#pragma unroll
    for (int k = 1; k < W; ++k) ea0[k] = ea0[k - 1] * step.a;
  } else {
#pragma unroll
    for (int k = 0; k < W; ++k) ea0[k] = phase(water.eta, om[k], water.h);
  }
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const cd top = (ea0[k] * ea0[k]) * r[k];
    r[k] = free_surface ? top / (cd{1.0, 0.0} + top) : top;
  }
}

// Full lane: PP reflectivity at W frequencies (UNIFORM and the table
// pointers as sweep_rrd00).
template <int W, bool GRID, bool UNIFORM = true, class LP, class FP, class SP>
__host__ __device__ __forceinline__ void sweep_lanes(LP L, FP F, SP steps, int nl,
                                                     const cd (&om)[W], int free_surface,
                                                     cd (&out)[W]) {
  sweep_rrd00<W, GRID, UNIFORM>(L, F, steps, nl, om, out);
  LayerStep s0{{1.0, 0.0}, {1.0, 0.0}};
  if (GRID) s0 = steps[0];
  surface_pp<W, GRID>(L[0], s0, om, out, free_surface);
}

// Single frequency, generic omega (host tests and simple callers; the
// tables may differ from lane to lane).
__host__ __device__ __forceinline__ cd sweep_lane(const LayerPt* __restrict__ L,
                                                  const Iface* __restrict__ F, int nl, cd om,
                                                  int free_surface) {
  const cd o[1] = {om};
  cd r[1];
  sweep_lanes<1, false, false>(L, F, static_cast<const LayerStep*>(nullptr), nl, o,
                               free_surface, r);
  return r[0];
}

}  // namespace kh
