// The layer and interface physics of kh_physics.hpp written once over a
// scalar type Z (cd, Dual<cd>, Dual<Dual<cd>>), for the derivative kernels.
// Real layer parameters are carried as Z with zero imaginary parts; every
// primal operation is the one kh_physics.hpp performs (real quotients via
// div_real, complex ones via Smith division), so the primal values agree.
//
//   complex slowness ........ Kennett_Reflectivity/layer_model.py:42-45
//   vertical slowness ....... Kennett_Reflectivity/layer_model.py:69-74
//   solid-solid interface ... Kennett_Reflectivity/scattering_matrices.py:82-161
//   ocean-bottom interface .. Kennett_Reflectivity/scattering_matrices.py:209-298
#pragma once

#include "kh_dual.hpp"
#include "kh_physics.hpp"

namespace kh {

template <class Z>
struct LayerZ {
  Z eta, neta, beta_c, rho, h;
  int halfspace;
};

template <class Z>
struct IfaceZ {
  Z rd00, rd01, rd11, ru00, ru01, ru11, td00, td01, td10, td11;
};

template <class Z>
__host__ __device__ __forceinline__ Z complex_slowness_t(const Z& v, const Z& q) {
  const Z two_q = q * 2.0;
  const Z two_q_sq = two_q * two_q;
  const Z denom = (lift<Z>(1.0) + two_q_sq) * v;
  return div_real(two_q_sq, denom) + mul_i(div_real(two_q, denom));
}

template <class Z>
__host__ __device__ __forceinline__ Z vertical_slowness_t(const Z& s, double p) {
  const Z cp = lift<Z>(p);
  const Z t = (s + cp) * (s - cp);
  Z e = csqrt(t);
  if (prim(e).i <= 0.0) e = -e;
  return e;
}

// Layer l at slowness p from its six parameters x[COL_*] (as Z).
template <class Z>
__host__ __device__ __forceinline__ LayerZ<Z> layer_t(const Z (&x)[NCOL], int l, double p) {
  LayerZ<Z> L;
  const Z sp = complex_slowness_t(x[COL_ALPHA], x[COL_QA]);
  Z ss = lift<Z>(0.0);
  if (prim(x[COL_BETA]).r > 0.0) ss = complex_slowness_t(x[COL_BETA], x[COL_QB]);
  const cd ssp = prim(ss);
  L.beta_c = (ssp.r != 0.0 || ssp.i != 0.0) ? recip(ss) : lift<Z>(0.0);
  L.eta = vertical_slowness_t(sp, p);
  L.neta = (l == 0) ? lift<Z>(0.0) : vertical_slowness_t(ss, p);
  L.rho = x[COL_RHO];
  L.h = x[COL_H];
  L.halfspace = is_inf(prim(x[COL_H]).r) ? 1 : 0;
  return L;
}

template <class Z>
__host__ __device__ inline IfaceZ<Z> solid_solid_t(double p, const LayerZ<Z>& l1,
                                                   const LayerZ<Z>& l2) {
  const Z eta1 = l1.eta, neta1 = l1.neta, beta1 = l1.beta_c;
  const Z eta2 = l2.eta, neta2 = l2.neta, beta2 = l2.beta_c;
  const Z rtrho1 = csqrt(l1.rho), rtrho2 = csqrt(l2.rho);
  const Z rteta1 = csqrt(eta1), rteta2 = csqrt(eta2);
  const Z rtneta1 = csqrt(neta1), rtneta2 = csqrt(neta2);
  const Z rtza1 = rteta1 * rtrho1, rtza2 = rteta2 * rtrho2;
  const Z rtzb1 = rtneta1 * rtrho1, rtzb2 = rtneta2 * rtrho2;
  const Z psq = lift<Z>(p * p);
  const Z cp = lift<Z>(p);
  const Z crho1 = l1.rho, crho2 = l2.rho;
  const Z drho = crho2 - crho1;
  const Z dmu = crho2 * beta2 * beta2 - crho1 * beta1 * beta1;
  const Z d = dmu * 2.0;
  const Z psqd = psq * d;
  const Z a = drho - psqd, b = crho2 - psqd, c = crho1 + psqd;
  Z E = b * eta1 + c * eta2;
  Z F = b * neta1 + c * neta2;
  Z G = a - d * eta1 * neta2;
  Z H = a - d * eta2 * neta1;
  const Z det = E * F + G * H * psq;
  E = E / det;
  F = F / det;
  G = G / det;
  H = H / det;
  const Z q_v = (b * eta1 - c * eta2) * F;
  const Z r_v = (a + d * eta1 * neta2) * H * psq;
  const Z s_v = (a * b + c * d * eta2 * neta2) * cp / det;
  const Z t_v = (b * neta1 - c * neta2) * E;
  const Z u_v = (a + d * eta2 * neta1) * G * psq;
  const Z v_v = (a * c + b * d * eta1 * neta1) * cp / det;
  const Z m2ci = lift_c<Z>(cd{0.0, -2.0});
  const Z two = lift<Z>(2.0);
  IfaceZ<Z> f;
  f.rd00 = q_v - r_v;
  f.rd01 = m2ci * rteta1 * rtneta1 * s_v;
  f.rd11 = t_v - u_v;
  f.td00 = two * rtza1 * rtza2 * F;
  f.td01 = m2ci * rtzb1 * rtza2 * G * cp;
  f.td10 = m2ci * rtza1 * rtzb2 * H * cp;
  f.td11 = two * rtzb1 * rtzb2 * E;
  f.ru00 = -(q_v + u_v);
  f.ru01 = m2ci * rteta2 * rtneta2 * v_v;
  f.ru11 = -(t_v + r_v);
  return f;
}

template <class Z>
__host__ __device__ inline IfaceZ<Z> ocean_bottom_t(double p, const LayerZ<Z>& l1,
                                                    const LayerZ<Z>& l2) {
  const Z eta1 = l1.eta;
  const Z eta2 = l2.eta, neta2 = l2.neta, beta2 = l2.beta_c;
  const Z rtrho1 = csqrt(l1.rho), rtrho2 = csqrt(l2.rho);
  const Z rteta1 = csqrt(eta1), rteta2 = csqrt(eta2), rtneta2 = csqrt(neta2);
  const Z rtza1 = rtrho1 * rteta1, rtza2 = rtrho2 * rteta2, rtzb2 = rtrho2 * rtneta2;
  const Z psq = lift<Z>(p * p);
  const Z cp = lift<Z>(p);
  const Z crho1 = l1.rho, crho2 = l2.rho;
  const Z drho = crho2 - crho1;
  const Z dmu = crho2 * beta2 * beta2;
  const Z d = dmu * 2.0;
  const Z psqd = psq * d;
  const Z a = drho - psqd, b = crho2 - psqd, c = crho1 + psqd;
  Z E = b * eta1 + c * eta2;
  Z F = b;
  Z G = a - d * eta1 * neta2;
  Z H = (-d) * eta2;
  const Z det = E * F + G * H * psq;
  E = E / det;
  F = F / det;
  G = G / det;
  H = H / det;
  const Z t1 = (b * eta1 - c * eta2) * F;
  const Z t2 = (a + d * eta1 * neta2) * H * psq;
  const Z t4 = b * E;
  const Z t5 = d * eta2 * G * psq;
  const Z t6 = b * d * eta1 * cp / det;
  const Z m2i = lift_c<Z>(cd{0.0, -2.0});
  const Z zero = lift<Z>(0.0);
  IfaceZ<Z> f;
  f.rd00 = t1 - t2;
  f.rd01 = zero;
  f.rd11 = zero;
  f.td00 = rtza1 * 2.0 * rtza2 * F;
  f.td01 = zero;
  f.td10 = m2i * rtza1 * rtzb2 * H * cp;
  f.td11 = zero;
  f.ru00 = -(t1 + t5);
  f.ru01 = m2i * rteta2 * rtneta2 * t6;
  f.ru11 = -(t2 + t4);
  return f;
}

template <class Z>
__host__ __device__ __forceinline__ IfaceZ<Z> interface_t(int i, double p, const LayerZ<Z>& up,
                                                          const LayerZ<Z>& down) {
  return (i == 0) ? ocean_bottom_t(p, up, down) : solid_solid_t(p, up, down);
}

// exp(i omega eta h) with the reference's association ((i eta) omega) h.
template <class Z>
__host__ __device__ __forceinline__ Z phase_t(const Z& eta, cd omega, const Z& h) {
  return cexp((mul_i(eta) * lift_c<Z>(omega)) * h);
}

}  // namespace kh
