// Per-(layer, slowness) and per-(interface, slowness) physics of the Kennett
// P-SV recursion, evaluated on the device.
//
//   complex slowness ........ Kennett_Reflectivity/layer_model.py:42-45
//   vertical slowness ....... Kennett_Reflectivity/layer_model.py:69-74
//   solid-solid interface ... Kennett_Reflectivity/scattering_matrices.py:82-161
//   ocean-bottom interface .. Kennett_Reflectivity/scattering_matrices.py:209-298
//
// The interface formulas keep the reference's operation order (association
// of products, division by the determinant) so that results agree with the
// NumPy restatement to rounding.  Coefficients are stored compactly: Rd and
// Ru are symmetric (Rd12 = Rd21 and Ru12 = Ru21 are the same expression in
// the reference) and Tu = Td^T holds for both interface types, so an
// interface is 10 complex numbers.
#pragma once

#include "kh_complex.hpp"

namespace kh {

// Row layout of the layer table passed through the C-ABI.
enum { COL_ALPHA = 0, COL_BETA, COL_RHO, COL_H, COL_QA, COL_QB, NCOL };

struct LayerPt {  // one layer at one slowness
  cd eta;        // vertical P slowness, Im > 0
  cd neta;       // vertical S slowness (0 in the water layer)
  cd beta_c;     // complex S velocity 1/s_s (0 for acoustic layers)
  double rho;
  double h;
  int halfspace;  // h == +inf: unit phase factors
  int pad_;
};

struct Iface {  // one interface at one slowness
  cd rd00, rd01, rd11;  // downgoing reflection (symmetric)
  cd ru00, ru01, ru11;  // upgoing reflection (symmetric)
  cd td00, td01, td10, td11;  // downgoing transmission; Tu = Td^T
};

__host__ __device__ __forceinline__ cd complex_slowness(double v, double q) {
  const double two_q = 2.0 * q;
  const double two_q_sq = two_q * two_q;
  const double denom = (1.0 + two_q_sq) * v;
  return cd{two_q_sq / denom, two_q / denom};
}

__host__ __device__ __forceinline__ cd vertical_slowness(cd s, double p) {
  const cd cp{p, 0.0};
  const cd t = (s + cp) * (s - cp);
  cd e = csqrt(t);
  if (e.i <= 0.0) e = -e;
  return e;
}

// Layer `l` of a row-major (nl, 6) table at slowness p.
__host__ __device__ __forceinline__ LayerPt layer_at(const double* row, int l, double p) {
  const double alpha = row[COL_ALPHA], beta = row[COL_BETA];
  LayerPt L;
  const cd sp = complex_slowness(alpha, row[COL_QA]);
  cd ss{0.0, 0.0};
  if (beta > 0.0) ss = complex_slowness(beta, row[COL_QB]);
  L.beta_c = (ss.r != 0.0 || ss.i != 0.0) ? recip(ss) : cd{0.0, 0.0};
  L.eta = vertical_slowness(sp, p);
  L.neta = (l == 0) ? cd{0.0, 0.0} : vertical_slowness(ss, p);
  L.rho = row[COL_RHO];
  L.h = row[COL_H];
  L.halfspace = is_inf(L.h) ? 1 : 0;
  L.pad_ = 0;
  return L;
}

// Solid-solid interface between layers 1 (above) and 2 (below).
__host__ __device__ inline Iface solid_solid(double p, const LayerPt& l1, const LayerPt& l2) {
  const cd eta1 = l1.eta, neta1 = l1.neta, beta1 = l1.beta_c;
  const cd eta2 = l2.eta, neta2 = l2.neta, beta2 = l2.beta_c;
  const double rtrho1 = sqrt(l1.rho), rtrho2 = sqrt(l2.rho);
  const cd rteta1 = csqrt(eta1), rteta2 = csqrt(eta2);
  const cd rtneta1 = csqrt(neta1), rtneta2 = csqrt(neta2);
  const cd rtza1 = rteta1 * rtrho1, rtza2 = rteta2 * rtrho2;
  const cd rtzb1 = rtneta1 * rtrho1, rtzb2 = rtneta2 * rtrho2;
  const cd psq{p * p, 0.0};
  const cd crho1{l1.rho, 0.0}, crho2{l2.rho, 0.0};
  const cd cp{p, 0.0};
  const cd drho = crho2 - crho1;
  const cd dmu = crho2 * beta2 * beta2 - crho1 * beta1 * beta1;
  const cd d = 2.0 * dmu;
  const cd psqd = psq * d;
  const cd a = drho - psqd, b = crho2 - psqd, c = crho1 + psqd;
  cd E = b * eta1 + c * eta2;
  cd F = b * neta1 + c * neta2;
  cd G = a - d * eta1 * neta2;
  cd H = a - d * eta2 * neta1;
  const cd det = E * F + G * H * psq;
  E = E / det;
  F = F / det;
  G = G / det;
  H = H / det;
  const cd q_v = (b * eta1 - c * eta2) * F;
  const cd r_v = (a + d * eta1 * neta2) * H * psq;
  const cd s_v = (a * b + c * d * eta2 * neta2) * cp / det;
  const cd t_v = (b * neta1 - c * neta2) * E;
  const cd u_v = (a + d * eta2 * neta1) * G * psq;
  const cd v_v = (a * c + b * d * eta1 * neta1) * cp / det;
  const cd m2ci{0.0, -2.0};
  const cd two{2.0, 0.0};
  Iface f;
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

// Acoustic (water, layer 1) over elastic (layer 2).
__host__ __device__ inline Iface ocean_bottom(double p, const LayerPt& l1, const LayerPt& l2) {
  const cd eta1 = l1.eta;
  const cd eta2 = l2.eta, neta2 = l2.neta, beta2 = l2.beta_c;
  const double rtrho1 = sqrt(l1.rho), rtrho2 = sqrt(l2.rho);
  const cd rteta1 = csqrt(eta1), rteta2 = csqrt(eta2), rtneta2 = csqrt(neta2);
  const cd rtza1 = rtrho1 * rteta1, rtza2 = rtrho2 * rteta2, rtzb2 = rtrho2 * rtneta2;
  const double psq = p * p;
  const cd crho1{l1.rho, 0.0}, crho2{l2.rho, 0.0};
  const cd cp{p, 0.0};
  const cd drho = crho2 - crho1;
  const cd dmu = crho2 * beta2 * beta2;
  const cd d = 2.0 * dmu;
  const cd psqd = psq * d;
  const cd a = drho - psqd, b = crho2 - psqd, c = crho1 + psqd;
  cd E = b * eta1 + c * eta2;
  cd F = b;
  cd G = a - d * eta1 * neta2;
  cd H = (-d) * eta2;
  const cd det = E * F + G * H * psq;
  E = E / det;
  F = F / det;
  G = G / det;
  H = H / det;
  const cd t1 = (b * eta1 - c * eta2) * F;
  const cd t2 = (a + d * eta1 * neta2) * H * psq;
  const cd t4 = b * E;
  const cd t5 = d * eta2 * G * psq;
  const cd t6 = b * d * eta1 * cp / det;
  const cd m2i{0.0, -2.0};  // 2.0 * (0, -1), exact
  Iface f;
  f.rd00 = t1 - t2;
  f.rd01 = cd{0.0, 0.0};
  f.rd11 = cd{0.0, 0.0};
  f.td00 = 2.0 * rtza1 * rtza2 * F;
  f.td01 = cd{0.0, 0.0};
  f.td10 = m2i * rtza1 * rtzb2 * H * cp;
  f.td11 = cd{0.0, 0.0};
  f.ru00 = -(t1 + t5);
  f.ru01 = m2i * rteta2 * rtneta2 * t6;
  f.ru11 = -(t2 + t4);
  return f;
}

__host__ __device__ __forceinline__ Iface interface_at(int i, double p, const LayerPt& up,
                                                       const LayerPt& down) {
  return (i == 0) ? ocean_bottom(p, up, down) : solid_solid(p, up, down);
}

// Phase factor exp(i * omega * eta * h), with the reference's association
// ((i*eta) * omega) * h (kennett_reflectivity_gpu.py:201-214).
__host__ __device__ __forceinline__ cd phase(cd eta, cd omega, double h) {
  return cexp((mul_i(eta) * omega) * h);
}

}  // namespace kh
