// The Kennett interface step and its hand-written adjoint, over a scalar
// type Z (cd for values and gradients, Dual<cd> for tangents and
// tangent-over-adjoint Hessian-vector products).
//
// Forward step at interface i (layer b = i+1 below it), as in
// Kennett_Reflectivity/kennett_reflectivity_gpu.py:235-257:
//   M = E2 o S_{i+1}  (E2 = [[ea^2, ea eb], [ea eb, eb^2]], elementwise)
//   A = I - Ru M,  U = adj(A) / det(A),  X = M U,  Y = Tu X  (Tu = Td^T)
//   S_i = Rd + Y Td
//
// Reverse step: for complex intermediates the adjoint z_bar follows torch's
// convention (dL/dRe z + i dL/dIm z); every operation is holomorphic, so for
// w = f(z): z_bar += conj(f'(z)) w_bar, and for matrix products C = A B:
// A_bar += C_bar B^H, B_bar += A^H C_bar.  Coefficient cotangents are kept in
// the compact form (Rd01 and Ru01 collect both off-diagonal entries, Tu's
// contributions are folded into Td's).
#pragma once

#include "kh_physics_t.hpp"

namespace kh {

template <class Z>
struct M2 {
  Z a00, a01, a10, a11;
};

template <class Z>
__host__ __device__ __forceinline__ M2<Z> step_fwd(const M2<Z>& r, const Z& ea, const Z& eb,
                                                   const IfaceZ<Z>& f) {
  const Z eaea = ea * ea, ebeb = eb * eb, eaeb = ea * eb;
  const Z m00 = eaea * r.a00, m01 = eaeb * r.a01, m10 = eaeb * r.a10, m11 = ebeb * r.a11;
  const Z one = lift<Z>(1.0);
  const Z a00 = one - (f.ru00 * m00 + f.ru01 * m10);
  const Z a01 = -(f.ru00 * m01 + f.ru01 * m11);
  const Z a10 = -(f.ru01 * m00 + f.ru11 * m10);
  const Z a11 = one - (f.ru01 * m01 + f.ru11 * m11);
  const Z rdet = recip_det(a00 * a11 - a01 * a10);
  const Z u00 = a11 * rdet, u01 = -a01 * rdet, u10 = -a10 * rdet, u11 = a00 * rdet;
  const Z x00 = m00 * u00 + m01 * u10, x01 = m00 * u01 + m01 * u11;
  const Z x10 = m10 * u00 + m11 * u10, x11 = m10 * u01 + m11 * u11;
  const Z y00 = f.td00 * x00 + f.td10 * x10, y01 = f.td00 * x01 + f.td10 * x11;
  const Z y10 = f.td01 * x00 + f.td11 * x10, y11 = f.td01 * x01 + f.td11 * x11;
  M2<Z> s;
  s.a00 = f.rd00 + (y00 * f.td00 + y01 * f.td10);
  s.a01 = f.rd01 + (y00 * f.td01 + y01 * f.td11);
  s.a10 = f.rd01 + (y10 * f.td00 + y11 * f.td10);
  s.a11 = f.rd11 + (y10 * f.td01 + y11 * f.td11);
  return s;
}

// Adjoint of step_fwd.  Given r = S_{i+1}, the phases, f and sb = S_i_bar:
// returns rb = S_{i+1}_bar, adds the coefficient cotangents into fb and
// returns the phase cotangents eab, ebb.
template <class Z>
__host__ __device__ __forceinline__ void step_rev(const M2<Z>& r, const Z& ea, const Z& eb,
                                                  const IfaceZ<Z>& f, const M2<Z>& sb,
                                                  M2<Z>& rb, IfaceZ<Z>& fb, Z& eab, Z& ebb) {
  // recompute the forward intermediates
  const Z eaea = ea * ea, ebeb = eb * eb, eaeb = ea * eb;
  const Z m00 = eaea * r.a00, m01 = eaeb * r.a01, m10 = eaeb * r.a10, m11 = ebeb * r.a11;
  const Z one = lift<Z>(1.0);
  const Z a00 = one - (f.ru00 * m00 + f.ru01 * m10);
  const Z a01 = -(f.ru00 * m01 + f.ru01 * m11);
  const Z a10 = -(f.ru01 * m00 + f.ru11 * m10);
  const Z a11 = one - (f.ru01 * m01 + f.ru11 * m11);
  const Z rdet = recip_det(a00 * a11 - a01 * a10);
  const Z u00 = a11 * rdet, u01 = -a01 * rdet, u10 = -a10 * rdet, u11 = a00 * rdet;
  const Z x00 = m00 * u00 + m01 * u10, x01 = m00 * u01 + m01 * u11;
  const Z x10 = m10 * u00 + m11 * u10, x11 = m10 * u01 + m11 * u11;
  const Z y00 = f.td00 * x00 + f.td10 * x10, y01 = f.td00 * x01 + f.td10 * x11;
  const Z y10 = f.td01 * x00 + f.td11 * x10, y11 = f.td01 * x01 + f.td11 * x11;

  // S = Rd + Y Td
  fb.rd00 = fb.rd00 + sb.a00;
  fb.rd01 = fb.rd01 + (sb.a01 + sb.a10);
  fb.rd11 = fb.rd11 + sb.a11;
  const Z yb00 = sb.a00 * conj(f.td00) + sb.a01 * conj(f.td01);
  const Z yb01 = sb.a00 * conj(f.td10) + sb.a01 * conj(f.td11);
  const Z yb10 = sb.a10 * conj(f.td00) + sb.a11 * conj(f.td01);
  const Z yb11 = sb.a10 * conj(f.td10) + sb.a11 * conj(f.td11);
  Z tdb00 = conj(y00) * sb.a00 + conj(y10) * sb.a10;
  Z tdb01 = conj(y00) * sb.a01 + conj(y10) * sb.a11;
  Z tdb10 = conj(y01) * sb.a00 + conj(y11) * sb.a10;
  Z tdb11 = conj(y01) * sb.a01 + conj(y11) * sb.a11;
  // Y = Tu X, Tu = Td^T: Tu_bar = Yb X^H (transposed into Td_bar), X_bar = Tu^H Yb
  tdb00 = tdb00 + (yb00 * conj(x00) + yb01 * conj(x01));
  tdb10 = tdb10 + (yb00 * conj(x10) + yb01 * conj(x11));
  tdb01 = tdb01 + (yb10 * conj(x00) + yb11 * conj(x01));
  tdb11 = tdb11 + (yb10 * conj(x10) + yb11 * conj(x11));
  const Z xb00 = conj(f.td00) * yb00 + conj(f.td01) * yb10;
  const Z xb01 = conj(f.td00) * yb01 + conj(f.td01) * yb11;
  const Z xb10 = conj(f.td10) * yb00 + conj(f.td11) * yb10;
  const Z xb11 = conj(f.td10) * yb01 + conj(f.td11) * yb11;
  fb.td00 = fb.td00 + tdb00;
  fb.td01 = fb.td01 + tdb01;
  fb.td10 = fb.td10 + tdb10;
  fb.td11 = fb.td11 + tdb11;
  // X = M U
  Z mb00 = xb00 * conj(u00) + xb01 * conj(u01);
  Z mb01 = xb00 * conj(u10) + xb01 * conj(u11);
  Z mb10 = xb10 * conj(u00) + xb11 * conj(u01);
  Z mb11 = xb10 * conj(u10) + xb11 * conj(u11);
  const Z ub00 = conj(m00) * xb00 + conj(m10) * xb10;
  const Z ub01 = conj(m00) * xb01 + conj(m10) * xb11;
  const Z ub10 = conj(m01) * xb00 + conj(m11) * xb10;
  const Z ub11 = conj(m01) * xb01 + conj(m11) * xb11;
  // U = adj(A) rdet
  const Z crd = conj(rdet);
  Z ab00 = crd * ub11, ab11 = crd * ub00;
  Z ab01 = -(crd * ub01), ab10 = -(crd * ub10);
  const Z rdetb = conj(a11) * ub00 - conj(a01) * ub01 - conj(a10) * ub10 + conj(a00) * ub11;
  // rdet = 1/det, det = a00 a11 - a01 a10
  const Z detb = -(conj(rdet * rdet) * rdetb);
  ab00 = ab00 + conj(a11) * detb;
  ab11 = ab11 + conj(a00) * detb;
  ab01 = ab01 - conj(a10) * detb;
  ab10 = ab10 - conj(a01) * detb;
  // A = I - Ru M
  fb.ru00 = fb.ru00 - (ab00 * conj(m00) + ab01 * conj(m01));
  fb.ru01 = fb.ru01 - (ab00 * conj(m10) + ab01 * conj(m11)) - (ab10 * conj(m00) + ab11 * conj(m01));
  fb.ru11 = fb.ru11 - (ab10 * conj(m10) + ab11 * conj(m11));
  mb00 = mb00 - (conj(f.ru00) * ab00 + conj(f.ru01) * ab10);
  mb01 = mb01 - (conj(f.ru00) * ab01 + conj(f.ru01) * ab11);
  mb10 = mb10 - (conj(f.ru01) * ab00 + conj(f.ru11) * ab10);
  mb11 = mb11 - (conj(f.ru01) * ab01 + conj(f.ru11) * ab11);
  // M = E2 o r
  rb.a00 = conj(eaea) * mb00;
  rb.a01 = conj(eaeb) * mb01;
  rb.a10 = conj(eaeb) * mb10;
  rb.a11 = conj(ebeb) * mb11;
  const Z eaeab = conj(r.a00) * mb00;
  const Z eaebb = conj(r.a01) * mb01 + conj(r.a10) * mb10;
  const Z ebebb = conj(r.a11) * mb11;
  eab = conj(ea * 2.0) * eaeab + conj(eb) * eaebb;
  ebb = conj(eb * 2.0) * ebebb + conj(ea) * eaebb;
}

// Cotangents of eta (complex) and h (real, returned in .r with 0 imaginary
// parts at every dual level) of ea = exp(((i eta) omega) h), given ea_bar.
template <class Z>
__host__ __device__ __forceinline__ void phase_rev(const Z& eta, cd omega, const Z& h, const Z& ea,
                                                   const Z& eab, Z& etab, Z& hb) {
  const Z zb = conj(ea) * eab;  // z = ((i eta) omega) h, ea = exp(z)
  const Z iw = lift_c<Z>(mul_i(omega));
  etab = etab + conj(iw * h) * zb;
  hb = hb + re_part(conj(mul_i(eta) * lift_c<Z>(omega)) * zb);
}

}  // namespace kh
