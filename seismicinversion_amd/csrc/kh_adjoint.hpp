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

#include <type_traits>

#include "kh_physics_t.hpp"

namespace kh {

template <class Z>
struct M2 {
  Z a00, a01, a10, a11;
};

// Intermediates of one step (shared by the forward step and its adjoint so
// that the reverse pass recomputes exactly the forward values).  Sums of
// products are fused multiply-add chains, as in kennett_step.
//
// The phases (type EZ) and the coefficients (IfaceZ<FZ>) may be plain cd
// while the state is dual (Z = Dual<cd>): at a level whose layers the
// direction does not move they have no tangent, and the mixed operations
// (kh_dual.hpp) give the full dual step's numbers bit for bit with a third
// fewer products in every coefficient and phase product.
template <class Z, class EZ = Z>
struct StepVals {
  EZ eaea, ebeb, eaeb;
  Z m00, m01, m10, m11, a00, a01, a10, a11, rdet, u00, u01, u10, u11;
  Z x00, x01, x10, x11, y00, y01, y10, y11;
};

// A coefficient as the state's type (zero tangent when lifted).
template <class Z, class X>
__host__ __device__ __forceinline__ Z as_z(const X& x) {
  if constexpr (std::is_same<X, Z>::value) return x;
  else return lift_c<Z>(x);
}

template <class Z, class EZ, class FZ>
__host__ __device__ __forceinline__ StepVals<Z, EZ> step_vals(const M2<Z>& r, const EZ& ea,
                                                              const EZ& eb, const IfaceZ<FZ>& f) {
  StepVals<Z, EZ> v;
  v.eaea = ea * ea;
  v.ebeb = eb * eb;
  v.eaeb = ea * eb;
  v.m00 = v.eaea * r.a00;
  v.m01 = v.eaeb * r.a01;
  v.m10 = v.eaeb * r.a10;
  v.m11 = v.ebeb * r.a11;
  const Z one = lift<Z>(1.0);
  v.a00 = cfms(f.ru01, v.m10, cfms(f.ru00, v.m00, one));
  v.a01 = cfms(f.ru01, v.m11, -(f.ru00 * v.m01));
  v.a10 = cfms(f.ru11, v.m10, -(f.ru01 * v.m00));
  v.a11 = cfms(f.ru11, v.m11, cfms(f.ru01, v.m01, one));
  v.rdet = recip_det(cfms(v.a01, v.a10, v.a00 * v.a11));
  v.u00 = v.a11 * v.rdet;
  v.u01 = -(v.a01 * v.rdet);
  v.u10 = -(v.a10 * v.rdet);
  v.u11 = v.a00 * v.rdet;
  v.x00 = cfma(v.m01, v.u10, v.m00 * v.u00);
  v.x01 = cfma(v.m01, v.u11, v.m00 * v.u01);
  v.x10 = cfma(v.m11, v.u10, v.m10 * v.u00);
  v.x11 = cfma(v.m11, v.u11, v.m10 * v.u01);
  v.y00 = cfma(f.td10, v.x10, f.td00 * v.x00);
  v.y01 = cfma(f.td10, v.x11, f.td00 * v.x01);
  v.y10 = cfma(f.td11, v.x10, f.td01 * v.x00);
  v.y11 = cfma(f.td11, v.x11, f.td01 * v.x01);
  return v;
}

template <class Z, class EZ, class FZ>
__host__ __device__ __forceinline__ M2<Z> step_fwd(const M2<Z>& r, const EZ& ea, const EZ& eb,
                                                   const IfaceZ<FZ>& f) {
  const StepVals<Z, EZ> v = step_vals(r, ea, eb, f);
  M2<Z> s;
  s.a00 = cfma(v.y01, f.td10, cfma(v.y00, f.td00, as_z<Z>(f.rd00)));
  s.a01 = cfma(v.y01, f.td11, cfma(v.y00, f.td01, as_z<Z>(f.rd01)));
  s.a10 = cfma(v.y11, f.td10, cfma(v.y10, f.td00, as_z<Z>(f.rd01)));
  s.a11 = cfma(v.y11, f.td11, cfma(v.y10, f.td01, as_z<Z>(f.rd11)));
  return s;
}

// Adjoint of step_fwd.  Given r = S_{i+1}, the phases, f and sb = S_i_bar:
// returns rb = S_{i+1}_bar, the coefficient cotangents in fb (the Rd and Td
// ones stored, the Ru ones added to fb's) and the phase cotangents eab, ebb.
template <class Z, class EZ, class FZ>
__host__ __device__ __forceinline__ void step_rev(const M2<Z>& r, const EZ& ea, const EZ& eb,
                                                  const IfaceZ<FZ>& f, const M2<Z>& sb,
                                                  M2<Z>& rb, IfaceZ<Z>& fb, Z& eab, Z& ebb) {
  const StepVals<Z, EZ> v = step_vals(r, ea, eb, f);  // recompute the forward intermediates

  // S = Rd + Y Td
  fb.rd00 = sb.a00;
  fb.rd01 = sb.a01 + sb.a10;
  fb.rd11 = sb.a11;
  const Z yb00 = cfma(sb.a01, conj(f.td01), sb.a00 * conj(f.td00));
  const Z yb01 = cfma(sb.a01, conj(f.td11), sb.a00 * conj(f.td10));
  const Z yb10 = cfma(sb.a11, conj(f.td01), sb.a10 * conj(f.td00));
  const Z yb11 = cfma(sb.a11, conj(f.td11), sb.a10 * conj(f.td10));
  // Td_bar from S = Rd + Y Td, then from Y = Tu X (Tu = Td^T: Tu_bar = Yb X^H transposed)
  Z tdb00 = cfma(conj(v.y10), sb.a10, conj(v.y00) * sb.a00);
  Z tdb01 = cfma(conj(v.y10), sb.a11, conj(v.y00) * sb.a01);
  Z tdb10 = cfma(conj(v.y11), sb.a10, conj(v.y01) * sb.a00);
  Z tdb11 = cfma(conj(v.y11), sb.a11, conj(v.y01) * sb.a01);
  tdb00 = cfma(yb01, conj(v.x01), cfma(yb00, conj(v.x00), tdb00));
  tdb10 = cfma(yb01, conj(v.x11), cfma(yb00, conj(v.x10), tdb10));
  tdb01 = cfma(yb11, conj(v.x01), cfma(yb10, conj(v.x00), tdb01));
  tdb11 = cfma(yb11, conj(v.x11), cfma(yb10, conj(v.x10), tdb11));
  const Z xb00 = cfma(conj(f.td01), yb10, conj(f.td00) * yb00);
  const Z xb01 = cfma(conj(f.td01), yb11, conj(f.td00) * yb01);
  const Z xb10 = cfma(conj(f.td11), yb10, conj(f.td10) * yb00);
  const Z xb11 = cfma(conj(f.td11), yb11, conj(f.td10) * yb01);
  fb.td00 = tdb00;
  fb.td01 = tdb01;
  fb.td10 = tdb10;
  fb.td11 = tdb11;
  // X = M U
  Z mb00 = cfma(xb01, conj(v.u01), xb00 * conj(v.u00));
  Z mb01 = cfma(xb01, conj(v.u11), xb00 * conj(v.u10));
  Z mb10 = cfma(xb11, conj(v.u01), xb10 * conj(v.u00));
  Z mb11 = cfma(xb11, conj(v.u11), xb10 * conj(v.u10));
  const Z ub00 = cfma(conj(v.m10), xb10, conj(v.m00) * xb00);
  const Z ub01 = cfma(conj(v.m10), xb11, conj(v.m00) * xb01);
  const Z ub10 = cfma(conj(v.m11), xb10, conj(v.m01) * xb00);
  const Z ub11 = cfma(conj(v.m11), xb11, conj(v.m01) * xb01);
  // U = adj(A) rdet
  const Z crd = conj(v.rdet);
  Z ab00 = crd * ub11, ab11 = crd * ub00;
  Z ab01 = -(crd * ub01), ab10 = -(crd * ub10);
  const Z rdetb = cfma(conj(v.a00), ub11,
                       cfms(conj(v.a10), ub10, cfms(conj(v.a01), ub01, conj(v.a11) * ub00)));
  // rdet = 1/det, det = a00 a11 - a01 a10
  const Z detb = -(conj(v.rdet * v.rdet) * rdetb);
  ab00 = cfma(conj(v.a11), detb, ab00);
  ab11 = cfma(conj(v.a00), detb, ab11);
  ab01 = cfms(conj(v.a10), detb, ab01);
  ab10 = cfms(conj(v.a01), detb, ab10);
  // A = I - Ru M
  fb.ru00 = cfms(ab01, conj(v.m01), cfms(ab00, conj(v.m00), fb.ru00));
  fb.ru01 = cfms(ab11, conj(v.m01), cfms(ab10, conj(v.m00),
                 cfms(ab01, conj(v.m11), cfms(ab00, conj(v.m10), fb.ru01))));
  fb.ru11 = cfms(ab11, conj(v.m11), cfms(ab10, conj(v.m10), fb.ru11));
  mb00 = cfms(conj(f.ru01), ab10, cfms(conj(f.ru00), ab00, mb00));
  mb01 = cfms(conj(f.ru01), ab11, cfms(conj(f.ru00), ab01, mb01));
  mb10 = cfms(conj(f.ru11), ab10, cfms(conj(f.ru01), ab00, mb10));
  mb11 = cfms(conj(f.ru11), ab11, cfms(conj(f.ru01), ab01, mb11));
  // M = E2 o r
  rb.a00 = conj(v.eaea) * mb00;
  rb.a01 = conj(v.eaeb) * mb01;
  rb.a10 = conj(v.eaeb) * mb10;
  rb.a11 = conj(v.ebeb) * mb11;
  const Z eaeab = conj(r.a00) * mb00;
  const Z eaebb = cfma(conj(r.a10), mb10, conj(r.a01) * mb01);
  const Z ebebb = conj(r.a11) * mb11;
  eab = cfma(conj(eb), eaebb, conj(ea * 2.0) * eaeab);
  ebb = cfma(conj(ea), eaebb, conj(eb * 2.0) * ebebb);
}

// Cotangents of eta (complex) and h (real, returned in .r with 0 imaginary
// parts at every dual level) of ea = exp(((i eta) omega) h), given ea_bar.
// eta, h and ea may be plain cd (a layer without tangent) while the
// cotangents are dual.
template <class Z, class EZ>
__host__ __device__ __forceinline__ void phase_rev(const EZ& eta, cd omega, const EZ& h,
                                                   const EZ& ea, const Z& eab, Z& etab, Z& hb) {
  const Z zb = conj(ea) * eab;  // z = ((i eta) omega) h, ea = exp(z)
  const EZ iw = lift_c<EZ>(mul_i(omega));
  etab = etab + conj(iw * h) * zb;
  hb = hb + re_part(conj(mul_i(eta) * lift_c<EZ>(omega)) * zb);
}

}  // namespace kh
