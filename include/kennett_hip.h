/*
 * kennett_hip.h — C ABI of the MI355X Kennett reflectivity library
 * (libkennett_hip.so, built for gfx950).
 *
 * Conventions shared by every entry point:
 *  - All array arguments are DEVICE pointers owned by the caller (for example
 *    torch tensors' data_ptr()).  Nothing is allocated inside a call unless
 *    the entry point says so.
 *  - Complex numbers are interleaved (re, im) float64 pairs (complex128).
 *  - `layers` is a row-major (nl, 6) float64 table, one row per layer:
 *        [alpha, beta, rho, h, Q_alpha, Q_beta]
 *    layer 0 is the acoustic water layer, h = +inf marks a half-space
 *    (unit phase factor), nl >= 2.
 *  - `stream` is a hipStream_t (NULL = the legacy default stream).  Calls only
 *    enqueue work; they return before it completes.
 *  - Return value: 0 on success, otherwise a KH_ERR_* code; kh_last_error()
 *    then describes the failure (per calling thread).
 *  - Lanes are independent: results do not depend on launch shape.
 */
#ifndef KENNETT_HIP_H
#define KENNETT_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KH_OK 0
#define KH_ERR_INVALID_ARGUMENT 1
#define KH_ERR_OUT_OF_MEMORY 2
#define KH_ERR_LAUNCH 3
#define KH_ERR_UNSUPPORTED 4

/* Largest layer count one launch accepts (the per-slowness layer and
 * interface tables are staged in LDS). */
#define KH_MAX_LAYERS 240

/* Library version string. */
const char* kh_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* kh_last_error(void);

/* Vertical P and S slownesses of every layer at every slowness:
 * eta, neta are (nl, np) complex128, row-major.  neta of layer 0 is 0.
 * Replaces the per-layer, per-p loop of
 * Kennett_Reflectivity/kennett_reflectivity_gpu.py:139-147
 * (vertical_slowness, layer_model.py:48-74). */
int kh_vertical_slowness(const double* layers, int nl, const double* p, int np,
                         double* eta, double* neta, void* stream);

/* Interface coefficients Rd, Ru, Tu, Td at every interface and slowness,
 * each (nl-1, np, 2, 2) complex128, row-major.  Interface 0 is the ocean
 * bottom.  Replaces the per-(p, interface) loop of
 * Kennett_Reflectivity/kennett_reflectivity_gpu.py:150-191
 * (solid_solid_interface / ocean_bottom_interface,
 *  scattering_matrices.py:42-298). */
int kh_interface_coefficients(const double* layers, int nl, const double* p, int np,
                              double* Rd, double* Ru, double* Tu, double* Td,
                              void* stream);

/* PP reflectivity R (np, nw) complex128 (row-major, slowness-major) for
 * slownesses p (np,) float64 and frequencies omega (nw,) complex128 (may be
 * damped, omega + i*gamma; must not be 0).  free_surface != 0 applies the
 * water-column reverberation R/(1+R).
 * Replaces kennett_reflectivity_batch
 * (Kennett_Reflectivity/kennett_reflectivity_gpu.py:88-275) and, at np = 1,
 * kennett_reflectivity (Kennett_Reflectivity/kennett_reflectivity.py:73-328). */
int kh_forward(const double* layers, int nl, const double* p, int np,
               const double* omega, int nw, int free_surface, double* R,
               void* stream);

/* kh_forward with an explicit launch shape: `block` work-items per
 * workgroup (64, 128, 256, 512 or 1024) and `omega_per_item` frequencies
 * walked by each work-item (1, 2, 3, 4 or 8).  Results are identical to
 * kh_forward's for every shape. */
int kh_forward_ex(const double* layers, int nl, const double* p, int np,
                  const double* omega, int nw, int free_surface, double* R,
                  int block, int omega_per_item, void* stream);

/* kh_forward on the uniform frequency grid omega_j = (k0 + j)*dw + i*gamma,
 * j = 0..nw-1 — the grids the reference builds:
 * omega = arange(1, nw)*dw + 1j*gamma (Kennett_Reflectivity/kennett_gather_gpu.py:226-231)
 * and arange(1, nfreq+1)*dw (Kennett_Reflectivity/taup_inversion.py:298-301).
 * No frequency array is read; phase factors of the frequencies a work-item
 * carries beyond its first come from exp(i(w + s)eta h) = exp(i w eta h) exp(i s eta h). */
int kh_forward_grid(const double* layers, int nl, const double* p, int np,
                    double k0, double dw, double gamma, int nw, int free_surface,
                    double* R, void* stream);

/* kh_forward_grid with an explicit launch shape (as kh_forward_ex). */
int kh_forward_grid_ex(const double* layers, int nl, const double* p, int np,
                       double k0, double dw, double gamma, int nw, int free_surface,
                       double* R, int block, int omega_per_item, void* stream);

/* ---- gather: discrete-wavenumber sum and inverse FFT to traces ---- */

/* Workspace bytes kh_bessel_sum needs for np slownesses, nw - 1 frequencies
 * and nr offsets (0 if any is empty). */
size_t kh_bessel_workspace_bytes(int np, int nw, int nr);

/* U[r, k] = S_k * sum_j R[j, k] * taper_j * p_j * w_k * dp * J0(w_k p_j r) in
 * float64, for k = 0..nw-2 with real w_k = (k0 + k)*dw, taper_j =
 * (1 + cos(pi p_j / p_max))/2.  R (np, nw-1) complex128, p (np,), offsets
 * (nr,), S (nw-1,) complex128 or NULL (S = 1), U (nr, nw-1) complex128.
 * The slowness sum is reduced in a fixed order (bitwise reproducible).
 * Replaces _gpu_bessel_summation (Kennett_Reflectivity/kennett_gather_gpu.py:56-166,
 * float32 there) and the per-p accumulation of compute_gather
 * (Kennett_Reflectivity/kennett_gather.py:100-124, 271-293). */
int kh_bessel_sum(const double* R, int np, int nw, const double* p, double dp,
                  double p_max, double k0, double dw, const double* offsets, int nr,
                  const double* S, double* U, void* workspace, size_t workspace_bytes,
                  void* stream);

/* Real traces (nr, nt = 2*nw) float64 from U (nr, nw-1) complex128:
 * trace_r[n] = Re(FFT_nt([0, U_r, 0, conj(reversed U_r)])[n]) * exp(-gamma n T/nt),
 * nw a power of two with nt*16 bytes <= the device's LDS per workgroup.
 * Replaces the per-trace np.fft.fft loop of
 * Kennett_Reflectivity/kennett_gather_gpu.py:273-282 (and kennett_gather.py:295-303). */
int kh_hermitian_traces(const double* U, int nr, int nw, double gamma, double T,
                        double* traces, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* KENNETT_HIP_H */
