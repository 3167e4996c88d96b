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

#ifdef __cplusplus
}
#endif

#endif /* KENNETT_HIP_H */
