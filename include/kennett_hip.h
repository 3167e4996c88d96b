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

/* Largest layer count the derivative entry points accept. */
#define KH_MAX_LAYERS_DERIV 120

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

/* Coefficients of n independent interfaces from given vertical slownesses:
 * p, rho1, rho2 (n,) float64; eta1, neta1, beta1, eta2, neta2, beta2 (n,)
 * complex128 (beta = complex S velocity 1/s_s); ocean != 0 selects the
 * acoustic-over-elastic form (neta1, beta1 ignored, may be NULL).  Outputs
 * Rd, Ru, Tu, Td (n, 2, 2) complex128.  Replaces solid_solid_interface and
 * ocean_bottom_interface (Kennett_Reflectivity/scattering_matrices.py:42-298). */
int kh_interfaces_direct(int n, int ocean, const double* p, const double* eta1,
                         const double* neta1, const double* rho1, const double* beta1,
                         const double* eta2, const double* neta2, const double* rho2,
                         const double* beta2, double* Rd, double* Ru, double* Tu, double* Td,
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
 * workgroup (64, 128 or 256) and `omega_per_item` frequencies
 * walked by each work-item (1, 2, 3, 4 or 8).  Results are identical to
 * kh_forward's for every shape (every phase factor is evaluated directly). */
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

/* kh_forward_grid with an explicit launch shape (as kh_forward_ex).  The
 * phase recurrence steps by block*dw from each lane's first frequency in a
 * tile, so results are bitwise identical for equal (block, omega_per_item)
 * and agree to rounding (about 1e-14 relative) across shapes. */
int kh_forward_grid_ex(const double* layers, int nl, const double* p, int np,
                       double k0, double dw, double gamma, int nw, int free_surface,
                       double* R, int block, int omega_per_item, void* stream);

/* ---- many models in one persistent launch ---- */

/* One model of a kh_forward_multi batch (96 bytes, all offsets in elements). */
typedef struct {
  long long layer_off; /* first row of this model in layers_all ((rows, 6) float64) */
  long long p_off;     /* first slowness of this model in p_all */
  long long r_off;     /* first complex element of this model's (np, nw) block in R_all */
  long long item_off;  /* first work item of this model (prefix sum of the items per model) */
  double k0, dw, gamma; /* frequencies omega_j = (k0 + j)*dw + i*gamma, j < nw */
  int nl, np, nw;
  int ntiles;          /* kh_multi_tiles(nw, block, omega_per_item, half_tile) */
  int free_surface;
  int half_tile;       /* 0: omega_per_item frequencies per lane; nonzero (omega_per_item
                          >= 2): omega_per_item / 2, for models whose nw fills at most
                          about half a tile (the lanes would otherwise idle on the rest) */
  int tiles_per_item;  /* <= 0: one work item per slowness (all ntiles tiles); else
                          each slowness's tiles in runs of this many, one item each */
  int reserved;        /* zero */
  long long table_off; /* first table record of this model (prefix sum of np * nl) */
} kh_model_desc;

/* Frequency tiles of a model's row: ceil(nw / (block * f)) with f the
 * frequencies per lane (omega_per_item, halved with half_tile); 0 if any
 * argument is <= 0. */
int kh_multi_tiles(int nw, int block, int omega_per_item, int half_tile);

/* Work items of one model: np * ceil(ntiles / tiles_per_item) (np when
 * tiles_per_item <= 0); 0 if np or ntiles is <= 0. */
long long kh_multi_items(int np, int ntiles, int tiles_per_item);

/* Bytes of the table workspace kh_forward_multi needs for `records` table
 * records (the sum of np * nl over the batch's models). */
size_t kh_multi_table_bytes(long long records);

/* Workgroups a kh_forward_multi launch with workgroups <= 0 starts on the
 * current device (CUs x resident workgroups, before the cap at total_items)
 * on the staged-table path (tables != 0) or the in-LDS path (whose LDS grows
 * with max_nl), written to *workgroups; for sizing tiles_per_item. */
int kh_multi_workgroups(int block, int omega_per_item, int max_nl, int tables, int* workgroups);

/* Reflectivity of nmodels models in one persistent launch: workgroups take
 * (model, slowness, run of tiles) items from a device counter until all
 * total_items are done.  With `tables` (caller-owned device memory of at
 * least kh_multi_table_bytes(records) bytes, aligned to 256 bytes, one
 * 256-byte record per (model, slowness, layer); records = the sum of np * nl,
 * each model's records starting at its table_off), every (model, slowness,
 * layer) record of the layer, its phase step and the interface below it is
 * computed there first and the sweep reads it with scalar loads; with
 * tables = NULL (records and table_bytes ignored) each item's workgroup
 * stages its slowness's tables in LDS, and the batch is a single kernel.
 * desc (nmodels,) is a DEVICE array sorted by item_off and by table_off
 * (list the models with the costliest items, tiles per item * (nl - 1),
 * first for the best drain order); max_nl = the largest nl;
 * counter = 8 bytes of device memory (reset by the call); nonfinite = NULL
 * or a device int array (nmodels,) that receives, for each desc entry k,
 * the number of NaN/Inf outputs of that model (reset by the call);
 * workgroups <= 0 picks CUs x resident workgroups.  block in {64, 128, 256},
 * omega_per_item in {1, 2, 4, 8}.  Each model's output equals kh_forward_grid_ex
 * of that model alone with the same block and its frequencies per lane
 * (omega_per_item, or omega_per_item / 2 with half_tile), bit for bit.
 * Replaces a loop of kennett_reflectivity_batch calls, one per model
 * (Kennett_Reflectivity/kennett_reflectivity_gpu.py:88). */
int kh_forward_multi(const kh_model_desc* desc, int nmodels, long long total_items, int max_nl,
                     const double* layers_all, const double* p_all, double* R_all, int block,
                     int omega_per_item, int workgroups, long long records, void* tables,
                     size_t table_bytes, void* counter, int* nonfinite, void* stream);

/* Scratch that kh_forward_multi_static keeps in the library, one copy per
 * device, for the life of the process: a 256-byte work counter and up to
 * KH_STATIC_MAX_RECORDS 256-byte table records (32 MiB in all). */
#define KH_STATIC_SCRATCH_BYTES (32u << 20)
#define KH_STATIC_COUNTER_BYTES 256u
#define KH_STATIC_MAX_RECORDS ((KH_STATIC_SCRATCH_BYTES - KH_STATIC_COUNTER_BYTES) / 256u)

/* kh_forward_multi that allocates nothing and takes no workspace: the work
 * counter and, with use_tables != 0, the table records live in the
 * library's own device scratch (KH_STATIC_SCRATCH_BYTES per device).  With
 * use_tables != 0 the staged-table path runs and records (the sum of
 * np * nl) must not exceed KH_STATIC_MAX_RECORDS (KH_ERR_INVALID_ARGUMENT
 * otherwise, before any launch); with use_tables = 0 the in-LDS path runs,
 * for any batch.  Other arguments and the output are kh_forward_multi's
 * (same kernels, bit for bit).  The first call on a device looks up the
 * scratch address (make it outside stream capture; a call with nmodels = 0
 * does only that).  Every call on a device shares that device's scratch:
 * calls that overlap on one device (concurrent streams or threads) race.
 * The counter is reset by the call with a kernel, not a memset.
 * No counterpart in the reference: this entry point is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). */
int kh_forward_multi_static(const kh_model_desc* desc, int nmodels, long long total_items,
                            int max_nl, const double* layers_all, const double* p_all,
                            double* R_all, int block, int omega_per_item, int workgroups,
                            long long records, int use_tables, int* nonfinite, void* stream);

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
 * nw a power of two with nw*16 bytes <= the device's LDS per workgroup (the
 * Hermitian input is transformed as one complex FFT of length nt/2).
 * Replaces the per-trace np.fft.fft loop of
 * Kennett_Reflectivity/kennett_gather_gpu.py:273-282 (and kennett_gather.py:295-303). */
int kh_hermitian_traces(const double* U, int nr, int nw, double gamma, double T,
                        double* traces, void* stream);

/* ---- derivatives ----
 * Parameters are the six columns of every layer row; derivative tables are
 * row-major (nl, 6) per slowness: d/d[alpha, beta, rho, h, Q_alpha, Q_beta].
 * Directions `dlayers` are (ndir, nl, 6) float64.  Cotangents follow torch's
 * convention for complex outputs of real inputs:
 *     grad_x = Re( sum_w conj(G_w) dR_w/dx ).
 * Thickness derivatives of half-space layers (h = inf) are 0.  nl must not
 * exceed KH_MAX_LAYERS_DERIV.  The adjoint entry points keep the whole
 * forward tape (every interface's state) in the workspace when it is large
 * enough, and otherwise the states every ceil(sqrt(nl-1)) interfaces,
 * recomputing the rest (the same numbers to rounding); per-frequency
 * cotangents are reduced in a fixed order. */

/* Workspace bytes for kh_vjp / kh_misfit_grad (dual = 0) or kh_misfit_hvp /
 * kh_vjp_hvp (dual = 1) with ndir directions, keeping the whole forward
 * tape (the faster layout). */
size_t kh_deriv_workspace_bytes(int nl, int np, int nw, int ndir, int dual);

/* The smallest workspace the same calls accept: sqrt(N) checkpoints (or the
 * tape, where that is smaller).  A workspace of at least
 * kh_deriv_workspace_bytes gets the tape, a smaller one of at least this
 * size the checkpoints.  Replaces, with kh_deriv_workspace_bytes, the
 * memory the reference's torch autograd keeps for the backward of
 * Kennett_Reflectivity/kennett_torch.py:312-480. */
size_t kh_deriv_workspace_bytes_min(int nl, int np, int nw, int ndir, int dual);

/* Tangent-linear (forward-mode) products dR[d] = J dlayers[d], (ndir, np, nw)
 * complex128; R (np, nw) is written too when non-NULL.  One launch gives the
 * analytical Frechet derivatives along ndir directions
 * (Kennett_Reflectivity/frechet_analytical.py:583-911, one column per direction). */
int kh_jvp(const double* layers, const double* dlayers, int ndir, int nl,
           const double* p, int np, const double* omega, int nw, int free_surface,
           double* R, double* dR, void* stream);

/* Vector-Jacobian product grad (np, nl, 6) for the cotangent G (np, nw)
 * complex128 (the backward of kennett_reflectivity_torch,
 * Kennett_Reflectivity/kennett_torch.py:312-480, under torch.autograd). */
int kh_vjp(const double* layers, int nl, const double* p, int np,
           const double* omega, int nw, int free_surface, const double* G,
           double* grad, double* R, void* workspace, size_t workspace_bytes,
           void* stream);

/* Misfit chi_p = sum_w |R - R_obs|^2 per slowness (chi (np, 2): value, 0) and
 * its gradient (np, nl, 6) (the per-p terms of taup_misfit_grad_hessian,
 * Kennett_Reflectivity/taup_inversion.py:154-174). */
int kh_misfit_grad(const double* layers, int nl, const double* p, int np,
                   const double* omega, int nw, int free_surface, const double* R_obs,
                   double* chi, double* grad, double* R, void* workspace,
                   size_t workspace_bytes, void* stream);

/* Tangent-over-adjoint Hessian-vector products of the misfit: for every
 * direction d, hvp[d] (np, nl, 6) = H dlayers[d] per slowness (exact Hessian,
 * as torch.func.hessian in taup_inversion.py:177), chi (ndir, np, 2) =
 * [chi, dchi], grad (np, nl, 6) (optional), dR (ndir, np, nw) = J dlayers[d]
 * (optional).  second_only != 0 drops the Gauss-Newton part 2 Re(J^H J dm),
 * leaving 2 Re(sum conj(R - R_obs) d^2R[dm]). */
int kh_misfit_hvp(const double* layers, const double* dlayers, int ndir, int nl,
                  const double* p, int np, const double* omega, int nw,
                  int free_surface, const double* R_obs, int second_only, double* chi,
                  double* grad, double* hvp, double* dR, void* workspace,
                  size_t workspace_bytes, void* stream);

/* kh_misfit_hvp computing, for each direction d, only the entries of the
 * layers 0 .. l_d + 1 of hvp[d], with l_d the deepest layer in which
 * dlayers[d] is nonzero; the entries of deeper layers are written as 0.
 * When grad is given, direction 0 (whose sweep also yields grad) is
 * complete.
 * The reverse sweep stops at interface l_d + 1, so for the basis direction
 * of a parameter of layer l the launch does the sweep work of layers 0..l+1
 * only.  Since the Hessian is symmetric, a caller passing the basis
 * direction of every parameter assembles the whole matrix: entry (j, k)
 * with layer(j) > layer(k) + 1 is entry (k, j) of direction j.  chi, grad
 * and dR are as kh_misfit_hvp's (complete).  Replaces, with that assembly,
 * the torch.func.hessian of the misfit (Kennett_Reflectivity/taup_inversion.py:177). */
int kh_misfit_hvp_partial(const double* layers, const double* dlayers, int ndir, int nl,
                          const double* p, int np, const double* omega, int nw,
                          int free_surface, const double* R_obs, int second_only, double* chi,
                          double* grad, double* hvp, double* dR, void* workspace,
                          size_t workspace_bytes, void* stream);

/* Directional derivative of the vector-Jacobian product: for cotangent G
 * (np, nw) with tangents dG (ndir, np, nw) (or NULL = 0) and parameter
 * directions dlayers, dgrad[d] = d/dt grad(m + t dm_d, G + t dG_d)
 * (ndir, np, nl, 6); grad (np, nl, 6) and dR (ndir, np, nw) optional.
 * This is the forward-mode derivative of kennett_reflectivity_torch's
 * backward, needed by forward-over-reverse Hessians. */
int kh_vjp_hvp(const double* layers, const double* dlayers, int ndir, int nl,
               const double* p, int np, const double* omega, int nw, int free_surface,
               const double* G, const double* dG, double* grad, double* dgrad,
               double* dR, void* workspace, size_t workspace_bytes, void* stream);

/* ---- Levenberg-Marquardt support ---- */

/* Workspace bytes of kh_gram for a (P, K) input. */
size_t kh_gram_workspace_bytes(int P, long K);

/* G (P, P) = scale * X X^T for X (P, K) float64 row-major, on the FP64
 * matrix cores; chunks of K reduced in a fixed order.  With X the real view
 * of Jacobian columns this is the Gauss-Newton block Re(J^H J) of the
 * Hessian that Kennett_Reflectivity/taup_inversion.py:177 forms through
 * torch.func.hessian. */
int kh_gram(const double* X, int P, long K, double scale, double* G, void* workspace,
            size_t workspace_bytes, void* stream);

/* ---- Global Matrix Method ---- */

/* PP reflectivity R (np, nw) complex128 by the Global Matrix Method's
 * block-Riccati sweep: per (p, omega) lane, a 4x4 complex system with two
 * right-hand sides per finite layer (LU with partial pivoting), then the
 * ocean-bottom 2x2 system; arguments as kh_forward.  An independent
 * formulation of the same boundary-value problem as the Kennett recursion
 * (they agree to about 1e-10).  Replaces gmm_reflectivity(..., solver="riccati")
 * (GlobalMatrix/global_matrix.py:297-348, riccati_solver.py:25-194). */
int kh_gmm_forward(const double* layers, int nl, const double* p, int np,
                   const double* omega, int nw, int free_surface, double* R, void* stream);

/* kh_gmm_forward on the uniform grid omega_j = (k0 + j)*dw + i*gamma. */
int kh_gmm_forward_grid(const double* layers, int nl, const double* p, int np, double k0,
                        double dw, double gamma, int nw, int free_surface, double* R,
                        void* stream);

/* Misfit per slowness chi (np,) float64: chi[j] = sum_w |R[j, w] - R_obs[j, w]|^2
 * for R, R_obs (np, nw) complex128, in a fixed summation order.  The trial
 * misfit of the LM damping loop (_eval_misfit_only,
 * Kennett_Reflectivity/taup_inversion.py:190-216) after a kh_forward of the
 * trial model. */
int kh_residual_norms(const double* R, const double* R_obs, int np, int nw, double* chi,
                      void* stream);

/* ---- Device-resident Levenberg-Marquardt iteration ----
 *
 * The host loop of Kennett_Reflectivity/taup_inversion.py:389-439 as four
 * kernels that follow the misfit / gradient / Hessian evaluation of the
 * current iterate, so an iteration needs no host round trip (and can be
 * captured as one HIP graph).  Packed log-parameters x (P = 4 nl - 5):
 * m = exp(x) = [alpha_1..N, beta_1..N, rho_1..N, h_1..N-1], N = nl - 1
 * (kennett_torch.py:552-608).
 *
 * Loop state: KH_LM_STATE_LEN float64 values in device memory,
 *   [0] lambda (set it < 0 before the first iteration: kh_lm_prepare then
 *       sets 0.01 max|diag H_log|, floor 1e-10, taup_inversion.py:406-407)
 *   [1] misfit at the current iterate   [2] misfit at the previous iterate
 *   [3] iterations evaluated            [4] status: 0 running, 1 converged
 *       (gradient or relative-misfit test), 2 line search failed
 *   [5] trial accepted in this iteration (-1: none yet)   [6..7] unused.
 * Every kernel returns at once when the status is not 0, so iterations
 * enqueued after the loop has stopped change nothing. */
#define KH_LM_STATE_LEN 8

/* tables (nsets, nl, 6) of packed log-parameters x (nsets, P): columns 0-3
 * of layers 1..nl-1 from exp(x) (the half-space keeps its thickness), the
 * water layer and the Q columns from ref_layers; thicknesses >= 1e30 become
 * +inf.  `state` may be NULL (always run).  Replaces np.exp and
 * _unpack_params (taup_inversion.py:88-95, 147, 203-215). */
int kh_lm_tables(const double* ref_layers, int nl, const double* x, int nsets, double* tables,
                 const double* state, void* stream);

/* The packed gradient g (P,) and Hessian H (P, P) summed over slownesses
 * (p ascending) from kh_misfit_hvp_partial over the P basis directions:
 * rows (P, np, nl, 6) are its hvp output, grad (np, nl, 6) its gradient,
 * and direction order[k] (int64) moves packed parameter k.  Where a row is
 * partial the entry comes from its mirror (H is symmetric).  G (P, P), the
 * Gauss-Newton block 2 Re(J^H J) over the directions (kh_gram of their
 * dR), is added when not NULL.  Replaces the per-slowness accumulation of
 * Kennett_Reflectivity/taup_inversion.py:163-178. */
int kh_lm_assemble(int nl, int np, const double* rows, const double* grad, const double* G,
                   const long long* order, double* g, double* H, void* stream);

/* One workgroup: chi = sum of chi_rows[p * chi_stride] over p ascending;
 * g_log = m g, H_log = (m_i m_j) H_ij + delta_ij g_j m_j (P x P row-major);
 * history[iter] = [chi, |g_log|, |m - x_true| / |x_true|] while
 * iter < max_hist; first lambda; status 1 when |g_log| < tol or, after the
 * first iteration, |chi_prev - chi| / max(|chi|, 1e-30) < tol.  Replaces
 * taup_inversion.py:180-185 and 389-420. */
int kh_lm_prepare(int nl, int np, const double* chi_rows, int chi_stride, const double* g,
                  const double* H, const double* log_x, const double* x_true, double tol,
                  int max_hist, double* state, double* g_log, double* H_log, double* history,
                  void* stream);

/* Workspace bytes of kh_lm_trials for a launch of ntrials trials (0 when
 * P <= 83: the systems fit in LDS). */
size_t kh_lm_trials_workspace_bytes(int nl, int ntrials);

/* Trials t = first..first+ntrials-1 of the damping schedule at once, one
 * workgroup each: (H_log + lambda 5^t I) dx = -g_log by LU with partial
 * pivoting, trial_x[t] = log_x + dx and trial_tables[t] as kh_lm_tables
 * (trial_x (T, P) and trial_tables (T, nl, 6) hold all T trials of the
 * schedule).  With first > 0 nothing runs once a trial of this iteration
 * has been accepted, so a caller can launch trial 0 alone and the rest
 * after it.  A singular system yields non-finite entries (a trial that is
 * never accepted) where np.linalg.solve raises.  Replaces newton_lm_step
 * (taup_inversion.py:238-254) inside the trial loop (:425-435). */
int kh_lm_trials(int nl, int first, int ntrials, const double* H_log, const double* g_log,
                 const double* log_x, const double* ref_layers, const double* state,
                 double* trial_x, double* trial_tables, void* workspace, size_t workspace_bytes,
                 void* stream);

/* chi_trials (T, np), rows first..first+ntrials-1: per trial and slowness
 * sum_w |R - R_obs|^2 of the trial table, equal bit for bit to kh_forward
 * followed by kh_residual_norms.  R_obs (np, nw) complex128, omega (nw,)
 * complex128; skipped as kh_lm_trials.  Replaces _eval_misfit_only
 * (taup_inversion.py:190-230) for those trials. */
int kh_lm_trial_misfits(const double* trial_tables, int nl, int first, int ntrials,
                        const double* p, int np, const double* omega, int nw, int free_surface,
                        const double* R_obs, const double* state, double* chi_trials,
                        void* stream);

/* Accepts the first of trials first..first+ntrials-1 whose misfit (rows
 * summed over p ascending) is below the current one: log_x = trial_x[t],
 * lambda = max(lambda 5^t / 2, 1e-15).  If none is and `last` is nonzero
 * (the schedule ends with these trials): lambda 5^(first+ntrials) and
 * status 2.  Skipped as kh_lm_trials.  Replaces taup_inversion.py:425-439. */
int kh_lm_decide(int nl, int first, int ntrials, int last, int np, const double* chi_trials,
                 const double* trial_x, double* state, double* log_x, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* KENNETT_HIP_H */
