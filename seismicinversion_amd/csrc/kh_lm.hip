// Device-resident Levenberg-Marquardt iteration of the tau-p inversion
// (Kennett_Reflectivity/taup_inversion.py:389-439), so that an iteration
// needs no host round trip and can be captured whole as one HIP graph.
//
// One iteration, after the misfit / gradient / exact Hessian evaluation of
// the current iterate (the derivative entry points):
//   lm_prepare_kernel   the log-parameter transform of g and H
//                       (taup_inversion.py:180-185), the history row, the
//                       first damping value and the two convergence tests
//                       (:389-420);
//   lm_trials_kernel    every trial of the damping schedule at once: trial t
//                       uses lambda * 5^t (the schedule multiplies by 5 after
//                       each rejection, :433), so the trials are independent
//                       and one workgroup per trial solves (H + lambda_t I)
//                       dx = -g by LU with partial pivoting (np.linalg.solve,
//                       :238-254) and writes the trial's layer table;
//   lm_trial_misfit_kernel  the misfit of every trial model
//                       (_eval_misfit_only, :190-230), one workgroup per
//                       (trial, slowness) row, with the arithmetic and the
//                       summation order of kh_forward + kh_residual_norms;
//   lm_decide_kernel    the first trial whose misfit is below the current
//                       one is accepted (lambda halved, floor 1e-15),
//                       otherwise the line search has failed (:425-439).
// The sequential loop's result follows: it accepts the same first trial
// with the same damping.  Every kernel reads the loop status first and does
// nothing once the loop has stopped, so iterations launched past the end
// change nothing.
//
// Loop state: KH_LM_STATE_LEN float64 values (see kennett_hip.h).
#include "kh_physics.hpp"
#include "kh_runtime.hpp"
#include "kh_wave.hpp"
#include "kh_sweep.hpp"

namespace kh {

enum { LM_LAM = 0, LM_CHI, LM_CHI_PREV, LM_ITER, LM_STATUS, LM_TRIAL };

constexpr int LM_MAX_P = 4 * KH_MAX_LAYERS_DERIV - 5;  // packed parameters at the layer cap
constexpr int LM_BLOCK = 512;
constexpr int LM_LDS_MAX_P = 83;  // P x (P + 1) doubles and the pivot tables within 64 KB of LDS

static_assert(KH_LM_STATE_LEN >= LM_TRIAL + 1, "loop state layout");

__host__ __device__ __forceinline__ int lm_params(int nl) { return 4 * nl - 5; }

// Column c of layer l of the table of the packed physical parameters
// m = exp(x) = [alpha_1..N, beta_1..N, rho_1..N, h_1..N-1], N = nl - 1
// (kennett_torch.py:552-608): the water layer, the Q columns and the
// half-space thickness come from the reference table; thicknesses of
// 1e30 or more mean a half-space (+inf), as the host table builder maps them.
__device__ __forceinline__ double lm_table_entry(const double* __restrict__ ref, int nl,
                                                 const double* __restrict__ x, int l, int c) {
  const int n = nl - 1;
  double v = ref[l * NCOL + c];
  if (l > 0 && c < COL_QA && !(c == COL_H && l == nl - 1)) v = exp(x[c * n + (l - 1)]);
  if (c == COL_H && !(v < 1e30)) v = INFINITY;
  return v;
}

__global__ void lm_tables_kernel(const double* __restrict__ ref, int nl,
                                 const double* __restrict__ x, int nsets,
                                 double* __restrict__ tables, const double* __restrict__ state) {
  if (state != nullptr && state[LM_STATUS] != 0.0) return;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)nl * NCOL;
  if (idx >= per * nsets) return;
  const int s = (int)(idx / per);
  const int r = (int)(idx - (long)s * per);
  tables[idx] = lm_table_entry(ref, nl, x + (size_t)s * lm_params(nl), r / NCOL, r % NCOL);
}

// Parameter k of the packed vector is column c of layer l
// (kennett_torch.py:552-571: all alphas, all betas, all rhos, then the
// thicknesses of layers 1..nl-2).
__device__ __forceinline__ void packed_lc(int k, int n, int& l, int& c) {
  c = k / n;
  l = k - c * n + 1;
}

// src[0] + src[stride] + ... + src[(n - 1) stride], added in that order
// from 0.0, with eight loads in flight at a time.
__device__ __forceinline__ double lm_strided_sum(const double* __restrict__ src, size_t stride,
                                                 int n) {
  double s = 0.0;
  for (int q0 = 0; q0 < n; q0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = q0 + u < n ? src[(size_t)(q0 + u) * stride] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (q0 + u < n) s += v[u];
  }
  return s;
}

// The gradient and Hessian in the packed parameters, summed over the
// slownesses in ascending order, from the per-slowness outputs of
// kh_misfit_hvp_partial over a basis (rows (P, np, nl, 6); direction
// order[k] moves packed parameter k):
//   g[k]    = sum_p grad[p, l_k, c_k]
//   H[a][b] = G[order[a]][order[b]] + r(b, a)  if l_a <= l_b + 1
//           = G[order[a]][order[b]] + r(a, b)  otherwise,
//   r(k, j) = sum_p rows[order[k], p, l_j, c_j]
// (r(k, j) is whole where l_j <= l_k + 1; the other half comes from the
// symmetry of H.)  G (the Gauss-Newton block over the directions) may be
// NULL.
__global__ void lm_assemble_kernel(int nl, int np, const double* __restrict__ rows,
                                   const double* __restrict__ grad, const double* __restrict__ G,
                                   const long long* __restrict__ order, double* __restrict__ g,
                                   double* __restrict__ H) {
  const int n = nl - 1, P = lm_params(nl);
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long PP = (long)P * P;
  const size_t per = (size_t)nl * NCOL;
  if (t < PP) {
    const int a = (int)(t / P), b = (int)(t - (long)(t / P) * P);
    int la, ca, lb, cb;
    packed_lc(a, n, la, ca);
    packed_lc(b, n, lb, cb);
    const bool mirror = la <= lb + 1;
    const int k = mirror ? b : a;
    const int lj = mirror ? la : lb, cj = mirror ? ca : cb;
    const double* src = rows + (size_t)order[k] * np * per + (size_t)lj * NCOL + cj;
    const double s = lm_strided_sum(src, per, np);
    H[t] = G != nullptr ? G[(size_t)order[a] * P + order[b]] + s : s;
  } else if (t < PP + P) {
    const int k = (int)(t - PP);
    int l, c;
    packed_lc(k, n, l, c);
    g[k] = lm_strided_sum(grad + (size_t)l * NCOL + c, per, np);
  }
}

// One workgroup.  g_log = m g, H_log = (m_i m_j) H_ij + delta_ij g_j m_j;
// chi summed over the slowness rows in ascending order; history row
// [chi, |g_log|, |m - m_true| / |m_true|].
__global__ void __launch_bounds__(LM_BLOCK) lm_prepare_kernel(
    int P, int np, const double* __restrict__ chi_rows, int chi_stride,
    const double* __restrict__ g, const double* __restrict__ H, const double* __restrict__ log_x,
    const double* __restrict__ x_true, double tol, int max_hist, double* __restrict__ state,
    double* __restrict__ g_log, double* __restrict__ H_log, double* __restrict__ hist) {
  __shared__ double s_m[LM_MAX_P];
  const bool run = state[LM_STATUS] == 0.0;
  __syncthreads();  // every wave has read the status before thread 0 may change it
  if (!run) return;
  for (int j = threadIdx.x; j < P; j += blockDim.x) {
    const double m = exp(log_x[j]);
    s_m[j] = m;
    g_log[j] = m * g[j];
  }
  __syncthreads();
  for (long idx = threadIdx.x; idx < (long)P * P; idx += blockDim.x) {
    const int i = (int)(idx / P), j = (int)(idx - (long)i * P);
    double v = (s_m[i] * s_m[j]) * H[idx];
    if (i == j) v += g[j] * s_m[j];
    H_log[idx] = v;
  }
  if (threadIdx.x != 0) return;
  const double chi = lm_strided_sum(chi_rows, (size_t)chi_stride, np);
  double gg = 0.0, ee = 0.0, tt = 0.0, dmax = 0.0;
  for (int j = 0; j < P; ++j) {
    const double m = s_m[j];
    const double gl = m * g[j];
    gg += gl * gl;
    const double e = m - x_true[j];
    ee += e * e;
    tt += x_true[j] * x_true[j];
    double d = (m * m) * H[(size_t)j * P + j];
    d += g[j] * m;
    dmax = fmax(dmax, fabs(d));
  }
  const double gnorm = sqrt(gg);
  const double perr = sqrt(ee) / sqrt(tt);
  const int it = (int)state[LM_ITER];
  if (it < max_hist) {
    hist[3 * it + 0] = chi;
    hist[3 * it + 1] = gnorm;
    hist[3 * it + 2] = perr;
  }
  if (state[LM_LAM] < 0.0) state[LM_LAM] = 0.01 * fmax(dmax, 1e-10);
  double status = 0.0;
  if (gnorm < tol) status = 1.0;
  else if (it > 0 && fabs(state[LM_CHI_PREV] - chi) / fmax(fabs(chi), 1e-30) < tol) status = 1.0;
  state[LM_CHI_PREV] = chi;
  state[LM_CHI] = chi;
  state[LM_ITER] = (double)(it + 1);
  state[LM_TRIAL] = -1.0;
  state[LM_STATUS] = status;
}

// Whether trial kernels of a launch starting at trial `first` have work:
// the loop runs, and no earlier trial of this iteration was accepted.
__device__ __forceinline__ bool lm_trials_open(const double* state, int first) {
  return state[LM_STATUS] == 0.0 && (first == 0 || state[LM_TRIAL] < 0.0);
}

// One workgroup per trial t = first + blockIdx.x: A = [H_log + lambda_t I |
// -g_log] (row-major, leading dimension P + 1) in LDS or in the caller's
// workspace, LU with partial pivoting (the first row of largest magnitude
// in the column, as LAPACK's idamax picks it), the right-hand side
// eliminated along, then back substitution.  Rows are exchanged through a
// permutation in LDS rather than moved.  A column step is two phases: the
// first wave searches the pivot, copies the pivot row to LDS and forms the
// multipliers (one reciprocal of the pivot per column, as LAPACK's dgetf2
// scales the column); then every work-item updates its elements of the
// trailing matrix, one read and one write of A each.  Back substitution
// runs in the first wave with the right-hand side in LDS.  Writes trial_x[t]
// = log_x + dx and the trial's layer table.  A zero pivot gives inf/NaN
// entries, whose trial misfit is NaN and is never accepted.
constexpr int LM_LU_BLOCK = 1024;

// The largest |pivot| candidate over the wave and its row: the first row of
// largest magnitude (ties to the lower row, NaN never chosen), the order
// LAPACK's idamax uses, reduced over DPP and permlane exchanges (no LDS).
template <int S>
__device__ __forceinline__ void lm_argmax_step(double& best, int& bi) {
  const long long b = __double_as_longlong(best);
  const unsigned lo = wave_partner_b32<S>((unsigned)b);
  const unsigned hi = wave_partner_b32<S>((unsigned)((unsigned long long)b >> 32));
  const double ob = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
  const int oi = (int)wave_partner_b32<S>((unsigned)bi);
  if (ob > best || (ob == best && oi < bi)) {
    best = ob;
    bi = oi;
  }
}
__device__ __forceinline__ void lm_wave_argmax(double& best, int& bi) {
  lm_argmax_step<0>(best, bi);
  lm_argmax_step<1>(best, bi);
  lm_argmax_step<2>(best, bi);
  lm_argmax_step<3>(best, bi);
  lm_argmax_step<4>(best, bi);
  lm_argmax_step<5>(best, bi);
}

__device__ __forceinline__ void lm_wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

template <bool SHARED>
__global__ void __launch_bounds__(LM_LU_BLOCK) lm_trials_kernel(
    int nl, int first, const double* __restrict__ H, const double* __restrict__ g,
    const double* __restrict__ log_x, const double* __restrict__ ref,
    const double* __restrict__ state, double* __restrict__ trial_x,
    double* __restrict__ tables, double* __restrict__ ws) {
  extern __shared__ double smem[];
  __shared__ int s_perm[LM_MAX_P];
  __shared__ double s_l[LM_MAX_P];        // multipliers of the current column, then x
  __shared__ double s_row[LM_MAX_P + 1];  // the pivot row, then the right-hand side
  if (!lm_trials_open(state, first)) return;  // this kernel never writes the state
  const int P = lm_params(nl);
  const int ld = P + 1;
  const int t = first + (int)blockIdx.x;
  const int tid = threadIdx.x, nthr = blockDim.x;
  double* A = SHARED ? smem : ws + (size_t)blockIdx.x * P * ld;
  double lam = state[LM_LAM];
  for (int k = 0; k < t; ++k) lam *= 5.0;

  for (long idx = tid; idx < (long)P * ld; idx += nthr) {
    const int i = (int)(idx / ld), j = (int)(idx - (long)i * ld);
    double v;
    if (j < P) {
      v = H[(size_t)i * P + j];
      if (i == j) v += lam;
    } else {
      v = -g[i];
    }
    A[idx] = v;
  }
  for (int i = tid; i < P; i += nthr) s_perm[i] = i;
  __syncthreads();

  for (int k = 0; k < P; ++k) {
    if (tid < 64) {
      double best = -1.0;
      int bi = P;
      for (int i = k + tid; i < P; i += 64) {
        const double v = fabs(A[(size_t)s_perm[i] * ld + k]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      lm_wave_argmax(best, bi);
      if (tid == 0 && bi < P && bi != k) {  // an all-NaN column keeps its row
        const int r = s_perm[k];
        s_perm[k] = s_perm[bi];
        s_perm[bi] = r;
      }
      lm_wave_lds_sync();
      const double* rowk = A + (size_t)s_perm[k] * ld;
      const double inv = 1.0 / rowk[k];
      for (int j = k + 1 + tid; j < ld; j += 64) s_row[j] = rowk[j];
      for (int i = k + 1 + tid; i < P; i += 64) s_l[i] = A[(size_t)s_perm[i] * ld + k] * inv;
    }
    __syncthreads();
    // each work-item keeps one column of the trailing matrix and strides
    // over its rows (every element updated once, as before)
    static_assert(LM_LU_BLOCK >= LM_LDS_MAX_P, "one work-item per trailing column at least");
    const int cols = ld - 1 - k, per = nthr / cols;
    const int c = tid % cols, r0 = tid / cols;
    if (r0 < per) {
      const int j = k + 1 + c;
      const double pj = s_row[j];
#pragma unroll 4
      for (int r = k + 1 + r0; r < P; r += per) {
        double* a = A + (size_t)s_perm[r] * ld + j;
        *a = fma(-s_l[r], pj, *a);
      }
    }
    __syncthreads();
  }
  if (tid >= 64) return;  // back substitution and the outputs: the first wave

  for (int i = tid; i < P; i += 64) s_row[i] = A[(size_t)s_perm[i] * ld + P];
  lm_wave_lds_sync();
  for (int i = P - 1; i >= 0; --i) {
    const double xi = s_row[i] / A[(size_t)s_perm[i] * ld + i];
    if (tid == 0) s_l[i] = xi;
    for (int j = tid; j < i; j += 64) s_row[j] = fma(-A[(size_t)s_perm[j] * ld + i], xi, s_row[j]);
    lm_wave_lds_sync();
  }

  double* xt = trial_x + (size_t)t * P;
  for (int j = tid; j < P; j += 64) xt[j] = log_x[j] + s_l[j];
  __threadfence_block();
  lm_wave_lds_sync();
  double* tab = tables + (size_t)t * nl * NCOL;
  for (int r = tid; r < nl * NCOL; r += 64) tab[r] = lm_table_entry(ref, nl, xt, r / NCOL, r % NCOL);
}

// The same solve for systems too large for LDS (A in the caller's
// workspace), blocked: panels of LM_NB columns are factorised in LDS by the
// first wave (pivot search, row exchange, multipliers, update of the
// panel's later columns), then every work-item takes a column to the right
// of the panel (the right-hand side included): it forms that column's
// entries in the panel's pivot rows (forward substitution with the panel's
// multipliers) and updates the rows below with them.  Each element of A
// receives the same fused multiply-adds in the same order as in
// lm_trials_kernel (the same arithmetic per element), while the trailing
// matrix is read and written once per panel instead of once per column.
constexpr int LM_NB = 8;
// panel row stride in LDS: one double of padding, so that the lanes of a
// wave reading one column of consecutive rows hit different banks
constexpr int LM_PS = LM_NB + 1;

__global__ void __launch_bounds__(LM_LU_BLOCK) lm_trials_blocked_kernel(
    int nl, int first, const double* __restrict__ H, const double* __restrict__ g,
    const double* __restrict__ log_x, const double* __restrict__ ref,
    const double* __restrict__ state, double* __restrict__ trial_x,
    double* __restrict__ tables, double* __restrict__ ws) {
  __shared__ int s_perm[LM_MAX_P];
  __shared__ double s_pan[LM_MAX_P * LM_PS];  // panel: logical rows kb.., LM_NB columns
  __shared__ double s_l[LM_MAX_P];            // x
  __shared__ double s_row[LM_MAX_P + 1];      // the right-hand side in back substitution
  if (!lm_trials_open(state, first)) return;  // this kernel never writes the state
  const int P = lm_params(nl);
  const int ld = P + 1;
  const int t = first + (int)blockIdx.x;
  const int tid = threadIdx.x, nthr = blockDim.x;
  double* A = ws + (size_t)blockIdx.x * P * ld;
  double lam = state[LM_LAM];
  for (int k = 0; k < t; ++k) lam *= 5.0;

  for (long idx = tid; idx < (long)P * ld; idx += nthr) {
    const int i = (int)(idx / ld), j = (int)(idx - (long)i * ld);
    double v;
    if (j < P) {
      v = H[(size_t)i * P + j];
      if (i == j) v += lam;
    } else {
      v = -g[i];
    }
    A[idx] = v;
  }
  for (int i = tid; i < P; i += nthr) s_perm[i] = i;
  __syncthreads();

  for (int kb = 0; kb < P; kb += LM_NB) {
    const int nbk = min(LM_NB, P - kb), m = P - kb;
    for (int e = tid; e < m * nbk; e += nthr) {
      const int r = e / nbk, c = e - (e / nbk) * nbk;
      s_pan[r * LM_PS + c] = A[(size_t)s_perm[kb + r] * ld + kb + c];
    }
    __syncthreads();
    if (tid < 64) {  // the panel, column by column
      for (int c = 0; c < nbk; ++c) {
        double best = -1.0;
        int bi = m;
        for (int r = c + tid; r < m; r += 64) {
          const double v = fabs(s_pan[r * LM_PS + c]);
          if (v > best) {
            best = v;
            bi = r;
          }
        }
        lm_wave_argmax(best, bi);
        if (bi < m && bi != c) {  // an all-NaN column keeps its row
          if (tid < nbk) {
            const double a = s_pan[c * LM_PS + tid];
            s_pan[c * LM_PS + tid] = s_pan[bi * LM_PS + tid];
            s_pan[bi * LM_PS + tid] = a;
          }
          if (tid == 0) {
            const int r = s_perm[kb + c];
            s_perm[kb + c] = s_perm[kb + bi];
            s_perm[kb + bi] = r;
          }
        }
        lm_wave_lds_sync();
        const double inv = 1.0 / s_pan[c * LM_PS + c];
        // the pivot row in registers; each row read whole, updated, written
        double prow[LM_NB];
#pragma unroll
        for (int cc = 0; cc < LM_NB; ++cc) prow[cc] = s_pan[c * LM_PS + cc];
        for (int r = c + 1 + tid; r < m; r += 64) {
          double v[LM_NB];
#pragma unroll
          for (int cc = 0; cc < LM_NB; ++cc) v[cc] = s_pan[r * LM_PS + cc];
          const double l = s_pan[r * LM_PS + c] * inv;
#pragma unroll
          for (int cc = 0; cc < LM_NB; ++cc)
            if (cc > c && cc < nbk) s_pan[r * LM_PS + cc] = fma(-l, prow[cc], v[cc]);
          s_pan[r * LM_PS + c] = l;
        }
        lm_wave_lds_sync();
      }
    }
    __syncthreads();
    // the panel's pivot rows (upper part) back to A, for the back substitution
    for (int e = tid; e < nbk * nbk; e += nthr) {
      const int i = e / nbk, c = e - (e / nbk) * nbk;
      if (c >= i) A[(size_t)s_perm[kb + i] * ld + kb + c] = s_pan[i * LM_PS + c];
    }
    // the columns right of the panel: pivot-row entries, then the rows below.
    // G work-items per column (each takes every G-th row below the panel)
    // all read the column's pivot-row entries, so those are written back
    // only after every work-item has read them.  ncol * G <= nthr: one
    // column per work-item at most.
    const int ncol = ld - (kb + nbk);
    const int G = ncol > 0 && nthr / ncol > 0 ? nthr / ncol : 1;
    const bool mine = ncol > 0 && tid < ncol * G;
    const int j = mine ? kb + nbk + tid % ncol : 0, grp = mine ? tid / ncol : 0;
    double u[LM_NB];
    if (mine) {
#pragma unroll
      for (int i = 0; i < LM_NB; ++i) u[i] = i < nbk ? A[(size_t)s_perm[kb + i] * ld + j] : 0.0;
#pragma unroll
      for (int i = 1; i < LM_NB; ++i)
#pragma unroll
        for (int c = 0; c < LM_NB; ++c)
          if (i < nbk && c < i) u[i] = fma(-s_pan[i * LM_PS + c], u[c], u[i]);
#pragma unroll 8
      for (int r = nbk + grp; r < m; r += G) {
        double* a = A + (size_t)s_perm[kb + r] * ld + j;
        double v = *a;
#pragma unroll
        for (int c = 0; c < LM_NB; ++c)
          if (c < nbk) v = fma(-s_pan[r * LM_PS + c], u[c], v);
        *a = v;
      }
    }
    __syncthreads();
    if (mine && grp == 0) {
#pragma unroll
      for (int i = 0; i < LM_NB; ++i)
        if (i < nbk) A[(size_t)s_perm[kb + i] * ld + j] = u[i];
    }
    __syncthreads();
  }
  if (tid >= 64) return;  // back substitution and the outputs: the first wave

  // the right-hand side and U's diagonal in LDS (the panel buffer is free);
  // lane l keeps the rows j = l + 64 k of the column being eliminated in
  // registers and loads the next column while this one is applied (the
  // same operations in the same order as a column-by-column sweep)
  constexpr int KB = (LM_MAX_P + 63) / 64;
  double* s_diag = s_pan;
  for (int i = tid; i < P; i += 64) {
    const double* row = A + (size_t)s_perm[i] * ld;
    s_row[i] = row[P];
    s_diag[i] = row[i];
  }
  const double* rowp[KB];
#pragma unroll
  for (int k = 0; k < KB; ++k) {
    const int j = tid + 64 * k;
    rowp[k] = A + (size_t)s_perm[j < P ? j : 0] * ld;
  }
  double col[KB];
#pragma unroll
  for (int k = 0; k < KB; ++k) {
    const int j = tid + 64 * k;
    col[k] = j < P - 1 ? rowp[k][P - 1] : 0.0;
  }
  lm_wave_lds_sync();
  for (int i = P - 1; i >= 0; --i) {
    double nxt[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int j = tid + 64 * k;
      nxt[k] = j < i - 1 ? rowp[k][i - 1] : 0.0;
    }
    const double xi = s_row[i] / s_diag[i];
    if (tid == 0) s_l[i] = xi;
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int j = tid + 64 * k;
      if (j < i) s_row[j] = fma(-col[k], xi, s_row[j]);
    }
    lm_wave_lds_sync();
#pragma unroll
    for (int k = 0; k < KB; ++k) col[k] = nxt[k];
  }

  double* xt = trial_x + (size_t)t * P;
  for (int j = tid; j < P; j += 64) xt[j] = log_x[j] + s_l[j];
  __threadfence_block();
  lm_wave_lds_sync();
  double* tab = tables + (size_t)t * nl * NCOL;
  for (int r = tid; r < nl * NCOL; r += 64) tab[r] = lm_table_entry(ref, nl, xt, r / NCOL, r % NCOL);
}

// chi[t, p] = sum_w |R_t(p, w) - R_obs(p, w)|^2 for trial table t: the lane
// arithmetic of kh_forward (explicit frequencies) and the summation order of
// kh_residual_norms (work-item j adds frequencies j, j + 256, ... in order,
// then a fixed LDS tree), so each row equals that pair of launches bit for bit.
constexpr int LM_MISFIT_BLOCK = 256;

__global__ void __launch_bounds__(LM_MISFIT_BLOCK) lm_trial_misfit_kernel(
    const double* __restrict__ tables, int nl, int first, const double* __restrict__ p, int np,
    const cd* __restrict__ omega, int nw, int free_surface, const cd* __restrict__ R_obs,
    const double* __restrict__ state, double* __restrict__ chi) {
  extern __shared__ double smem_d[];
  __shared__ double s_red[LM_MISFIT_BLOCK];
  if (!lm_trials_open(state, first)) return;
  LayerPt* s_layer = reinterpret_cast<LayerPt*>(smem_d);
  Iface* s_if = reinterpret_cast<Iface*>(s_layer + nl);
  const int t = first + (int)(blockIdx.x / np);
  const int ip = (int)(blockIdx.x % np);
  stage_slowness(tables + (size_t)t * nl * NCOL, nl, p[ip], s_layer, s_if);
  const cd* obs = R_obs + (size_t)ip * nw;
  double acc = 0.0;
  // frequencies w and w + LM_MISFIT_BLOCK swept together (two independent
  // lane sweeps, each the single-frequency arithmetic), added in order
  int w = threadIdx.x;
  for (; w + LM_MISFIT_BLOCK < nw; w += 2 * LM_MISFIT_BLOCK) {
    const cd om[2] = {omega[w], omega[w + LM_MISFIT_BLOCK]};
    cd out[2];
    sweep_lanes<2, false>(s_layer, s_if, static_cast<const LayerStep*>(nullptr), nl, om,
                          free_surface, out);
    acc = add_sq_diff(acc, out[0], obs[w]);
    acc = add_sq_diff(acc, out[1], obs[w + LM_MISFIT_BLOCK]);
  }
  if (w < nw) {
    const cd om[1] = {omega[w]};
    cd out[1];
    sweep_lanes<1, false>(s_layer, s_if, static_cast<const LayerStep*>(nullptr), nl, om,
                          free_surface, out);
    acc = add_sq_diff(acc, out[0], obs[w]);
  }
  s_red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = LM_MISFIT_BLOCK / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s_red[threadIdx.x] += s_red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) chi[(size_t)t * np + ip] = s_red[0];
}

// One workgroup: among trials first..first+ntrials-1, the first whose
// misfit (rows summed in ascending slowness order) is below the current
// misfit becomes the iterate, with lambda 5^t / 2 (floor 1e-15) from the
// iteration's lambda.  If none is and `last` is set (the schedule ends with
// this batch), lambda 5^(first+ntrials) and status 2; otherwise the state
// is left for the next batch.
__global__ void __launch_bounds__(256) lm_decide_kernel(int P, int first, int ntrials, int last,
                                                        int np,
                                                        const double* __restrict__ chi_trials,
                                                        const double* __restrict__ trial_x,
                                                        double* __restrict__ state,
                                                        double* __restrict__ log_x) {
  __shared__ int s_acc;
  const bool run = lm_trials_open(state, first);
  __syncthreads();  // every wave has read the state before thread 0 may change it
  if (!run) return;
  if (threadIdx.x == 0) {
    const double chi = state[LM_CHI];
    double lam = state[LM_LAM];
    for (int k = 0; k < first; ++k) lam *= 5.0;
    int acc = -1;
    for (int t = first; t < first + ntrials; ++t) {
      const double s = lm_strided_sum(chi_trials + (size_t)t * np, 1, np);
      if (s < chi) {
        acc = t;
        break;
      }
      lam *= 5.0;
    }
    if (acc >= 0) {
      state[LM_LAM] = fmax(lam / 2.0, 1e-15);
      state[LM_TRIAL] = (double)acc;
    } else if (last) {
      state[LM_LAM] = lam;
      state[LM_STATUS] = 2.0;
    }
    s_acc = acc;
  }
  __syncthreads();
  const int acc = s_acc;
  if (acc < 0) return;
  for (int j = threadIdx.x; j < P; j += blockDim.x) log_x[j] = trial_x[(size_t)acc * P + j];
}

static int check_lm_layers(const char* what, int nl) {
  KH_REQUIRE(nl >= 2, KH_ERR_INVALID_ARGUMENT, "%s: need at least 2 layers, got %d", what, nl);
  KH_REQUIRE(nl <= KH_MAX_LAYERS_DERIV, KH_ERR_UNSUPPORTED,
             "%s: %d layers exceeds KH_MAX_LAYERS_DERIV=%d", what, nl, KH_MAX_LAYERS_DERIV);
  return KH_OK;
}

}  // namespace kh

using namespace kh;

extern "C" int kh_lm_tables(const double* ref_layers, int nl, const double* x, int nsets,
                            double* tables, const double* state, void* stream) {
  clear_error();
  KH_TRY(check_lm_layers("kh_lm_tables", nl));
  KH_REQUIRE(nsets >= 0, KH_ERR_INVALID_ARGUMENT, "kh_lm_tables: nsets < 0");
  if (nsets == 0) return KH_OK;
  KH_REQUIRE(ref_layers && x && tables, KH_ERR_INVALID_ARGUMENT, "kh_lm_tables: NULL pointer");
  const long n = (long)nsets * nl * NCOL;
  hipLaunchKernelGGL(lm_tables_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     as_stream(stream), ref_layers, nl, x, nsets, tables, state);
  return check_launch("kh_lm_tables");
}

extern "C" int kh_lm_assemble(int nl, int np, const double* rows, const double* grad,
                              const double* G, const long long* order, double* g, double* H,
                              void* stream) {
  clear_error();
  KH_TRY(check_lm_layers("kh_lm_assemble", nl));
  KH_REQUIRE(np >= 1, KH_ERR_INVALID_ARGUMENT, "kh_lm_assemble: np %d", np);
  KH_REQUIRE(rows && grad && order && g && H, KH_ERR_INVALID_ARGUMENT,
             "kh_lm_assemble: NULL pointer");
  const int P = lm_params(nl);
  const long n = (long)P * P + P;
  hipLaunchKernelGGL(lm_assemble_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     as_stream(stream), nl, np, rows, grad, G, order, g, H);
  return check_launch("kh_lm_assemble");
}

extern "C" int kh_lm_prepare(int nl, int np, const double* chi_rows, int chi_stride,
                             const double* g, const double* H, const double* log_x,
                             const double* x_true, double tol, int max_hist, double* state,
                             double* g_log, double* H_log, double* history, void* stream) {
  clear_error();
  KH_TRY(check_lm_layers("kh_lm_prepare", nl));
  KH_REQUIRE(np >= 1 && chi_stride >= 1 && max_hist >= 0, KH_ERR_INVALID_ARGUMENT,
             "kh_lm_prepare: np %d, chi_stride %d, max_hist %d", np, chi_stride, max_hist);
  KH_REQUIRE(chi_rows && g && H && log_x && x_true && state && g_log && H_log &&
                 (history || max_hist == 0),
             KH_ERR_INVALID_ARGUMENT, "kh_lm_prepare: NULL pointer");
  hipLaunchKernelGGL(lm_prepare_kernel, dim3(1), dim3(LM_BLOCK), 0, as_stream(stream),
                     lm_params(nl), np, chi_rows, chi_stride, g, H, log_x, x_true, tol, max_hist,
                     state, g_log, H_log, history);
  return check_launch("kh_lm_prepare");
}

extern "C" size_t kh_lm_trials_workspace_bytes(int nl, int ntrials) {
  if (nl < 2 || ntrials <= 0) return 0;
  const int P = lm_params(nl);
  if (P <= LM_LDS_MAX_P) return 0;
  return (size_t)ntrials * P * (P + 1) * sizeof(double);
}

extern "C" int kh_lm_trials(int nl, int first, int ntrials, const double* H_log,
                            const double* g_log, const double* log_x, const double* ref_layers,
                            const double* state, double* trial_x, double* trial_tables,
                            void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  KH_TRY(check_lm_layers("kh_lm_trials", nl));
  KH_REQUIRE(first >= 0 && ntrials >= 0, KH_ERR_INVALID_ARGUMENT,
             "kh_lm_trials: first %d, ntrials %d", first, ntrials);
  if (ntrials == 0) return KH_OK;
  KH_REQUIRE(H_log && g_log && log_x && ref_layers && state && trial_x && trial_tables,
             KH_ERR_INVALID_ARGUMENT, "kh_lm_trials: NULL pointer");
  const int P = lm_params(nl);
  hipStream_t st = as_stream(stream);
  if (P <= LM_LDS_MAX_P) {
    const size_t lds = (size_t)P * (P + 1) * sizeof(double);
    hipLaunchKernelGGL(lm_trials_kernel<true>, dim3((unsigned)ntrials), dim3(LM_LU_BLOCK), lds, st,
                       nl, first, H_log, g_log, log_x, ref_layers, state, trial_x, trial_tables,
                       static_cast<double*>(nullptr));
  } else {
    const size_t need = kh_lm_trials_workspace_bytes(nl, ntrials);
    KH_REQUIRE(workspace && workspace_bytes >= need, KH_ERR_INVALID_ARGUMENT,
               "kh_lm_trials: workspace of %zu bytes needed, %zu given", need, workspace_bytes);
    hipLaunchKernelGGL(lm_trials_blocked_kernel, dim3((unsigned)ntrials), dim3(LM_LU_BLOCK), 0, st,
                       nl, first, H_log, g_log, log_x, ref_layers, state, trial_x, trial_tables,
                       static_cast<double*>(workspace));
  }
  return check_launch("kh_lm_trials");
}

extern "C" int kh_lm_trial_misfits(const double* trial_tables, int nl, int first, int ntrials,
                                   const double* p, int np, const double* omega, int nw,
                                   int free_surface, const double* R_obs, const double* state,
                                   double* chi_trials, void* stream) {
  clear_error();
  KH_TRY(check_lm_layers("kh_lm_trial_misfits", nl));
  KH_REQUIRE(first >= 0 && ntrials >= 0 && np >= 0 && nw >= 0, KH_ERR_INVALID_ARGUMENT,
             "kh_lm_trial_misfits: negative size");
  if (ntrials == 0 || np == 0) return KH_OK;
  KH_REQUIRE(trial_tables && p && state && chi_trials && (nw == 0 || (omega && R_obs)),
             KH_ERR_INVALID_ARGUMENT, "kh_lm_trial_misfits: NULL pointer");
  const long nblocks = (long)ntrials * np;
  KH_REQUIRE(nblocks <= 0x7fffffffL, KH_ERR_UNSUPPORTED, "kh_lm_trial_misfits: grid too large");
  const size_t lds = (size_t)nl * sizeof(LayerPt) + (size_t)(nl - 1) * sizeof(Iface);
  hipLaunchKernelGGL(lm_trial_misfit_kernel, dim3((unsigned)nblocks), dim3(LM_MISFIT_BLOCK), lds,
                     as_stream(stream), trial_tables, nl, first, p, np,
                     reinterpret_cast<const cd*>(omega), nw, free_surface,
                     reinterpret_cast<const cd*>(R_obs), state, chi_trials);
  return check_launch("kh_lm_trial_misfits");
}

extern "C" int kh_lm_decide(int nl, int first, int ntrials, int last, int np,
                            const double* chi_trials, const double* trial_x, double* state,
                            double* log_x, void* stream) {
  clear_error();
  KH_TRY(check_lm_layers("kh_lm_decide", nl));
  KH_REQUIRE(first >= 0 && ntrials >= 0 && np >= 1, KH_ERR_INVALID_ARGUMENT,
             "kh_lm_decide: bad size");
  KH_REQUIRE(state && log_x && (ntrials == 0 || (chi_trials && trial_x)), KH_ERR_INVALID_ARGUMENT,
             "kh_lm_decide: NULL pointer");
  hipLaunchKernelGGL(lm_decide_kernel, dim3(1), dim3(256), 0, as_stream(stream), lm_params(nl),
                     first, ntrials, last, np, chi_trials, trial_x, state, log_x);
  return check_launch("kh_lm_decide");
}
