// Global Matrix Method reflectivity by the block-Riccati sweep, per lane
// (GlobalMatrix/riccati_solver.py:25-194, global_matrix.py:24-61, 279-348).
//
// For a solid layer at slowness p the displacement-stress eigenvector
// matrices (rows u_x, u_z, sigma_zz/(-i w), sigma_xz/(-i w); columns P, S;
// layer_matrix.py:44-68) are
//     E_d = [[p, nu], [eta, -p], [-g, 2 mu p nu], [2 mu p eta, g]],
//     E_u = E_d with the sign of entry (r, c) flipped when r + c is odd,
// mu = rho beta~^2, g = rho (1 - 2 beta~^2 p^2).  The ocean has
// e_d = (eta0, -rho0), e_u = (-eta0, -rho0).
//
// Upward sweep over the finite layers k = M..1 (M = nl - 2), Y = 0 below:
//     E_eff = E_d[k+1] + E_u[k+1] diag(ph[k+1]) Y      (E_d[k+1] for the half-space)
//     [E_u[k] | -E_eff] X = -E_d[k] diag(ph[k]),  Y = X[0:2, :]
// then at the ocean bottom, with E1 = E_d[1] + E_u[1] diag(ph[1]) Y:
//     [e_u0 E1[2,:] - e_u1 E1[1,:]; E1[3,:]] D = [(e_u0 e_d1 - e_u1 e_d0) e0, 0]
//     U0 = (E1[1,:] D - e_d0 e0) / e_u0,   R = e0 U0   (R / (1 + R) with the free surface)
// ph[k] = (exp(i w eta_k h_k), exp(i w nu_k h_k)), e0 = exp(i w eta_0 h_0).
// Linear systems are solved by LU with partial pivoting on |re| + |im| (the
// pivot rule of LAPACK's zgetrf, which numpy.linalg.solve calls); row swaps
// are conditional moves, so no register array is indexed at run time.
#pragma once

#include "kh_physics.hpp"

namespace kh {

struct EigD {  // E_d of one solid layer at one slowness (E_u follows by signs)
  cd e[4][2];
};

__host__ __device__ __forceinline__ EigD eig_d(double p, const LayerPt& L) {
  const cd bc2 = L.beta_c * L.beta_c;
  const cd mu = bc2 * L.rho;
  const cd g = (cd{1.0, 0.0} - (bc2 * 2.0) * (p * p)) * L.rho;
  const cd tmp = mu * (2.0 * p);
  EigD E;
  E.e[0][0] = cd{p, 0.0};
  E.e[0][1] = L.neta;
  E.e[1][0] = L.eta;
  E.e[1][1] = cd{-p, 0.0};
  E.e[2][0] = -g;
  E.e[2][1] = tmp * L.neta;
  E.e[3][0] = tmp * L.eta;
  E.e[3][1] = g;
  return E;
}

// E_u entry from the E_d entry.
template <int R, int C>
__host__ __device__ __forceinline__ cd eu(const EigD& E) {
  return ((R + C) & 1) ? -E.e[R][C] : E.e[R][C];
}

__host__ __device__ __forceinline__ double cabs1(const cd& z) { return fabs(z.r) + fabs(z.i); }

// Solves A X = B in place (X overwrites B) for an N x N complex A and NR
// right-hand sides: LU with partial pivoting (largest |re| + |im|, first of
// equals), then forward and back substitution.
template <int N, int NR>
__host__ __device__ __forceinline__ void lu_solve(cd (&A)[N][N], cd (&B)[N][NR]) {
  // Every loop has a constant trip count and guards on the outer index, so
  // full unrolling leaves only compile-time register indices.
  cd inv[N];
#pragma unroll
  for (int c = 0; c < N; ++c) {
    int piv = c;
    double best = cabs1(A[c][c]);
#pragma unroll
    for (int r = 0; r < N; ++r) {
      if (r > c) {
        const double v = cabs1(A[r][c]);
        if (v > best) {
          best = v;
          piv = r;
        }
      }
    }
    // exchange rows c and piv: the new row c is chosen among the candidates
    // by a chain of value selects, then the old row c moves to the pivot row
#pragma unroll
    for (int j = 0; j < N + NR; ++j) {
      if (j >= c) {
        const cd old = j < N ? A[c][j] : B[c][j - N];
        cd pick = old;
#pragma unroll
        for (int r = 0; r < N; ++r) {
          if (r > c) {
            const cd cand = j < N ? A[r][j] : B[r][j - N];
            pick.r = piv == r ? cand.r : pick.r;
            pick.i = piv == r ? cand.i : pick.i;
          }
        }
#pragma unroll
        for (int r = 0; r < N; ++r) {
          if (r > c) {
            cd& dst = j < N ? A[r][j] : B[r][j - N];
            dst.r = piv == r ? old.r : dst.r;
            dst.i = piv == r ? old.i : dst.i;
          }
        }
        (j < N ? A[c][j] : B[c][j - N]) = pick;
      }
    }
    inv[c] = recip(A[c][c]);
#pragma unroll
    for (int r = 0; r < N; ++r) {
      if (r > c) {
        const cd f = A[r][c] * inv[c];
#pragma unroll
        for (int j = 0; j < N; ++j)
          if (j > c) A[r][j] = A[r][j] - f * A[c][j];
#pragma unroll
        for (int j = 0; j < NR; ++j) B[r][j] = B[r][j] - f * B[c][j];
      }
    }
  }
#pragma unroll
  for (int cc = 0; cc < N; ++cc) {
    const int c = N - 1 - cc;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      cd s = B[c][j];
#pragma unroll
      for (int k = 0; k < N; ++k)
        if (k > c) s = s - A[c][k] * B[k][j];
      B[c][j] = s * inv[c];
    }
  }
}

// One Riccati step at interface k for one frequency: Y (2x2) in and out.
// Eb: E_d of layer k+1, phb: its phases (ignored when below_half is true),
// Ek, phk: layer k.
__host__ __device__ __forceinline__ void riccati_step(cd (&Y)[2][2], const EigD& Ek,
                                                      const cd (&phk)[2], const EigD& Eb,
                                                      const cd (&phb)[2], bool below_half) {
  cd A[4][4];
  cd B[4][2];
  // E_eff = E_d[k+1] + E_u[k+1] diag(ph[k+1]) Y
  cd eup[4][2];
  eup[0][0] = eu<0, 0>(Eb) * phb[0];
  eup[0][1] = eu<0, 1>(Eb) * phb[1];
  eup[1][0] = eu<1, 0>(Eb) * phb[0];
  eup[1][1] = eu<1, 1>(Eb) * phb[1];
  eup[2][0] = eu<2, 0>(Eb) * phb[0];
  eup[2][1] = eu<2, 1>(Eb) * phb[1];
  eup[3][0] = eu<3, 0>(Eb) * phb[0];
  eup[3][1] = eu<3, 1>(Eb) * phb[1];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      cd eff = Eb.e[r][c];
      if (!below_half) eff = eff + (eup[r][0] * Y[0][c] + eup[r][1] * Y[1][c]);
      A[r][2 + c] = -eff;
      B[r][c] = -(Ek.e[r][c] * phk[c]);
    }
  }
  A[0][0] = eu<0, 0>(Ek);
  A[0][1] = eu<0, 1>(Ek);
  A[1][0] = eu<1, 0>(Ek);
  A[1][1] = eu<1, 1>(Ek);
  A[2][0] = eu<2, 0>(Ek);
  A[2][1] = eu<2, 1>(Ek);
  A[3][0] = eu<3, 0>(Ek);
  A[3][1] = eu<3, 1>(Ek);
  lu_solve<4, 2>(A, B);
  Y[0][0] = B[0][0];
  Y[0][1] = B[0][1];
  Y[1][0] = B[1][0];
  Y[1][1] = B[1][1];
}

// R at the ocean bottom from Y at the top of layer 1 (Y ignored when layer 1
// is the half-space).
__host__ __device__ __forceinline__ cd riccati_surface(const cd (&Y)[2][2], const EigD& E1d,
                                                       const cd (&ph1)[2], bool layer1_half,
                                                       cd eta0, double rho0, cd e0,
                                                       int free_surface) {
  cd E1[4][2];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int c = 0; c < 2; ++c) E1[r][c] = E1d.e[r][c];
  }
  if (!layer1_half) {
    const cd u[4][2] = {{eu<0, 0>(E1d) * ph1[0], eu<0, 1>(E1d) * ph1[1]},
                        {eu<1, 0>(E1d) * ph1[0], eu<1, 1>(E1d) * ph1[1]},
                        {eu<2, 0>(E1d) * ph1[0], eu<2, 1>(E1d) * ph1[1]},
                        {eu<3, 0>(E1d) * ph1[0], eu<3, 1>(E1d) * ph1[1]}};
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int c = 0; c < 2; ++c) E1[r][c] = E1[r][c] + (u[r][0] * Y[0][c] + u[r][1] * Y[1][c]);
    }
  }
  const cd ed0 = eta0, ed1{-rho0, 0.0};
  const cd eu0 = -eta0, eu1{-rho0, 0.0};
  cd M[2][2] = {{eu0 * E1[2][0] - eu1 * E1[1][0], eu0 * E1[2][1] - eu1 * E1[1][1]},
                {E1[3][0], E1[3][1]}};
  cd D[2][1] = {{(eu0 * ed1 - eu1 * ed0) * e0}, {cd{0.0, 0.0}}};
  lu_solve<2, 1>(M, D);
  const cd uz = E1[1][0] * D[0][0] + E1[1][1] * D[1][0];
  const cd U0 = (uz - ed0 * e0) / eu0;
  const cd R = e0 * U0;
  return free_surface ? R / (cd{1.0, 0.0} + R) : R;
}

}  // namespace kh
