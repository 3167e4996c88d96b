"""NumPy restatement of the Global Matrix Method reflectivity (TEST
INFRASTRUCTURE ONLY): the block-Riccati sweep and the dense global solve.

Follows:
  * eigenvector matrices ....... GlobalMatrix/layer_matrix.py:24-94
  * phases and eigenvectors .... GlobalMatrix/global_matrix.py:24-61
  * dense global system ........ GlobalMatrix/global_matrix.py:64-259 (no sources)
  * Riccati sweep .............. GlobalMatrix/riccati_solver.py:25-194 (no sources)
  * driver, free surface ....... GlobalMatrix/global_matrix.py:297-348

Pins (tests/test_gmm_cpu.py): the dense solve is an independent brute-force
statement of the same boundary-value problem, and both must agree with the
Kennett oracle (itself pinned by the reference's golden vectors) to the
reference's own test bars (GlobalMatrix/test_riccati.py:53-104, 212-348:
< 1e-10 relative, < 1e-12 absolute; < 1e-8 for the 22-layer model).
"""

from __future__ import annotations

import numpy as np

from .kennett_numpy import layer_slownesses

__all__ = ["layer_eigenvectors", "ocean_eigenvectors", "riccati_reflectivity",
           "dense_reflectivity"]


def layer_eigenvectors(p, eta, neta, rho, beta_c):
    """(E_d, E_u), each (4, 2): columns [P, S], rows
    [u_x, u_z, sigma_zz/(-i w), sigma_xz/(-i w)] (layer_matrix.py:24-68)."""
    mu = rho * beta_c ** 2
    gam = rho * (1.0 - 2.0 * beta_c ** 2 * p ** 2)
    E_d = np.array([[p, neta], [eta, -p], [-gam, 2.0 * mu * p * neta],
                    [2.0 * mu * p * eta, gam]], np.complex128)
    E_u = np.array([[p, -neta], [-eta, -p], [-gam, -2.0 * mu * p * neta],
                    [-2.0 * mu * p * eta, gam]], np.complex128)
    return E_d, E_u


def ocean_eigenvectors(p, eta, rho):
    """(e_d, e_u), each (2,): rows [u_z, sigma_zz/(-i w)] (layer_matrix.py:70-94)."""
    return (np.array([eta, -rho], np.complex128), np.array([-eta, -rho], np.complex128))


def _setup(model, p, omega):
    eta, neta, beta_c = layer_slownesses(model.alpha, model.beta, model.Q_alpha, model.Q_beta, p)
    eta, neta = eta[:, 0], neta[:, 0]
    nl = len(model.alpha)
    omega = np.asarray(omega).astype(np.complex128)
    cp = complex(p)
    E_d, E_u = {}, {}
    for j in range(1, nl):
        E_d[j], E_u[j] = layer_eigenvectors(cp, eta[j], neta[j], model.rho[j], beta_c[j])
    e0 = np.exp(1j * omega * eta[0] * model.thickness[0])
    phase = {}
    for j in range(1, nl - 1):
        phase[j] = np.stack([np.exp(1j * omega * eta[j] * model.thickness[j]),
                             np.exp(1j * omega * neta[j] * model.thickness[j])], axis=-1)
    e_d, e_u = ocean_eigenvectors(cp, eta[0], model.rho[0])
    return nl, E_d, E_u, phase, e_d, e_u, e0


def _ocean_bottom(E1, e_d, e_u, e0):
    """R from the effective eigenvectors at the top of layer 1
    (riccati_solver.py:129-194): u_z and sigma_zz continuity with the ocean,
    sigma_xz = 0."""
    row_a = e_u[0] * E1[:, 2, :] - e_u[1] * E1[:, 1, :]
    mat = np.stack([row_a, E1[:, 3, :]], axis=1)
    rhs = np.stack([(e_u[0] * e_d[1] - e_u[1] * e_d[0]) * e0, np.zeros_like(e0)], axis=1)
    D1 = np.linalg.solve(mat, rhs[..., None])[..., 0]
    uz = np.einsum("fi,fi->f", E1[:, 1, :], D1)
    U0 = (uz - e_d[0] * e0) / e_u[0]
    return e0 * U0


def riccati_reflectivity(model, p, omega, free_surface=False):
    """PP reflectivity (nw,) by the upward block-Riccati sweep."""
    nl, E_d, E_u, phase, e_d, e_u, e0 = _setup(model, p, omega)
    nf = e0.shape[0]
    M = nl - 2
    if M == 0:
        E1 = np.broadcast_to(E_d[1][None], (nf, 4, 2)).copy()
    else:
        Y = np.zeros((nf, 2, 2), np.complex128)
        for k in range(M, 0, -1):
            kb = k + 1
            if kb < nl - 1:
                Eup = E_u[kb][None] * phase[kb][:, None, :]
                Eeff = E_d[kb][None] + Eup @ Y
            else:
                Eeff = np.broadcast_to(E_d[kb][None], (nf, 4, 2)).copy()
            A_d = E_d[k][None] * phase[k][:, None, :]
            F = np.concatenate([np.broadcast_to(E_u[k][None], (nf, 4, 2)), -Eeff], axis=-1)
            Y = np.linalg.solve(F, -A_d)[:, :2, :]
        E1 = E_d[1][None] + (E_u[1][None] * phase[1][:, None, :]) @ Y
    R = _ocean_bottom(E1, e_d, e_u, e0)
    return R / (1.0 + R) if free_surface else R


def dense_reflectivity(model, p, omega, free_surface=False):
    """PP reflectivity (nw,) from the dense global system G x = b
    (global_matrix.py:64-276): unknowns U0^P, then (D^P, D^S, U^P, U^S) per
    finite layer, then (D^P, D^S) of the half-space."""
    nl, E_d, E_u, phase, e_d, e_u, e0 = _setup(model, p, omega)
    nf = e0.shape[0]
    nfin = nl - 2
    N = 1 + 4 * nfin + 2
    G = np.zeros((nf, N, N), np.complex128)
    b = np.zeros((nf, N), np.complex128)

    def col(j):
        return 0 if j == 0 else (1 + 4 * (j - 1) if j < nl - 1 else 1 + 4 * nfin)

    c1 = col(1)
    for r, erow in enumerate((1, 2, 3)):  # u_z, sigma_zz, sigma_xz at the ocean bottom
        if erow < 3:
            G[:, r, 0] = e_u[erow - 1]
            b[:, r] = -e_d[erow - 1] * e0
        G[:, r, c1] = -E_d[1][erow, 0]
        G[:, r, c1 + 1] = -E_d[1][erow, 1]
        if nl > 2:
            G[:, r, c1 + 2] = -E_u[1][erow, 0] * phase[1][:, 0]
            G[:, r, c1 + 3] = -E_u[1][erow, 1] * phase[1][:, 1]
    r = 3
    for k in range(1, nl - 1):
        ck, cn = col(k), col(k + 1)
        for e in range(4):
            G[:, r, ck] += E_d[k][e, 0] * phase[k][:, 0]
            G[:, r, ck + 1] += E_d[k][e, 1] * phase[k][:, 1]
            G[:, r, ck + 2] += E_u[k][e, 0]
            G[:, r, ck + 3] += E_u[k][e, 1]
            G[:, r, cn] -= E_d[k + 1][e, 0]
            G[:, r, cn + 1] -= E_d[k + 1][e, 1]
            if k + 1 < nl - 1:
                G[:, r, cn + 2] -= E_u[k + 1][e, 0] * phase[k + 1][:, 0]
                G[:, r, cn + 3] -= E_u[k + 1][e, 1] * phase[k + 1][:, 1]
            r += 1
    x = np.linalg.solve(G, b[..., None])[..., 0]
    R = e0 * x[:, 0]
    return R / (1.0 + R) if free_surface else R
