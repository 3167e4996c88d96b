"""The Kennett reflectivity in multiple precision (TEST INFRASTRUCTURE ONLY).

The same algorithm as kennett_numpy (complex slowness layer_model.py:42-45,
vertical slowness :69-74, coefficients scattering_matrices.py:82-161 and
:209-298, the upward sweep kennett_reflectivity_gpu.py:218-272) evaluated
with mpmath at DPS decimal digits, the float64 inputs taken as exact.  It
gives the value every float64 implementation of the reference algorithm
approximates, so each one's rounding error (the reference's own golden
vectors included) can be measured against the same exact result.  Pinned by
tests/test_oracle.py: it agrees with the float64 oracle to the float64
rounding level on well-conditioned rows, and satisfies the free-surface
identity R_fs = R / (1 + R) to its own precision.

Scalar and slow (about 0.1 ms per (p, omega, interface)): small cases only.
"""

from __future__ import annotations

import mpmath as mp
import numpy as np

DPS = 40


def _cslow(v, Q):
    two_q = 2 * mp.mpf(Q)
    d = (1 + two_q * two_q) * mp.mpf(v)
    return mp.mpc(two_q * two_q / d, two_q / d)


def _vslow(s, p):
    eta = mp.sqrt((s + p) * (s - p))
    return -eta if mp.im(eta) <= 0 else eta


def _solid(p, e1, n1, r1, b1, e2, n2, r2, b2):
    rr1, rr2 = mp.sqrt(r1), mp.sqrt(r2)
    re1, re2, rn1, rn2 = mp.sqrt(e1), mp.sqrt(e2), mp.sqrt(n1), mp.sqrt(n2)
    za1, za2, zb1, zb2 = re1 * rr1, re2 * rr2, rn1 * rr1, rn2 * rr2
    psq = p * p
    dmu = r2 * b2 * b2 - r1 * b1 * b1
    d = 2 * dmu
    a, b, c = (r2 - r1) - psq * d, r2 - psq * d, r1 + psq * d
    E, F = b * e1 + c * e2, b * n1 + c * n2
    G, H = a - d * e1 * n2, a - d * e2 * n1
    det = E * F + G * H * psq
    E, F, G, H = E / det, F / det, G / det, H / det
    q = (b * e1 - c * e2) * F
    r = (a + d * e1 * n2) * H * psq
    s = (a * b + c * d * e2 * n2) * p / det
    t = (b * n1 - c * n2) * E
    u = (a + d * e2 * n1) * G * psq
    v = (a * c + b * d * e1 * n1) * p / det
    m2i = mp.mpc(0, -2)
    Rd = [[q - r, m2i * re1 * rn1 * s], [m2i * re1 * rn1 * s, t - u]]
    Td = [[2 * za1 * za2 * F, m2i * zb1 * za2 * G * p], [m2i * za1 * zb2 * H * p, 2 * zb1 * zb2 * E]]
    Ru = [[-(q + u), m2i * re2 * rn2 * v], [m2i * re2 * rn2 * v, -(t + r)]]
    Tu = [[Td[0][0], Td[1][0]], [Td[0][1], Td[1][1]]]
    return Rd, Ru, Tu, Td


def _ocean(p, e1, r1, e2, n2, r2, b2):
    rr1, rr2 = mp.sqrt(r1), mp.sqrt(r2)
    re1, re2, rn2 = mp.sqrt(e1), mp.sqrt(e2), mp.sqrt(n2)
    za1, za2, zb2 = rr1 * re1, rr2 * re2, rr2 * rn2
    psq = p * p
    d = 2 * (r2 * b2 * b2)
    a, b, c = (r2 - r1) - psq * d, r2 - psq * d, r1 + psq * d
    E, F = b * e1 + c * e2, b
    G, H = a - d * e1 * n2, -d * e2
    det = E * F + G * H * psq
    E, F, G, H = E / det, F / det, G / det, H / det
    t1 = (b * e1 - c * e2) * F
    t2 = (a + d * e1 * n2) * H * psq
    t4 = b * E
    t5 = d * e2 * G * psq
    t6 = b * d * e1 * p / det
    mi = mp.mpc(0, -1)
    pdpd = 2 * za1 * za2 * F
    pdsd = 2 * mi * za1 * zb2 * H * p
    pusd = 2 * mi * re2 * rn2 * t6
    z = mp.mpc(0)
    Rd = [[t1 - t2, z], [z, z]]
    Td = [[pdpd, z], [pdsd, z]]
    Tu = [[pdpd, pdsd], [z, z]]
    Ru = [[-(t1 + t5), pusd], [pusd, -(t2 + t4)]]
    return Rd, Ru, Tu, Td


def _mm(A, B):
    return [[A[0][0] * B[0][0] + A[0][1] * B[1][0], A[0][0] * B[0][1] + A[0][1] * B[1][1]],
            [A[1][0] * B[0][0] + A[1][1] * B[1][0], A[1][0] * B[0][1] + A[1][1] * B[1][1]]]


def reflectivity(model, p_samples, omega, free_surface=False, dps: int = DPS) -> np.ndarray:
    """R (np, nw) complex128, rounded from the exact-arithmetic value."""
    with mp.workdps(dps):
        alpha, beta, rho = model.alpha, model.beta, model.rho
        h, qa, qb = model.thickness, model.Q_alpha, model.Q_beta
        nl = len(alpha)
        s_p = [_cslow(alpha[i], qa[i]) for i in range(nl)]
        s_s = [_cslow(beta[i], qb[i]) if beta[i] > 0 else mp.mpc(0) for i in range(nl)]
        beta_c = [1 / s if s != 0 else mp.mpc(0) for s in s_s]
        R = np.zeros((len(p_samples), len(omega)), np.complex128)
        for jp, pv in enumerate(p_samples):
            p = mp.mpf(float(pv))
            eta = [_vslow(s, p) for s in s_p]
            neta = [mp.mpc(0)] + [_vslow(s_s[i], p) for i in range(1, nl)]
            rr = [mp.mpf(float(x)) for x in rho]
            co = [_ocean(p, eta[0], rr[0], eta[1], neta[1], rr[1], beta_c[1])]
            co += [_solid(p, eta[i], neta[i], rr[i], beta_c[i], eta[i + 1], neta[i + 1],
                          rr[i + 1], beta_c[i + 1]) for i in range(1, nl - 1)]
            for k, w in enumerate(omega):
                om = mp.mpc(complex(w))

                def ph(i):
                    if np.isinf(h[i]):
                        return mp.mpc(1), mp.mpc(1)
                    hh = mp.mpf(float(h[i]))
                    return mp.exp(1j * eta[i] * om * hh), mp.exp(1j * neta[i] * om * hh)

                S = [[mp.mpc(0), mp.mpc(0)], [mp.mpc(0), mp.mpc(0)]]
                for i in range(nl - 2, -1, -1):
                    ea, eb = ph(i + 1)
                    MT = [[ea * ea * S[0][0], ea * eb * S[0][1]],
                          [ea * eb * S[1][0], eb * eb * S[1][1]]]
                    Rd, Ru, Tu, Td = co[i]
                    RM = _mm(Ru, MT)
                    A = [[1 - RM[0][0], -RM[0][1]], [-RM[1][0], 1 - RM[1][1]]]
                    det = A[0][0] * A[1][1] - A[0][1] * A[1][0]
                    U = [[A[1][1] / det, -A[0][1] / det], [-A[1][0] / det, A[0][0] / det]]
                    Y = _mm(_mm(Tu, _mm(MT, U)), Td)
                    S = [[Rd[0][0] + Y[0][0], Rd[0][1] + Y[0][1]],
                         [Rd[1][0] + Y[1][0], Rd[1][1] + Y[1][1]]]
                ea0, _ = ph(0)
                top = ea0 * ea0 * S[0][0]
                r = top / (1 + top) if free_surface else top
                R[jp, k] = complex(r)
        return R
