"""NumPy restatement of the Kennett P-SV reflectivity (TEST INFRASTRUCTURE ONLY).

Follows, operation for operation:
  * complex slowness ............ Kennett_Reflectivity/layer_model.py:42-45, 140-167
  * vertical slowness ........... Kennett_Reflectivity/layer_model.py:69-74
  * solid-solid coefficients .... Kennett_Reflectivity/scattering_matrices.py:82-161
  * ocean-bottom coefficients ... Kennett_Reflectivity/scattering_matrices.py:209-298
  * batched upward sweep ........ Kennett_Reflectivity/kennett_reflectivity_gpu.py:134-272
  * single-p forward ............ Kennett_Reflectivity/kennett_reflectivity.py:156-328
  * Ricker spectrum ............. Kennett_Reflectivity/source.py:39-45
  * gather (float64 twin) ....... Kennett_Reflectivity/kennett_gather.py:195-303
                                   and kennett_gather_gpu.py:96-107, 216-282

Everything is vectorised over slowness (no per-p Python loop), in the
reference's (n_p, n_omega, 2, 2) layout, so it doubles as the timed CPU
baseline.  ``ctype=np.clongdouble`` runs the same operations in x87 extended
precision (64-bit significand, 2048 times finer than float64) on the float64
inputs taken as exact: the reference value against which the rounding error
of any float64 evaluation, the reference's own golden vectors included, is
measured (pinned against 40-digit mpmath, oracle/kennett_mp.py, in
tests/test_oracle.py).
"""

from __future__ import annotations

import numpy as np
from scipy.special import j0

__all__ = [
    "complex_slowness",
    "vertical_slowness",
    "layer_slownesses",
    "interface_coefficients",
    "reflectivity_batch",
    "reflectivity",
    "ricker_spectrum",
    "bessel_sum",
    "hermitian_ifft",
    "gather",
]


def _real(ctype):
    return np.float64 if np.dtype(ctype) == np.complex128 else np.longdouble


def complex_slowness(v, Q, ctype=np.complex128):
    """layer_model.py:42-45 (vectorised)."""
    v = np.asarray(v, dtype=np.float64).astype(_real(ctype))
    Q = np.asarray(Q, dtype=np.float64).astype(_real(ctype))
    two_q = 2.0 * Q
    two_q_sq = two_q * two_q
    denom = (1.0 + two_q_sq) * v
    return ((two_q_sq / denom) + 1j * (two_q / denom)).astype(ctype)


def vertical_slowness(s, p):
    """layer_model.py:69-74: sqrt((s+p)(s-p)) with Im(eta) > 0 enforced."""
    t = (s + p) * (s - p)
    eta = np.sqrt(t)
    return np.where(eta.imag <= 0.0, -eta, eta)


def layer_slownesses(alpha, beta, Q_alpha, Q_beta, p, ctype=np.complex128):
    """Per-layer slownesses for every p.

    Returns (eta, neta, beta_c) with eta, neta of shape (nl, np) and beta_c
    (complex S velocity, 1/s_s or 0) of shape (nl,).  neta of layer 0 is 0
    (kennett_reflectivity_gpu.py:139-147)."""
    alpha = np.asarray(alpha, np.float64)
    beta = np.asarray(beta, np.float64)
    p = np.atleast_1d(np.asarray(p, np.float64)).astype(ctype)
    s_p = complex_slowness(alpha, Q_alpha, ctype)
    s_s = np.zeros(len(alpha), ctype)
    pos = beta > 0
    s_s[pos] = complex_slowness(beta[pos], np.asarray(Q_beta, np.float64)[pos], ctype)
    beta_c = np.zeros(len(alpha), ctype)
    nz = np.abs(s_s) > 0
    beta_c[nz] = 1.0 / s_s[nz]
    eta = vertical_slowness(s_p[:, None], p[None, :])
    neta = vertical_slowness(s_s[:, None], p[None, :])
    neta[0, :] = 0.0
    return eta, neta, beta_c


def _solid_solid(p, eta1, neta1, rho1, beta1, eta2, neta2, rho2, beta2):
    """scattering_matrices.py:82-161, vectorised over p (all args broadcast)."""
    rtrho1 = np.sqrt(rho1)
    rtrho2 = np.sqrt(rho2)
    rteta1 = np.sqrt(eta1)
    rteta2 = np.sqrt(eta2)
    rtneta1 = np.sqrt(neta1)
    rtneta2 = np.sqrt(neta2)
    rtza1 = rteta1 * rtrho1
    rtza2 = rteta2 * rtrho2
    rtzb1 = rtneta1 * rtrho1
    rtzb2 = rtneta2 * rtrho2
    ctype = np.result_type(eta1, np.complex128)
    psq = (p * p).astype(ctype)
    crho1 = np.asarray(rho1).astype(ctype)
    crho2 = np.asarray(rho2).astype(ctype)
    drho = crho2 - crho1
    dmu = crho2 * beta2 * beta2 - crho1 * beta1 * beta1
    d = 2.0 * dmu
    psqd = psq * d
    a = drho - psqd
    b = crho2 - psqd
    c = crho1 + psqd
    E = b * eta1 + c * eta2
    F = b * neta1 + c * neta2
    G = a - d * eta1 * neta2
    H = a - d * eta2 * neta1
    det = E * F + G * H * psq
    E = E / det
    F = F / det
    G = G / det
    H = H / det
    q_v = (b * eta1 - c * eta2) * F
    r_v = (a + d * eta1 * neta2) * H * psq
    s_v = (a * b + c * d * eta2 * neta2) * p / det
    t_v = (b * neta1 - c * neta2) * E
    u_v = (a + d * eta2 * neta1) * G * psq
    v_v = (a * c + b * d * eta1 * neta1) * p / det
    m2ci = complex(0.0, -2.0)
    n = p.shape[0]
    Rd = np.empty((n, 2, 2), ctype)
    Ru = np.empty((n, 2, 2), ctype)
    Td = np.empty((n, 2, 2), ctype)
    Rd[:, 0, 0] = q_v - r_v
    Rd[:, 0, 1] = m2ci * rteta1 * rtneta1 * s_v
    Rd[:, 1, 0] = m2ci * rteta1 * rtneta1 * s_v
    Rd[:, 1, 1] = t_v - u_v
    Td[:, 0, 0] = 2 * rtza1 * rtza2 * F
    Td[:, 0, 1] = m2ci * rtzb1 * rtza2 * G * p
    Td[:, 1, 0] = m2ci * rtza1 * rtzb2 * H * p
    Td[:, 1, 1] = 2 * rtzb1 * rtzb2 * E
    Ru[:, 0, 0] = -(q_v + u_v)
    Ru[:, 0, 1] = m2ci * rteta2 * rtneta2 * v_v
    Ru[:, 1, 0] = m2ci * rteta2 * rtneta2 * v_v
    Ru[:, 1, 1] = -(t_v + r_v)
    Tu = np.swapaxes(Td, 1, 2).copy()
    return Rd, Ru, Tu, Td


def _ocean_bottom(p, eta1, rho1, eta2, neta2, rho2, beta2):
    """scattering_matrices.py:209-298, vectorised over p."""
    rtrho1 = np.sqrt(rho1)
    rtrho2 = np.sqrt(rho2)
    rteta1 = np.sqrt(eta1)
    rteta2 = np.sqrt(eta2)
    rtneta2 = np.sqrt(neta2)
    rtza1 = rtrho1 * rteta1
    rtza2 = rtrho2 * rteta2
    rtzb2 = rtrho2 * rtneta2
    ctype = np.result_type(eta1, np.complex128)
    psq = p * p
    crho1 = np.asarray(rho1).astype(ctype)
    crho2 = np.asarray(rho2).astype(ctype)
    drho = crho2 - crho1
    dmu = crho2 * beta2 * beta2
    d = 2.0 * dmu
    psqd = psq * d
    a = drho - psqd
    b = crho2 - psqd
    c = crho1 + psqd
    E = b * eta1 + c * eta2
    F = b
    G = a - d * eta1 * neta2
    H = -d * eta2
    det = E * F + G * H * psq
    E = E / det
    F = F / det
    G = G / det
    H = H / det
    t1 = (b * eta1 - c * eta2) * F
    t2 = (a + d * eta1 * neta2) * H * psq
    t4 = b * E
    t5 = d * eta2 * G * psq
    t6 = b * d * eta1 * p / det
    mci = complex(0.0, -1.0)
    pdpu = t1 - t2
    pdpd = 2.0 * rtza1 * rtza2 * F
    pdsd = 2.0 * mci * rtza1 * rtzb2 * H * p
    pusd = 2.0 * mci * rteta2 * rtneta2 * t6
    n = p.shape[0]
    Rd = np.zeros((n, 2, 2), ctype)
    Ru = np.zeros((n, 2, 2), ctype)
    Tu = np.zeros((n, 2, 2), ctype)
    Td = np.zeros((n, 2, 2), ctype)
    Rd[:, 0, 0] = pdpu
    Td[:, 0, 0] = pdpd
    Td[:, 1, 0] = pdsd
    Tu[:, 0, 0] = pdpd
    Tu[:, 0, 1] = pdsd
    Ru[:, 0, 0] = -(t1 + t5)
    Ru[:, 0, 1] = pusd
    Ru[:, 1, 0] = pusd
    Ru[:, 1, 1] = -(t2 + t4)
    return Rd, Ru, Tu, Td


def interface_coefficients(model, p, ctype=np.complex128):
    """Rd, Ru, Tu, Td for every interface and p: each (nl-1, np, 2, 2).

    Interface 0 is the ocean bottom, 1..nl-2 solid-solid
    (kennett_reflectivity_gpu.py:150-191)."""
    p = np.atleast_1d(np.asarray(p, np.float64)).astype(_real(ctype))
    eta, neta, beta_c = layer_slownesses(model.alpha, model.beta, model.Q_alpha,
                                         model.Q_beta, p, ctype)
    nl = model.n_layers
    out = [np.zeros((nl - 1, len(p), 2, 2), ctype) for _ in range(4)]
    rho = np.asarray(model.rho, np.float64).astype(_real(ctype))
    coeffs = _ocean_bottom(p, eta[0], rho[0], eta[1], neta[1], rho[1], beta_c[1])
    for k in range(4):
        out[k][0] = coeffs[k]
    for il in range(1, nl - 1):
        coeffs = _solid_solid(p, eta[il], neta[il], rho[il], beta_c[il],
                              eta[il + 1], neta[il + 1], rho[il + 1], beta_c[il + 1])
        for k in range(4):
            out[k][il] = coeffs[k]
    return tuple(out)


def _inv2x2(M):
    """kennett_reflectivity_gpu.py:56-85."""
    a = M[..., 0, 0]
    b = M[..., 0, 1]
    c = M[..., 1, 0]
    d = M[..., 1, 1]
    det = a * d - b * c
    inv = np.empty_like(M)
    inv[..., 0, 0] = d / det
    inv[..., 0, 1] = -b / det
    inv[..., 1, 0] = -c / det
    inv[..., 1, 1] = a / det
    return inv


def reflectivity_batch(model, p_samples, omega, free_surface=False, ctype=np.complex128):
    """kennett_reflectivity_gpu.py:88-275: R of shape (np, nw), complex128
    (or ``ctype``)."""
    p_samples = np.atleast_1d(np.asarray(p_samples, np.float64))
    omega_c = np.asarray(omega).astype(np.complex128).astype(ctype)
    nl = model.n_layers
    nfreq = omega_c.shape[0]
    eta, neta, _ = layer_slownesses(model.alpha, model.beta, model.Q_alpha,
                                    model.Q_beta, p_samples, ctype)
    Rd_all, Ru_all, Tu_all, Td_all = interface_coefficients(model, p_samples, ctype)

    def phases(i):
        h = np.float64(model.thickness[i])
        if np.isinf(h):
            one = np.ones((len(p_samples), nfreq), ctype)
            return one, one
        h = _real(ctype)(h)
        ea = np.exp(1j * eta[i, :, None] * omega_c[None, :] * h)
        eb = np.exp(1j * neta[i, :, None] * omega_c[None, :] * h)
        return ea, eb

    RRd = np.zeros((len(p_samples), nfreq, 2, 2), ctype)
    eye = np.eye(2, dtype=ctype)
    for iface in range(nl - 2, -1, -1):
        ea, eb = phases(iface + 1)
        eaea = ea ** 2
        ebeb = eb ** 2
        eaeb = ea * eb
        MT = np.empty_like(RRd)
        MT[:, :, 0, 0] = eaea * RRd[:, :, 0, 0]
        MT[:, :, 0, 1] = eaeb * RRd[:, :, 0, 1]
        MT[:, :, 1, 0] = eaeb * RRd[:, :, 1, 0]
        MT[:, :, 1, 1] = ebeb * RRd[:, :, 1, 1]
        U = _inv2x2(eye - np.matmul(Ru_all[iface][:, None], MT))
        RRd = Rd_all[iface][:, None] + np.matmul(
            np.matmul(Tu_all[iface][:, None], np.matmul(MT, U)), Td_all[iface][:, None])
    ea0, _ = phases(0)
    top = (ea0 ** 2) * RRd[:, :, 0, 0]
    if free_surface:
        return top / (1.0 + top)
    return top


def reflectivity(model, p, omega, free_surface=False):
    """kennett_reflectivity.py:73-328 (one p): R of shape (nw,)."""
    return reflectivity_batch(model, np.array([float(p)]), omega, free_surface)[0]


def ricker_spectrum(omega, omega_max):
    """source.py:39-45."""
    alpha = 0.1 * omega_max
    t0 = 5.0 / (alpha * np.sqrt(2.0))
    z = (omega ** 2) / (4.0 * alpha ** 2)
    return z * np.exp(-z + 1j * omega * t0)


def bessel_sum(R, p_samples, omega_real, offsets, dp, p_max=None):
    """U[r, w] = sum_j R[j,w] taper_j p_j w dp J0(w p_j r) in float64.

    Same formula as kennett_gather.py:111-120 (per-p contribution) and
    kennett_gather_gpu.py:100-107 (Hann taper, p_max = last slowness)."""
    if p_max is None:
        p_max = p_samples[-1]
    taper = 0.5 * (1.0 + np.cos(np.pi * p_samples / p_max))
    U = np.zeros((len(offsets), len(omega_real)), np.complex128)
    for jp, pj in enumerate(p_samples):
        arg = omega_real[None, :] * pj * offsets[:, None]
        weight = taper[jp] * pj * omega_real * dp
        U += j0(arg) * (R[jp] * weight)[None, :]
    return U


def hermitian_ifft(U, nw, gamma, T):
    """kennett_gather_gpu.py:274-282: real traces of length nt = 2*nw,
    multiplied by exp(-gamma t)."""
    nr = U.shape[0]
    nt = 2 * nw
    time = np.arange(nt, dtype=np.float64) * (T / float(nt))
    X = np.zeros((nr, nt), np.complex128)
    X[:, 1:nw] = U
    X[:, nw + 1:] = np.conj(U[:, ::-1])
    return time, np.real(np.fft.fft(X, axis=1)) * np.exp(-gamma * time)[None, :]


def gather(model, offsets, T=64.0, nw=2048, np_slow=4096, p_max=1.2, gamma=None,
           source_func=None, free_surface=False):
    """compute_gather / compute_gather_gpu in float64 (kennett_gather.py:195-303)."""
    offsets = np.asarray(offsets, np.float64)
    if gamma is None:
        gamma = np.pi / T
    dw = 2.0 * np.pi / T
    omega_real = np.arange(1, nw, dtype=np.float64) * dw
    wmax = nw * dw
    S = ricker_spectrum(omega_real, wmax) if source_func is None else source_func(omega_real, wmax)
    dp = p_max / np_slow
    p_samples = np.arange(1, np_slow + 1, dtype=np.float64) * dp
    R = reflectivity_batch(model, p_samples, omega_real + 1j * gamma, free_surface)
    U = bessel_sum(R, p_samples, omega_real, offsets, dp, p_max=p_max) * S[None, :]
    time, traces = hermitian_ifft(U, nw, gamma, T)
    return time, offsets, traces
