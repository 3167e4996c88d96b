"""Pure-torch restatement of the differentiable Kennett forward (TEST INFRASTRUCTURE ONLY).

Follows Kennett_Reflectivity/kennett_torch.py:43-480 (complex slowness
:43-61, vertical slowness :69-86, interfaces :94-285, the sweep :312-480) and
its Jacobian/Hessian API (:552-684), on CPU tensors.  torch autograd through
it is the derivative oracle for the device adjoint, tangent and
Hessian-vector kernels: it is pinned against the reference's own golden
Jacobian/Hessian/Frechet vectors (tests/golden/torch_derivs.npz) and against
central finite differences of the NumPy oracle.
"""

from __future__ import annotations

import torch

C = torch.complex128
F = torch.float64


def complex_slowness(v, q):
    two_q = 2.0 * q
    two_q_sq = two_q * two_q
    denom = (1.0 + two_q_sq) * v
    return torch.complex(two_q_sq / denom, two_q / denom)


def vertical_slowness(s, p):
    t = (s + p) * (s - p)
    eta = torch.sqrt(t)
    return torch.where(eta.imag <= 0, -eta, eta)


def _solid(p, eta1, neta1, rho1, beta1, eta2, neta2, rho2, beta2):
    rtrho1, rtrho2 = torch.sqrt(rho1.to(C)), torch.sqrt(rho2.to(C))
    rteta1, rteta2 = torch.sqrt(eta1), torch.sqrt(eta2)
    rtneta1, rtneta2 = torch.sqrt(neta1), torch.sqrt(neta2)
    rtza1, rtza2 = rteta1 * rtrho1, rteta2 * rtrho2
    rtzb1, rtzb2 = rtneta1 * rtrho1, rtneta2 * rtrho2
    psq = (p * p).to(C)
    crho1, crho2 = rho1.to(C), rho2.to(C)
    drho = crho2 - crho1
    dmu = crho2 * beta2 * beta2 - crho1 * beta1 * beta1
    d = 2.0 * dmu
    psqd = psq * d
    a, b, c = drho - psqd, crho2 - psqd, crho1 + psqd
    E = b * eta1 + c * eta2
    Fv = b * neta1 + c * neta2
    G = a - d * eta1 * neta2
    H = a - d * eta2 * neta1
    det = E * Fv + G * H * psq
    E, Fv, G, H = E / det, Fv / det, G / det, H / det
    cp = p.to(C)
    q_v = (b * eta1 - c * eta2) * Fv
    r_v = (a + d * eta1 * neta2) * H * psq
    s_v = (a * b + c * d * eta2 * neta2) * cp / det
    t_v = (b * neta1 - c * neta2) * E
    u_v = (a + d * eta2 * neta1) * G * psq
    v_v = (a * c + b * d * eta1 * neta1) * cp / det
    m2ci = torch.tensor(-2.0j, dtype=C)
    rd01 = m2ci * rteta1 * rtneta1 * s_v
    Rd = torch.stack([torch.stack([q_v - r_v, rd01]), torch.stack([rd01, t_v - u_v])])
    Td = torch.stack([torch.stack([2 * rtza1 * rtza2 * Fv, m2ci * rtzb1 * rtza2 * G * cp]),
                      torch.stack([m2ci * rtza1 * rtzb2 * H * cp, 2 * rtzb1 * rtzb2 * E])])
    ru01 = m2ci * rteta2 * rtneta2 * v_v
    Ru = torch.stack([torch.stack([-(q_v + u_v), ru01]), torch.stack([ru01, -(t_v + r_v)])])
    return Rd, Ru, Td.T, Td


def _ocean_bottom(p, eta1, rho1, eta2, neta2, rho2, beta2):
    rtrho1, rtrho2 = torch.sqrt(rho1.to(C)), torch.sqrt(rho2.to(C))
    rteta1, rteta2, rtneta2 = torch.sqrt(eta1), torch.sqrt(eta2), torch.sqrt(neta2)
    rtza1, rtza2, rtzb2 = rtrho1 * rteta1, rtrho2 * rteta2, rtrho2 * rtneta2
    psq = (p * p).to(C)
    crho1, crho2 = rho1.to(C), rho2.to(C)
    drho = crho2 - crho1
    dmu = crho2 * beta2 * beta2
    d = 2.0 * dmu
    psqd = psq * d
    a, b, c = drho - psqd, crho2 - psqd, crho1 + psqd
    E = b * eta1 + c * eta2
    Fv = b
    G = a - d * eta1 * neta2
    H = -d * eta2
    det = E * Fv + G * H * psq
    E, Fv, G, H = E / det, Fv / det, G / det, H / det
    cp = p.to(C)
    t1 = (b * eta1 - c * eta2) * Fv
    t2 = (a + d * eta1 * neta2) * H * psq
    t4 = b * E
    t5 = d * eta2 * G * psq
    t6 = b * d * eta1 * cp / det
    mci = torch.tensor(-1.0j, dtype=C)
    zero = torch.zeros((), dtype=C)
    pdpd = 2.0 * rtza1 * rtza2 * Fv
    pdsd = 2.0 * mci * rtza1 * rtzb2 * H * cp
    pusd = 2.0 * mci * rteta2 * rtneta2 * t6
    Rd = torch.stack([torch.stack([t1 - t2, zero]), torch.stack([zero, zero])])
    Td = torch.stack([torch.stack([pdpd, zero]), torch.stack([pdsd, zero])])
    Tu = torch.stack([torch.stack([pdpd, pdsd]), torch.stack([zero, zero])])
    Ru = torch.stack([torch.stack([-(t1 + t5), pusd]), torch.stack([pusd, -(t2 + t4)])])
    return Rd, Ru, Tu, Td


def _inv2x2(M):
    a, b, c, d = M[:, 0, 0], M[:, 0, 1], M[:, 1, 0], M[:, 1, 1]
    det = a * d - b * c
    return torch.stack([torch.stack([d / det, -b / det], -1),
                        torch.stack([-c / det, a / det], -1)], -2)


def reflectivity(alpha, beta, rho, thickness, Q_alpha, Q_beta, p, omega, free_surface=False):
    """kennett_torch.py:312-480 (one p, vectorised over omega; thickness >= 1e30
    marks the half-space).  Differentiable in the six layer tensors."""
    p = torch.as_tensor(p, dtype=F)
    cp = p.to(C)
    nl = alpha.shape[0]
    nf = omega.shape[0]
    s_p = complex_slowness(alpha, Q_alpha)
    pos = beta > 0
    s_s_all = complex_slowness(torch.where(pos, beta, torch.ones_like(beta)),
                               torch.where(pos, Q_beta, torch.ones_like(Q_beta)))
    s_s = torch.where(pos, s_s_all, torch.zeros_like(s_s_all))
    nz = s_s.abs() > 0
    beta_c = torch.where(nz, 1.0 / torch.where(nz, s_s, torch.ones_like(s_s)),
                         torch.zeros_like(s_s))
    eta = vertical_slowness(s_p, cp)
    neta_all = vertical_slowness(s_s, cp)
    ocean = torch.zeros(nl, dtype=torch.bool)
    ocean[0] = True
    neta = torch.where(ocean, torch.zeros_like(neta_all), neta_all)
    finite = thickness < 1e30
    h = torch.where(finite, thickness, torch.zeros_like(thickness))
    ea = torch.where(finite[:, None], torch.exp(1j * omega[None, :] * (eta * h)[:, None]),
                     torch.ones((nl, nf), dtype=C))
    eb = torch.where(finite[:, None], torch.exp(1j * omega[None, :] * (neta * h)[:, None]),
                     torch.ones((nl, nf), dtype=C))
    coeffs = [_ocean_bottom(p, eta[0], rho[0], eta[1], neta[1], rho[1], beta_c[1])]
    for il in range(1, nl - 1):
        coeffs.append(_solid(p, eta[il], neta[il], rho[il], beta_c[il], eta[il + 1],
                             neta[il + 1], rho[il + 1], beta_c[il + 1]))
    RRd = torch.zeros((nf, 2, 2), dtype=C)
    eye = torch.eye(2, dtype=C)
    for iface in range(nl - 2, -1, -1):
        b = iface + 1
        Rd, Ru, Tu, Td = coeffs[iface]
        eaea, ebeb, eaeb = ea[b] ** 2, eb[b] ** 2, ea[b] * eb[b]
        MT = torch.stack([torch.stack([eaea * RRd[:, 0, 0], eaeb * RRd[:, 0, 1]], -1),
                          torch.stack([eaeb * RRd[:, 1, 0], ebeb * RRd[:, 1, 1]], -1)], -2)
        U = _inv2x2(eye[None] - torch.matmul(Ru[None], MT))
        RRd = Rd[None] + torch.matmul(torch.matmul(Tu[None], torch.matmul(MT, U)), Td[None])
    top = ea[0] ** 2 * RRd[:, 0, 0]
    return top / (1.0 + top) if free_surface else top


def pack_params(alpha, beta, rho, thickness):
    """kennett_torch.py:552-571."""
    return torch.cat([alpha[1:], beta[1:], rho[1:], thickness[1:-1]])


def unpack_params(params, alpha, beta, rho, thickness):
    """kennett_torch.py:574-608."""
    n = alpha.shape[0] - 1
    a = torch.cat([alpha[:1].detach(), params[:n]])
    b = torch.cat([beta[:1].detach(), params[n:2 * n]])
    r = torch.cat([rho[:1].detach(), params[2 * n:3 * n]])
    h = torch.cat([thickness[:1].detach(), params[3 * n:4 * n - 1], thickness[-1:].detach()])
    return a, b, r, h
