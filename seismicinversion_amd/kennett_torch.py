"""Drop-in for Kennett_Reflectivity/kennett_torch.py.

``kennett_reflectivity_torch`` keeps the reference signature
(kennett_torch.py:312-338) and is differentiable in alpha, beta, rho,
thickness, Q_alpha and Q_beta.  It is a torch.autograd.Function whose
derivative rules are the device kernels:

  backward          -> kh_vjp        (hand-written adjoint, sqrt(N) checkpoints)
  jvp               -> kh_jvp        (tangent-linear sweep)
  backward of backward and forward-over-reverse
                    -> kh_vjp_hvp    (tangent-over-adjoint)
  vmap of a backward over stacked cotangents (torch.func.jacrev) and of a jvp
  over stacked tangents (torch.func.jacfwd) batch into one launch each.

So ``.backward()``, ``torch.autograd.grad``, ``torch.func.jacrev`` and
``torch.func.hessian`` (forward-over-reverse) all work on closures that call
it, as the reference's callers do (taup_inversion.py:169-177).  Inputs may
live on the CPU (as the reference's callers keep them) or on the GPU; the
result comes back on the input device.  Slowness and frequencies are
constants (not differentiated).

``jacobian`` and ``hessian`` (kennett_torch.py:611-684) run one kernel
launch each: all P tangent columns of J at once (kh_jvp), all P
Hessian-vector products at once (kh_misfit_hvp).
"""

from __future__ import annotations

from typing import TYPE_CHECKING

import numpy as np
import torch

from . import _abi
from . import derivatives as D
from . import engine

if TYPE_CHECKING:
    from .layer_model import LayerModel

__all__ = [
    "complex_slowness_torch",
    "vertical_slowness_torch",
    "solid_solid_interface_torch",
    "ocean_bottom_interface_torch",
    "kennett_reflectivity_torch",
    "forward_model_torch",
    "jacobian",
    "hessian",
    "model_to_tensors",
    "_pack_params",
    "_unpack_params",
]

_CDTYPE = torch.complex128
_FDTYPE = torch.float64
HALFSPACE = 1e30  # thickness sentinel of the torch API (kennett_torch.py:374)


# ---------------------------------------------------------------- device Functions

def _dev():
    return _abi.device()


def _omega_c(omega):
    if isinstance(omega, torch.Tensor):
        engine.check_precision(omega, "omega")
        return omega.detach().to(_CDTYPE)
    return torch.as_tensor(np.asarray(omega), dtype=_CDTYPE)


def _pval(p) -> float:
    if isinstance(p, torch.Tensor):
        engine.check_precision(p, "p")
        return float(p.detach())
    return float(p)


class _Forward(torch.autograd.Function):
    """tab (nl, 6) -> R (nw,) complex128."""

    @staticmethod
    def forward(tab, p, omega, fs):
        dev = _dev()
        R = engine.forward(engine.layers_tensor(tab, dev), torch.tensor([p], dtype=_FDTYPE),
                           omega.to(dev), free_surface=fs)
        return R[0].to(tab.device)

    @staticmethod
    def setup_context(ctx, inputs, output):
        tab, p, omega, fs = inputs
        ctx.save_for_backward(tab, omega)
        ctx.save_for_forward(tab, omega)
        ctx.p, ctx.fs = p, fs

    @staticmethod
    def backward(ctx, G):
        tab, omega = ctx.saved_tensors
        return _Vjp.apply(tab, ctx.p, omega, ctx.fs, G), None, None, None

    @staticmethod
    def jvp(ctx, dtab, _dp, _domega, _dfs):
        tab, omega = ctx.saved_tensors
        return _Jvp.apply(tab, ctx.p, omega, ctx.fs, dtab)

    @staticmethod
    def vmap(info, in_dims, tab, p, omega, fs):
        if in_dims[0] is None:
            return _Forward.apply(tab, p, omega, fs), None
        tab = tab.movedim(in_dims[0], 0)
        return torch.stack([_Forward.apply(t, p, omega, fs) for t in tab]), 0


class _Jvp(torch.autograd.Function):
    """(tab, dtab) -> dR = J dtab (nw,)."""

    @staticmethod
    def forward(tab, p, omega, fs, dtab):
        dev = _dev()
        _, dR = D.jvp(engine.layers_tensor(tab, dev), dtab.reshape(1, *tab.shape),
                      torch.tensor([p], dtype=_FDTYPE), omega, free_surface=fs)
        return dR[0, 0].to(tab.device)

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def vmap(info, in_dims, tab, p, omega, fs, dtab):
        if in_dims[0] is not None:
            tab = tab.movedim(in_dims[0], 0)
            dt = dtab.movedim(in_dims[4], 0) if in_dims[4] is not None else \
                dtab.expand(tab.shape[0], *dtab.shape)
            return torch.stack([_Jvp.apply(t, p, omega, fs, d) for t, d in zip(tab, dt)]), 0
        if in_dims[4] is None:
            return _Jvp.apply(tab, p, omega, fs, dtab), None
        dt = dtab.movedim(in_dims[4], 0)
        dev = _dev()
        _, dR = D.jvp(engine.layers_tensor(tab, dev), dt, torch.tensor([p], dtype=_FDTYPE),
                      omega, free_surface=fs)
        return dR[:, 0].to(tab.device), 0


class _Vjp(torch.autograd.Function):
    """(tab, G) -> grad (nl, 6) = Re(sum_w conj(G_w) dR_w/dtab)."""

    @staticmethod
    def forward(tab, p, omega, fs, G):
        dev = _dev()
        grad, _ = D.vjp(engine.layers_tensor(tab, dev), torch.tensor([p], dtype=_FDTYPE), omega,
                        G.reshape(1, -1), free_surface=fs)
        return grad[0].to(tab.device)

    @staticmethod
    def setup_context(ctx, inputs, output):
        tab, p, omega, fs, G = inputs
        ctx.save_for_backward(tab, omega, G)
        ctx.save_for_forward(tab, omega, G)
        ctx.p, ctx.fs = p, fs

    @staticmethod
    def jvp(ctx, dtab, _dp, _domega, _dfs, dG):
        tab, omega, G = ctx.saved_tensors
        if dtab is None:
            dtab = torch.zeros_like(tab)
        return _VjpHvp.apply(tab, ctx.p, omega, ctx.fs, G, dtab, dG)

    @staticmethod
    def backward(ctx, gbar):
        # grad is symmetric-Hessian-linear in tab and linear in G:
        # tab-cotangent = d/dt grad(tab + t gbar, G);  G-cotangent = J gbar
        tab, omega, G = ctx.saved_tensors
        dtab = _VjpHvp.apply(tab, ctx.p, omega, ctx.fs, G, gbar, None)
        dG = _Jvp.apply(tab, ctx.p, omega, ctx.fs, gbar)
        return dtab, None, None, None, dG

    @staticmethod
    def vmap(info, in_dims, tab, p, omega, fs, G):
        if in_dims[0] is not None:
            tab = tab.movedim(in_dims[0], 0)
            Gs = G.movedim(in_dims[4], 0) if in_dims[4] is not None else \
                G.expand(tab.shape[0], *G.shape)
            return torch.stack([_Vjp.apply(t, p, omega, fs, g) for t, g in zip(tab, Gs)]), 0
        if in_dims[4] is None:
            return _Vjp.apply(tab, p, omega, fs, G), None
        # stacked cotangents of one model (torch.func.jacrev): one launch
        return _VjpMany.apply(tab, p, omega, fs, G.movedim(in_dims[4], 0)), 0


class _VjpMany(torch.autograd.Function):
    """(tab, Gs (B, nw)) -> grads (B, nl, 6): B cotangents of one model in one
    kh_vjp launch (the slowness repeated B times)."""

    @staticmethod
    def forward(tab, p, omega, fs, Gs):
        B = Gs.shape[0]
        grad, _ = D.vjp(engine.layers_tensor(tab, _dev()), torch.full((B,), p, dtype=_FDTYPE),
                        omega, Gs.reshape(B, -1), free_surface=fs)
        return grad.to(tab.device)

    @staticmethod
    def setup_context(ctx, inputs, output):
        tab, p, omega, fs, Gs = inputs
        ctx.save_for_backward(tab, omega, Gs)
        ctx.save_for_forward(tab, omega, Gs)
        ctx.p, ctx.fs = p, fs

    @staticmethod
    def jvp(ctx, dtab, _dp, _domega, _dfs, dGs):
        tab, omega, Gs = ctx.saved_tensors
        if dtab is None:
            dtab = torch.zeros_like(tab)
        return _VjpHvpMany.apply(tab, ctx.p, omega, ctx.fs, Gs, dtab, dGs)

    @staticmethod
    def vmap(info, in_dims, tab, p, omega, fs, Gs):
        B = info.batch_size
        tabs = list(tab.movedim(in_dims[0], 0)) if in_dims[0] is not None else [tab] * B
        Gss = list(Gs.movedim(in_dims[4], 0)) if in_dims[4] is not None else [Gs] * B
        return torch.stack([_VjpMany.apply(t, p, omega, fs, g) for t, g in zip(tabs, Gss)]), 0


class _VjpHvpMany(torch.autograd.Function):
    """(tab, Gs (B, nw), dtab (nl, 6), dGs (B, nw)) -> (B, nl, 6): the
    directional derivative of _VjpMany; batched directions (torch.func.jacfwd)
    become the direction axis of one kh_vjp_hvp launch."""

    @staticmethod
    def forward(tab, p, omega, fs, Gs, dtab, dGs):
        return _VjpHvpMany._run(tab, p, omega, fs, Gs, dtab.reshape(1, *tab.shape),
                                None if dGs is None else dGs.reshape(1, *Gs.shape))[0]

    @staticmethod
    def _run(tab, p, omega, fs, Gs, dtabs, dGss):
        B = Gs.shape[0]
        _, dgrad = D.vjp_hvp(engine.layers_tensor(tab, _dev()), dtabs,
                             torch.full((B,), p, dtype=_FDTYPE), omega, Gs.reshape(B, -1),
                             dGss, free_surface=fs)
        return dgrad.to(tab.device)

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def vmap(info, in_dims, tab, p, omega, fs, Gs, dtab, dGs):
        n = info.batch_size
        if in_dims[0] is None and in_dims[4] is None:
            dts = dtab.movedim(in_dims[5], 0) if in_dims[5] is not None else \
                dtab.expand(n, *dtab.shape)
            dGss = None
            if dGs is not None:
                dGss = dGs.movedim(in_dims[6], 0) if in_dims[6] is not None else \
                    dGs.expand(n, *dGs.shape)
            return _VjpHvpMany._run(tab, p, omega, fs, Gs, dts, dGss), 0

        def pick(x, dim):
            if x is None:
                return [None] * n
            return list(x.movedim(dim, 0)) if dim is not None else [x] * n

        outs = [_VjpHvpMany.apply(t, p, omega, fs, g, dt, dg) for t, g, dt, dg in zip(
            pick(tab, in_dims[0]), pick(Gs, in_dims[4]), pick(dtab, in_dims[5]),
            pick(dGs, in_dims[6]))]
        return torch.stack(outs), 0


class _VjpHvp(torch.autograd.Function):
    """(tab, G, dtab, dG) -> d/dt grad(tab + t dtab, G + t dG) (nl, 6)."""

    @staticmethod
    def forward(tab, p, omega, fs, G, dtab, dG):
        dev = _dev()
        _, dgrad = D.vjp_hvp(engine.layers_tensor(tab, dev), dtab.reshape(1, *tab.shape),
                             torch.tensor([p], dtype=_FDTYPE), omega, G.reshape(1, -1),
                             None if dG is None else dG.reshape(1, 1, -1), free_surface=fs)
        return dgrad[0, 0].to(tab.device)

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def vmap(info, in_dims, tab, p, omega, fs, G, dtab, dG):
        if in_dims[0] is not None or in_dims[4] is not None:
            B = info.batch_size

            def pick(x, d):
                if x is None:
                    return [None] * B
                return list(x.movedim(d, 0)) if d is not None else [x] * B

            outs = [_VjpHvp.apply(t, p, omega, fs, g, dt, dg) for t, g, dt, dg in zip(
                pick(tab, in_dims[0]), pick(G, in_dims[4]), pick(dtab, in_dims[5]),
                pick(dG, in_dims[6]))]
            return torch.stack(outs), 0
        if in_dims[5] is None and (dG is None or in_dims[6] is None):
            return _VjpHvp.apply(tab, p, omega, fs, G, dtab, dG), None
        B = info.batch_size
        dts = dtab.movedim(in_dims[5], 0) if in_dims[5] is not None else \
            dtab.expand(B, *dtab.shape)
        dGs = None
        if dG is not None:
            dGs = dG.movedim(in_dims[6], 0) if in_dims[6] is not None else \
                dG.expand(B, *dG.shape)
            dGs = dGs.reshape(B, 1, -1)
        dev = _dev()
        _, dgrad = D.vjp_hvp(engine.layers_tensor(tab, dev), dts,
                             torch.tensor([p], dtype=_FDTYPE), omega, G.reshape(1, -1), dGs,
                             free_surface=fs)
        return dgrad[:, 0].to(tab.device), 0


def _layer_table(alpha, beta, rho, thickness, Q_alpha, Q_beta) -> torch.Tensor:
    cols = [torch.as_tensor(x, dtype=_FDTYPE) for x in
            (alpha, beta, rho, thickness, Q_alpha, Q_beta)]
    for c in cols:
        engine.check_precision(c, "layer parameter")
    h = cols[3]
    h = torch.where(h < HALFSPACE, h, torch.full_like(h, float("inf")))
    return torch.stack([cols[0], cols[1], cols[2], h, cols[4], cols[5]], dim=1)


def kennett_reflectivity_torch(
    alpha: torch.Tensor,
    beta: torch.Tensor,
    rho: torch.Tensor,
    thickness: torch.Tensor,
    Q_alpha: torch.Tensor,
    Q_beta: torch.Tensor,
    p: float | torch.Tensor,
    omega: torch.Tensor,
    free_surface: bool = False,
) -> torch.Tensor:
    """Differentiable PP reflectivity at one slowness, shape (nw,) complex128.

    thickness >= 1e30 marks the half-space (kennett_torch.py:374)."""
    tab = _layer_table(alpha, beta, rho, thickness, Q_alpha, Q_beta)
    return _Forward.apply(tab, _pval(p), _omega_c(omega), bool(free_surface))


# ---------------------------------------------------------------- Jacobian / Hessian API

def _pack_params(alpha, beta, rho, thickness) -> torch.Tensor:
    """m = [Vp_1..N, Vs_1..N, rho_1..N, h_1..N-1] (kennett_torch.py:552-571)."""
    return torch.cat([alpha[1:], beta[1:], rho[1:], thickness[1:-1]])


def _unpack_params(params, alpha_full, beta_full, rho_full, thickness_full):
    """Inverse of _pack_params with the ocean and half-space fixed (kennett_torch.py:574-608)."""
    n_sub = alpha_full.shape[0] - 1
    n_thick = n_sub - 1
    a = params[:n_sub]
    b = params[n_sub:2 * n_sub]
    r = params[2 * n_sub:3 * n_sub]
    h = params[3 * n_sub:3 * n_sub + n_thick]
    return (torch.cat([alpha_full[:1].detach(), a]), torch.cat([beta_full[:1].detach(), b]),
            torch.cat([rho_full[:1].detach(), r]),
            torch.cat([thickness_full[:1].detach(), h, thickness_full[-1:].detach()]))


def packed_index(n_layers: int) -> list[tuple[int, int]]:
    """(layer, column) of each packed parameter in the (nl, 6) layer table."""
    rng = range(1, n_layers)
    return ([(l, 0) for l in rng] + [(l, 1) for l in rng] + [(l, 2) for l in rng]
            + [(l, 3) for l in range(1, n_layers - 1)])


def _basis(n_layers: int) -> torch.Tensor:
    idx = packed_index(n_layers)
    B = torch.zeros((len(idx), n_layers, 6), dtype=_FDTYPE)
    for k, (l, c) in enumerate(idx):
        B[k, l, c] = 1.0
    return B


def jacobian(alpha, beta, rho, thickness, Q_alpha, Q_beta, p, omega,
             free_surface: bool = False) -> torch.Tensor:
    """J (nw, P) complex128, J_ij = dR(omega_i)/dm_j for the packed parameters
    (kennett_torch.py:611-641), all P tangent columns in one kh_jvp launch."""
    tab = _layer_table(alpha, beta, rho, thickness, Q_alpha, Q_beta).detach()
    dev = _dev()
    _, dR = D.jvp(engine.layers_tensor(tab, dev), _basis(tab.shape[0]),
                  torch.tensor([_pval(p)], dtype=_FDTYPE), _omega_c(omega),
                  free_surface=free_surface)
    return dR[:, 0, :].T.contiguous().to(torch.as_tensor(alpha).device)


def hessian(alpha, beta, rho, thickness, Q_alpha, Q_beta, p, omega, R_obs=None,
            free_surface: bool = False) -> torch.Tensor:
    """H (P, P) of chi = ||R - R_obs||^2 in the packed parameters
    (kennett_torch.py:644-684): P exact Hessian-vector products
    (tangent-over-adjoint) in one kh_misfit_hvp_partial launch, each down to
    one layer below its parameter's layer, completed by symmetry."""
    tab = _layer_table(alpha, beta, rho, thickness, Q_alpha, Q_beta).detach()
    nl = tab.shape[0]
    om = _omega_c(omega)
    if R_obs is None:
        R_obs = torch.zeros(om.shape[0], dtype=_CDTYPE)
    dev = _dev()
    _, _, hvp, _ = D.misfit_hvp(engine.layers_tensor(tab, dev), _basis(nl),
                                torch.tensor([_pval(p)], dtype=_FDTYPE), om,
                                torch.as_tensor(R_obs).reshape(1, -1),
                                free_surface=free_surface, partial=True)
    idx = packed_index(nl)
    li = torch.tensor([l for l, _ in idx], device=hvp.device)
    ci = torch.tensor([c for _, c in idx], device=hvp.device)
    H = D.symmetric_from_partial(hvp[:, 0][:, li, ci], li)
    return H.contiguous().to(torch.as_tensor(alpha).device)


def model_to_tensors(model: "LayerModel", requires_grad: bool = False) -> dict[str, torch.Tensor]:
    """LayerModel -> dict of float64 tensors, half-space thickness -> 1e30
    (kennett_torch.py:692-725)."""
    thick = np.array(model.thickness, dtype=np.float64)
    thick[np.isinf(thick)] = HALFSPACE
    out = {}
    for name, arr in (("alpha", model.alpha), ("beta", model.beta), ("rho", model.rho),
                      ("thickness", thick), ("Q_alpha", model.Q_alpha),
                      ("Q_beta", model.Q_beta)):
        t = torch.tensor(np.asarray(arr, dtype=np.float64), dtype=_FDTYPE)
        if requires_grad and name in ("alpha", "beta", "rho", "thickness"):
            t = t.requires_grad_(True)
        out[name] = t
    return out


# ---------------------------------------------------------------- trace and scalar helpers

def forward_model_torch(alpha, beta, rho, thickness, Q_alpha, Q_beta, p, T: float = 64.0,
                        nw: int = 2048, free_surface: bool = False):
    """(time, seismogram), both (2*nw,): R * Ricker, Hermitian FFT
    (kennett_torch.py:488-544).  Differentiable through the reflectivity."""
    dw = 2.0 * np.pi / T
    wmax = nw * dw
    omega = torch.arange(1, nw, dtype=_FDTYPE) * dw
    a_src = 0.1 * wmax
    t0 = 5.0 / (a_src * (2.0 ** 0.5))
    Z = omega ** 2 / (4.0 * a_src ** 2)
    S = Z * torch.exp(-Z + 1j * omega * t0)
    R = kennett_reflectivity_torch(alpha, beta, rho, thickness, Q_alpha, Q_beta, p, omega,
                                   free_surface=free_surface)
    Y = R * S.to(R.device)
    nt = 2 * nw
    X = torch.zeros(nt, dtype=_CDTYPE, device=Y.device)
    X = torch.cat([X[:1], Y, X[:1], torch.conj(torch.flip(Y, [0]))])
    seismogram = torch.fft.fft(X).real
    time = torch.arange(nt, dtype=_FDTYPE) * (T / float(nt))
    return time, seismogram


def complex_slowness_torch(velocity: torch.Tensor, Q: torch.Tensor) -> torch.Tensor:
    """s = (4Q^2 + 2Qi)/((1 + 4Q^2) v) (kennett_torch.py:43-61)."""
    two_q = 2.0 * Q
    two_q_sq = two_q * two_q
    denom = (1.0 + two_q_sq) * velocity
    return torch.complex(two_q_sq / denom, two_q / denom)


def vertical_slowness_torch(slowness: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
    """sqrt((s+p)(s-p)) with Im > 0 (kennett_torch.py:69-86)."""
    eta = torch.sqrt((slowness + p) * (slowness - p))
    return torch.where(eta.imag <= 0, -eta, eta)


def solid_solid_interface_torch(p, eta1, neta1, rho1, beta1, eta2, neta2, rho2, beta2):
    """(Rd, Ru, Tu, Td) of a solid-solid interface from vertical slownesses,
    densities and complex S velocities (kennett_torch.py:94-183), as
    differentiable torch ops on scalar tensors.  A host utility for callers
    that assemble interfaces by hand; the reflectivity kernels evaluate the
    same formulas on the device."""
    C = _CDTYPE
    p = torch.as_tensor(p, dtype=_FDTYPE)
    rtrho1, rtrho2 = torch.sqrt(rho1.to(C)), torch.sqrt(rho2.to(C))
    rteta1, rteta2, rtneta1, rtneta2 = (torch.sqrt(eta1), torch.sqrt(eta2), torch.sqrt(neta1),
                                        torch.sqrt(neta2))
    psq, cp = (p * p).to(C), p.to(C)
    c1, c2 = rho1.to(C), rho2.to(C)
    d = 2.0 * (c2 * beta2 * beta2 - c1 * beta1 * beta1)
    a, b, c = (c2 - c1) - psq * d, c2 - psq * d, c1 + psq * d
    E, F = b * eta1 + c * eta2, b * neta1 + c * neta2
    G, H = a - d * eta1 * neta2, a - d * eta2 * neta1
    det = E * F + G * H * psq
    E, F, G, H = E / det, F / det, G / det, H / det
    q = (b * eta1 - c * eta2) * F
    r = (a + d * eta1 * neta2) * H * psq
    s_ = (a * b + c * d * eta2 * neta2) * cp / det
    t = (b * neta1 - c * neta2) * E
    u = (a + d * eta2 * neta1) * G * psq
    v = (a * c + b * d * eta1 * neta1) * cp / det
    m2ci = torch.tensor(-2.0j, dtype=C)
    rd01 = m2ci * rteta1 * rtneta1 * s_
    ru01 = m2ci * rteta2 * rtneta2 * v
    Rd = torch.stack([torch.stack([q - r, rd01]), torch.stack([rd01, t - u])])
    Td = torch.stack([
        torch.stack([2 * rteta1 * rtrho1 * rteta2 * rtrho2 * F,
                     m2ci * rtneta1 * rtrho1 * rteta2 * rtrho2 * G * cp]),
        torch.stack([m2ci * rteta1 * rtrho1 * rtneta2 * rtrho2 * H * cp,
                     2 * rtneta1 * rtrho1 * rtneta2 * rtrho2 * E])])
    Ru = torch.stack([torch.stack([-(q + u), ru01]), torch.stack([ru01, -(t + r)])])
    return Rd, Ru, Td.T, Td


def ocean_bottom_interface_torch(p, eta1, rho1, eta2, neta2, rho2, beta2):
    """(Rd, Ru, Tu, Td) of the acoustic-elastic interface (kennett_torch.py:191-285),
    differentiable torch ops on scalar tensors (host utility)."""
    C = _CDTYPE
    p = torch.as_tensor(p, dtype=_FDTYPE)
    rtrho1, rtrho2 = torch.sqrt(rho1.to(C)), torch.sqrt(rho2.to(C))
    rteta1, rteta2, rtneta2 = torch.sqrt(eta1), torch.sqrt(eta2), torch.sqrt(neta2)
    psq, cp = (p * p).to(C), p.to(C)
    c1, c2 = rho1.to(C), rho2.to(C)
    d = 2.0 * (c2 * beta2 * beta2)
    a, b, c = (c2 - c1) - psq * d, c2 - psq * d, c1 + psq * d
    E, F, G, H = b * eta1 + c * eta2, b, a - d * eta1 * neta2, -d * eta2
    det = E * F + G * H * psq
    E, F, G, H = E / det, F / det, G / det, H / det
    t1 = (b * eta1 - c * eta2) * F
    t2 = (a + d * eta1 * neta2) * H * psq
    t4, t5 = b * E, d * eta2 * G * psq
    t6 = b * d * eta1 * cp / det
    mci = torch.tensor(-1.0j, dtype=C)
    z = torch.zeros((), dtype=C)
    pdpd = 2.0 * rtrho1 * rteta1 * rtrho2 * rteta2 * F
    pdsd = 2.0 * mci * rtrho1 * rteta1 * rtrho2 * rtneta2 * H * cp
    pusd = 2.0 * mci * rteta2 * rtneta2 * t6
    Rd = torch.stack([torch.stack([t1 - t2, z]), torch.stack([z, z])])
    Td = torch.stack([torch.stack([pdpd, z]), torch.stack([pdsd, z])])
    Tu = torch.stack([torch.stack([pdpd, pdsd]), torch.stack([z, z])])
    Ru = torch.stack([torch.stack([-(t1 + t5), pusd]), torch.stack([pusd, -(t2 + t4)])])
    return Rd, Ru, Tu, Td
