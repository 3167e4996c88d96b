"""Device derivative entry points (kh_jvp, kh_vjp, kh_misfit_grad,
kh_misfit_hvp, kh_vjp_hvp) on torch tensors in HBM.

Parameter layout everywhere: per layer the six columns
[alpha, beta, rho, h, Q_alpha, Q_beta] of the (nl, 6) layer table; gradients
and Hessian-vector products come back as (np, nl, 6) per slowness (callers
sum over slowness themselves).  Directions are (ndir, nl, 6).
"""

from __future__ import annotations

import threading
from contextlib import contextmanager

import torch

from . import _abi
from .engine import check_precision

__all__ = ["jvp", "vjp", "misfit_grad", "misfit_hvp", "vjp_hvp", "paced", "WORKSPACE_LIMIT",
           "symmetric_from_partial"]

# Directions per launch are capped so that one launch's workspace stays below
# this (the 20-layer LM's 75 directions at 128 x 512 keep their whole forward
# tape in 12.0 GB: one launch).
WORKSPACE_LIMIT = 16 << 30


def _f64(x, dev):
    check_precision(x)
    return x.detach().to(device=dev, dtype=torch.float64).contiguous()


def _c128(x, dev):
    check_precision(x)
    return x.detach().to(device=dev, dtype=torch.complex128).contiguous()


def _ws_bytes(nl, n_p, nw, ndir, dual, tape):
    """Workspace bytes keeping the whole forward tape, or (tape False) the
    smallest workspace, sqrt(N) checkpoints."""
    lib = _abi.lib()
    fn = lib.kh_deriv_workspace_bytes if tape else lib.kh_deriv_workspace_bytes_min
    return int(fn(nl, n_p, nw, ndir, int(dual)))


def _tape_fits(nl, n_p, nw, dual):
    """Whether one direction's whole-tape workspace stays under WORKSPACE_LIMIT
    (otherwise the calls run on sqrt(N) checkpoints; the numbers agree to
    rounding)."""
    return _ws_bytes(nl, n_p, nw, 1, dual, True) <= WORKSPACE_LIMIT


def _workspace(nl, n_p, nw, ndir, dual, dev, tape=None):
    if tape is None:
        tape = _tape_fits(nl, n_p, nw, dual)
    nbytes = _ws_bytes(nl, n_p, nw, ndir, dual, tape)
    return torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)


_pacing = threading.local()


@contextmanager
def paced(max_lane_updates: float, gate=None):
    """On this thread, within the block: split direction batches so that one
    launch covers at most ``max_lane_updates`` (direction, p, omega,
    interface) steps, and call ``gate()`` before each launch.  A server uses
    it to keep an inversion's derivative kernels short and to let the forward
    batches queued before each one run first (``gate`` makes the current
    stream wait for them)."""
    prev = getattr(_pacing, "cfg", None)
    _pacing.cfg = (float(max_lane_updates), gate)
    try:
        yield
    finally:
        _pacing.cfg = prev


def _gate():
    cfg = getattr(_pacing, "cfg", None)
    if cfg is not None and cfg[1] is not None:
        cfg[1]()


def _dir_chunks(nl, n_p, nw, ndir):
    """Direction chunks whose workspace stays under WORKSPACE_LIMIT (and, when
    paced, whose launches stay under the pacing cap), and whether they keep
    the whole tape."""
    tape = _tape_fits(nl, n_p, nw, True)
    # the workspace is a part per call (the phase and shared-tape tables)
    # plus a part per direction
    one = _ws_bytes(nl, n_p, nw, 1, True, tape)
    per = max(1, _ws_bytes(nl, n_p, nw, 2, True, tape) - one)
    fixed = max(0, one - per)
    step = max(1, min(ndir, (WORKSPACE_LIMIT - fixed) // per))
    cfg = getattr(_pacing, "cfg", None)
    if cfg is not None:
        step = max(1, min(step, int(cfg[0] // max(1, n_p * nw * (nl - 1)))))
    return [(s, min(ndir, s + step)) for s in range(0, ndir, step)], tape


def jvp(layers, dlayers, p, omega, free_surface=False):
    """(R (np, nw), dR (ndir, np, nw)) with dR[d] = J dlayers[d]."""
    dev = layers.device
    p = _f64(p, dev)
    omega = _c128(omega, dev)
    dlayers = _f64(dlayers, dev).reshape(-1, layers.shape[0], 6)
    ndir, nl = dlayers.shape[0], layers.shape[0]
    n_p, nw = p.numel(), omega.numel()
    R = torch.empty((n_p, nw), dtype=torch.complex128, device=dev)
    dR = torch.empty((ndir, n_p, nw), dtype=torch.complex128, device=dev)
    rc = _abi.lib().kh_jvp(_abi.ptr(layers), _abi.ptr(dlayers), ndir, nl, _abi.ptr(p), n_p,
                           _abi.ptr(omega), nw, int(bool(free_surface)), _abi.ptr(R),
                           _abi.ptr(dR), _abi.stream_handle(dev))
    _abi.check(rc, "kh_jvp")
    return R, dR


def vjp(layers, p, omega, G, free_surface=False):
    """(grad (np, nl, 6), R (np, nw)) for the cotangent G (np, nw)."""
    dev = layers.device
    p = _f64(p, dev)
    omega = _c128(omega, dev)
    nl, n_p, nw = layers.shape[0], p.numel(), omega.numel()
    G = _c128(G, dev).reshape(n_p, nw)
    grad = torch.empty((n_p, nl, 6), dtype=torch.float64, device=dev)
    R = torch.empty((n_p, nw), dtype=torch.complex128, device=dev)
    ws = _workspace(nl, n_p, nw, 1, False, dev)
    rc = _abi.lib().kh_vjp(_abi.ptr(layers), nl, _abi.ptr(p), n_p, _abi.ptr(omega), nw,
                           int(bool(free_surface)), _abi.ptr(G), _abi.ptr(grad), _abi.ptr(R),
                           _abi.ptr(ws), ws.numel(), _abi.stream_handle(dev))
    _abi.check(rc, "kh_vjp")
    return grad, R


def misfit_grad(layers, p, omega, R_obs, free_surface=False):
    """(chi (np,), grad (np, nl, 6), R (np, nw)) of chi_p = sum_w |R - R_obs|^2."""
    dev = layers.device
    p = _f64(p, dev)
    omega = _c128(omega, dev)
    nl, n_p, nw = layers.shape[0], p.numel(), omega.numel()
    R_obs = _c128(R_obs, dev).reshape(n_p, nw)
    chi = torch.empty((n_p, 2), dtype=torch.float64, device=dev)
    grad = torch.empty((n_p, nl, 6), dtype=torch.float64, device=dev)
    R = torch.empty((n_p, nw), dtype=torch.complex128, device=dev)
    ws = _workspace(nl, n_p, nw, 1, False, dev)
    rc = _abi.lib().kh_misfit_grad(_abi.ptr(layers), nl, _abi.ptr(p), n_p, _abi.ptr(omega), nw,
                                   int(bool(free_surface)), _abi.ptr(R_obs), _abi.ptr(chi),
                                   _abi.ptr(grad), _abi.ptr(R), _abi.ptr(ws), ws.numel(),
                                   _abi.stream_handle(dev))
    _abi.check(rc, "kh_misfit_grad")
    return chi[:, 0], grad, R


def misfit_hvp(layers, dlayers, p, omega, R_obs, free_surface=False, second_only=False,
               want_dR=False, partial=False):
    """Exact misfit Hessian-vector products along each direction.

    Returns (chi (np,), grad (np, nl, 6), hvp (ndir, np, nl, 6), dR (ndir, np, nw) or None).
    With ``partial`` (kh_misfit_hvp_partial) hvp[d] holds only the entries of
    the layers down to one below the deepest layer direction d moves (the
    rest are 0), except hvp[0], which is complete; ``symmetric_from_partial``
    assembles the Hessian of a parameter basis from such rows."""
    dev = layers.device
    p = _f64(p, dev)
    omega = _c128(omega, dev)
    nl, n_p, nw = layers.shape[0], p.numel(), omega.numel()
    R_obs = _c128(R_obs, dev).reshape(n_p, nw)
    dlayers = _f64(dlayers, dev).reshape(-1, nl, 6)
    ndir = dlayers.shape[0]
    grad = torch.empty((n_p, nl, 6), dtype=torch.float64, device=dev)
    hvp = torch.empty((ndir, n_p, nl, 6), dtype=torch.float64, device=dev)
    dR = torch.empty((ndir, n_p, nw), dtype=torch.complex128, device=dev) if want_dR else None
    chi = None
    chunks, tape = _dir_chunks(nl, n_p, nw, ndir)
    for s, e in chunks:
        nd = e - s
        ws = _workspace(nl, n_p, nw, nd, True, dev, tape)
        chi_c = torch.empty((nd, n_p, 2), dtype=torch.float64, device=dev)
        _gate()
        fn = _abi.lib().kh_misfit_hvp_partial if partial else _abi.lib().kh_misfit_hvp
        rc = fn(
            _abi.ptr(layers), _abi.ptr(dlayers[s:e]), nd, nl, _abi.ptr(p), n_p, _abi.ptr(omega),
            nw, int(bool(free_surface)), _abi.ptr(R_obs), int(bool(second_only)),
            _abi.ptr(chi_c), _abi.ptr(grad) if s == 0 else None, _abi.ptr(hvp[s:e]),
            _abi.ptr(dR[s:e]) if want_dR else None, _abi.ptr(ws), ws.numel(),
            _abi.stream_handle(dev))
        _abi.check(rc, "kh_misfit_hvp")
        if chi is None:
            chi = chi_c[0, :, 0]
    return chi, grad, hvp, dR


def symmetric_from_partial(rows: torch.Tensor, li: torch.Tensor) -> torch.Tensor:
    """The (P, P) Hessian from partial rows: rows[k, j] = entry j of the
    Hessian-vector product of the basis direction of parameter k, computed
    where layer li[j] <= li[k] + 1 (kh_misfit_hvp_partial).  Entry (j, k)
    is rows[k, j] there and rows[j, k] otherwise (the Hessian is
    symmetric)."""
    valid = li[:, None] <= li[None, :] + 1  # valid[j, k]: rows[k, j] was computed
    return torch.where(valid, rows.T, rows)


def vjp_hvp(layers, dlayers, p, omega, G, dG=None, free_surface=False):
    """(grad (np, nl, 6), dgrad (ndir, np, nl, 6)): directional derivatives of
    the vector-Jacobian product along (dlayers[d], dG[d])."""
    dev = layers.device
    p = _f64(p, dev)
    omega = _c128(omega, dev)
    nl, n_p, nw = layers.shape[0], p.numel(), omega.numel()
    G = _c128(G, dev).reshape(n_p, nw)
    dlayers = _f64(dlayers, dev).reshape(-1, nl, 6)
    ndir = dlayers.shape[0]
    if dG is not None:
        dG = _c128(dG, dev).reshape(ndir, n_p, nw)
    grad = torch.empty((n_p, nl, 6), dtype=torch.float64, device=dev)
    dgrad = torch.empty((ndir, n_p, nl, 6), dtype=torch.float64, device=dev)
    chunks, tape = _dir_chunks(nl, n_p, nw, ndir)
    for s, e in chunks:
        nd = e - s
        ws = _workspace(nl, n_p, nw, nd, True, dev, tape)
        _gate()
        rc = _abi.lib().kh_vjp_hvp(
            _abi.ptr(layers), _abi.ptr(dlayers[s:e]), nd, nl, _abi.ptr(p), n_p, _abi.ptr(omega),
            nw, int(bool(free_surface)), _abi.ptr(G), _abi.ptr(dG[s:e]) if dG is not None else None,
            _abi.ptr(grad) if s == 0 else None, _abi.ptr(dgrad[s:e]), None, _abi.ptr(ws),
            ws.numel(), _abi.stream_handle(dev))
        _abi.check(rc, "kh_vjp_hvp")
    return grad, dgrad
