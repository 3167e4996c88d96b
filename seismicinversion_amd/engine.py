"""Device-level entry points: torch tensors in HBM in, torch tensors out.

These are the calls the drop-in modules, the gather and the benchmark build
on.  Every function launches on torch's current HIP stream of the tensors'
device and returns without synchronising.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _abi
from .layer_model import LayerModel, check_table, pack_layers

__all__ = [
    "layers_tensor",
    "uniform_grid",
    "forward",
    "forward_grid",
    "vertical_slowness",
    "interface_coefficients",
    "bessel_sum",
    "hermitian_traces",
    "interfaces_direct",
    "residual_norms",
    "gram",
    "lm_assemble",
    "gmm_forward",
    "gmm_forward_grid",
]


def layers_tensor(model: LayerModel | np.ndarray | torch.Tensor, dev=None) -> torch.Tensor:
    """The (nl, 6) float64 layer table of ``model`` on the device."""
    dev = _abi.device(dev)
    if isinstance(model, LayerModel):
        tab = torch.from_numpy(pack_layers(model))
    elif isinstance(model, np.ndarray):
        tab = torch.from_numpy(check_table(model, values=False))
    else:
        check_precision(model, "layer table")
        tab = model.detach().to(torch.float64)
    if tab.ndim != 2 or tab.shape[1] != 6:
        raise ValueError(f"layer table must be (n_layers, 6), got {tuple(tab.shape)}")
    return tab.to(dev).contiguous()


_LOW_PRECISION = (torch.float16, torch.bfloat16, torch.float32, torch.complex32, torch.complex64)


def check_precision(x, what="input"):
    """Reject reduced-precision tensors: the path computes in float64/complex128
    and a float32 slowness or frequency would silently change the result."""
    if isinstance(x, torch.Tensor) and x.dtype in _LOW_PRECISION:
        raise TypeError(f"{what} must be float64/complex128, got {x.dtype}")


def _as_dev(x, dtype, dev) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        check_precision(x)
        return x.detach().to(device=dev, dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x).astype(
        np.complex128 if dtype == torch.complex128 else np.float64))).to(dev)


def forward(layers: torch.Tensor, p: torch.Tensor, omega: torch.Tensor,
            free_surface: bool = False, out: torch.Tensor | None = None,
            block: int | None = None, omega_per_item: int | None = None) -> torch.Tensor:
    """R (np, nw) complex128 on the device for a layer table, slownesses and
    (possibly damped) frequencies already on the device."""
    dev = layers.device
    p = _as_dev(p, torch.float64, dev)
    omega = _as_dev(omega, torch.complex128, dev)
    n_p, n_w = p.numel(), omega.numel()
    if out is None:
        out = torch.empty((n_p, n_w), dtype=torch.complex128, device=dev)
    elif out.shape != (n_p, n_w) or out.dtype != torch.complex128 or not out.is_contiguous():
        raise ValueError("out must be a contiguous complex128 tensor of shape (np, nw)")
    lib = _abi.lib()
    s = _abi.stream_handle(dev)
    nl = layers.shape[0]
    if block is None and omega_per_item is None:
        rc = lib.kh_forward(_abi.ptr(layers), nl, _abi.ptr(p), n_p, _abi.ptr(omega), n_w,
                            int(bool(free_surface)), _abi.ptr(out), s)
    else:
        rc = lib.kh_forward_ex(_abi.ptr(layers), nl, _abi.ptr(p), n_p, _abi.ptr(omega), n_w,
                               int(bool(free_surface)), _abi.ptr(out), int(block or 256),
                               int(omega_per_item or 1), s)
    _abi.check(rc, "kh_forward")
    return out


def uniform_grid(omega) -> tuple[float, float, float] | None:
    """(k0, dw, gamma) if omega[j] == (k0 + j)*dw + i*gamma holds exactly for
    every j (the grids the reference builds with arange(...)*dw), else None."""
    w = np.asarray(omega).astype(np.complex128).reshape(-1)
    n = w.shape[0]
    if n < 2 or not np.all(w.imag == w.imag[0]) or not np.all(np.isfinite(w)):
        return None
    re = w.real
    d1 = re[1] - re[0]
    if not d1 > 0:
        return None
    k0 = float(np.round(re[0] / d1))
    cands = [d1, (re[-1] - re[0]) / (n - 1)]
    if k0 != 0:
        cands.append(re[0] / k0)
    j = np.arange(n, dtype=np.float64)
    for dw in cands:
        if np.array_equal((k0 + j) * dw, re):
            return k0, float(dw), float(w.imag[0])
    return None


def forward_grid(layers: torch.Tensor, p: torch.Tensor, k0: float, dw: float, gamma: float,
                 n_omega: int, free_surface: bool = False, out: torch.Tensor | None = None,
                 block: int | None = None, omega_per_item: int | None = None) -> torch.Tensor:
    """R (np, n_omega) on the uniform grid omega_j = (k0 + j)*dw + i*gamma."""
    dev = layers.device
    p = _as_dev(p, torch.float64, dev)
    n_p = p.numel()
    if out is None:
        out = torch.empty((n_p, n_omega), dtype=torch.complex128, device=dev)
    elif out.shape != (n_p, n_omega) or out.dtype != torch.complex128 or not out.is_contiguous():
        raise ValueError("out must be a contiguous complex128 tensor of shape (np, n_omega)")
    lib = _abi.lib()
    s = _abi.stream_handle(dev)
    nl = layers.shape[0]
    if block is None and omega_per_item is None:
        rc = lib.kh_forward_grid(_abi.ptr(layers), nl, _abi.ptr(p), n_p, float(k0), float(dw),
                                 float(gamma), int(n_omega), int(bool(free_surface)),
                                 _abi.ptr(out), s)
    else:
        rc = lib.kh_forward_grid_ex(_abi.ptr(layers), nl, _abi.ptr(p), n_p, float(k0),
                                    float(dw), float(gamma), int(n_omega),
                                    int(bool(free_surface)), _abi.ptr(out), int(block or 256),
                                    int(omega_per_item or 1), s)
    _abi.check(rc, "kh_forward_grid")
    return out


def vertical_slowness(layers: torch.Tensor, p) -> tuple[torch.Tensor, torch.Tensor]:
    """(eta, neta), each (nl, np) complex128 on the device."""
    dev = layers.device
    p = _as_dev(p, torch.float64, dev)
    nl = layers.shape[0]
    eta = torch.empty((nl, p.numel()), dtype=torch.complex128, device=dev)
    neta = torch.empty_like(eta)
    rc = _abi.lib().kh_vertical_slowness(_abi.ptr(layers), nl, _abi.ptr(p), p.numel(),
                                         _abi.ptr(eta), _abi.ptr(neta), _abi.stream_handle(dev))
    _abi.check(rc, "kh_vertical_slowness")
    return eta, neta


def interface_coefficients(layers: torch.Tensor, p):
    """(Rd, Ru, Tu, Td), each (nl-1, np, 2, 2) complex128 on the device."""
    dev = layers.device
    p = _as_dev(p, torch.float64, dev)
    nl = layers.shape[0]
    outs = [torch.empty((nl - 1, p.numel(), 2, 2), dtype=torch.complex128, device=dev)
            for _ in range(4)]
    rc = _abi.lib().kh_interface_coefficients(_abi.ptr(layers), nl, _abi.ptr(p), p.numel(),
                                              *[_abi.ptr(o) for o in outs],
                                              _abi.stream_handle(dev))
    _abi.check(rc, "kh_interface_coefficients")
    return tuple(outs)


def bessel_sum(R: torch.Tensor, p: torch.Tensor, dp: float, p_max: float, k0: float, dw: float,
               offsets: torch.Tensor, S: torch.Tensor | None = None,
               out: torch.Tensor | None = None) -> torch.Tensor:
    """U (nr, nw-1) complex128: Bessel-weighted slowness sum of R (np, nw-1)
    over real frequencies (k0 + k)*dw, times S if given."""
    dev = R.device
    n_p, nwm = R.shape
    p = _as_dev(p, torch.float64, dev)
    offsets = _as_dev(offsets, torch.float64, dev)
    nr = offsets.numel()
    if out is None:
        out = torch.empty((nr, nwm), dtype=torch.complex128, device=dev)
    lib = _abi.lib()
    nbytes = lib.kh_bessel_workspace_bytes(n_p, nwm + 1, nr)
    ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=dev)
    S_ptr = None if S is None else _abi.ptr(_as_dev(S, torch.complex128, dev))
    rc = lib.kh_bessel_sum(_abi.ptr(R.contiguous()), n_p, nwm + 1, _abi.ptr(p), float(dp),
                           float(p_max), float(k0), float(dw), _abi.ptr(offsets), nr, S_ptr,
                           _abi.ptr(out), _abi.ptr(ws), ws.numel(), _abi.stream_handle(dev))
    _abi.check(rc, "kh_bessel_sum")
    return out


def hermitian_traces(U: torch.Tensor, nw: int, gamma: float, T: float,
                     out: torch.Tensor | None = None) -> torch.Tensor:
    """Real traces (nr, 2*nw) float64 from U (nr, nw-1)."""
    dev = U.device
    nr = U.shape[0]
    if U.shape[1] != nw - 1:
        raise ValueError(f"U has {U.shape[1]} frequencies, expected nw-1 = {nw - 1}")
    if out is None:
        out = torch.empty((nr, 2 * nw), dtype=torch.float64, device=dev)
    rc = _abi.lib().kh_hermitian_traces(_abi.ptr(U.contiguous()), nr, int(nw), float(gamma),
                                        float(T), _abi.ptr(out), _abi.stream_handle(dev))
    _abi.check(rc, "kh_hermitian_traces")
    return out


def interfaces_direct(p, eta1, neta1, rho1, beta1, eta2, neta2, rho2, beta2, ocean: bool,
                      dev=None):
    """(Rd, Ru, Tu, Td) each (n, 2, 2) complex128 on the device for n interfaces
    given their vertical slownesses (kh_interfaces_direct)."""
    dev = _abi.device(dev)
    f = [_as_dev(np.atleast_1d(np.asarray(x)), torch.float64, dev) for x in (p, rho1, rho2)]
    c = [None if x is None else _as_dev(np.atleast_1d(np.asarray(x)), torch.complex128, dev)
         for x in (eta1, neta1, beta1, eta2, neta2, beta2)]
    n = f[0].numel()
    outs = [torch.empty((n, 2, 2), dtype=torch.complex128, device=dev) for _ in range(4)]
    ptr = lambda t: None if t is None else _abi.ptr(t)  # noqa: E731
    rc = _abi.lib().kh_interfaces_direct(
        n, int(bool(ocean)), ptr(f[0]), ptr(c[0]), ptr(c[1]), ptr(f[1]), ptr(c[2]), ptr(c[3]),
        ptr(c[4]), ptr(f[2]), ptr(c[5]), *[_abi.ptr(o) for o in outs], _abi.stream_handle(dev))
    _abi.check(rc, "kh_interfaces_direct")
    return tuple(outs)


def gmm_forward(layers: torch.Tensor, p: torch.Tensor, omega: torch.Tensor,
                free_surface: bool = False) -> torch.Tensor:
    """R (np, nw) by the Global Matrix Method's block-Riccati sweep (kh_gmm_forward)."""
    dev = layers.device
    p = _as_dev(p, torch.float64, dev)
    omega = _as_dev(omega, torch.complex128, dev)
    out = torch.empty((p.numel(), omega.numel()), dtype=torch.complex128, device=dev)
    rc = _abi.lib().kh_gmm_forward(_abi.ptr(layers), layers.shape[0], _abi.ptr(p), p.numel(),
                                   _abi.ptr(omega), omega.numel(), int(bool(free_surface)),
                                   _abi.ptr(out), _abi.stream_handle(dev))
    _abi.check(rc, "kh_gmm_forward")
    return out


def gmm_forward_grid(layers: torch.Tensor, p: torch.Tensor, k0: float, dw: float, gamma: float,
                     n_omega: int, free_surface: bool = False) -> torch.Tensor:
    """gmm_forward on omega_j = (k0 + j)*dw + i*gamma (kh_gmm_forward_grid)."""
    dev = layers.device
    p = _as_dev(p, torch.float64, dev)
    out = torch.empty((p.numel(), n_omega), dtype=torch.complex128, device=dev)
    rc = _abi.lib().kh_gmm_forward_grid(_abi.ptr(layers), layers.shape[0], _abi.ptr(p),
                                        p.numel(), float(k0), float(dw), float(gamma),
                                        int(n_omega), int(bool(free_surface)), _abi.ptr(out),
                                        _abi.stream_handle(dev))
    _abi.check(rc, "kh_gmm_forward_grid")
    return out


def residual_norms(R: torch.Tensor, R_obs: torch.Tensor,
                   out: torch.Tensor | None = None) -> torch.Tensor:
    """chi (np,) float64: sum over frequency of |R - R_obs|^2 per slowness row
    (kh_residual_norms, fixed summation order)."""
    check_precision(R, "R")
    check_precision(R_obs, "R_obs")
    if R.shape != R_obs.shape or R.dim() != 2:
        raise ValueError("R and R_obs must both be (np, nw)")
    R = R.to(torch.complex128).contiguous()
    R_obs = R_obs.to(device=R.device, dtype=torch.complex128).contiguous()
    n_p, nw = R.shape
    if out is None:
        out = torch.empty(n_p, dtype=torch.float64, device=R.device)
    rc = _abi.lib().kh_residual_norms(_abi.ptr(R), _abi.ptr(R_obs), n_p, nw, _abi.ptr(out),
                                      _abi.stream_handle(R.device))
    _abi.check(rc, "kh_residual_norms")
    return out


def lm_assemble(nl: int, rows: torch.Tensor, grad: torch.Tensor, G: torch.Tensor | None,
                order: torch.Tensor):
    """(g (P,), H (P, P)) in the packed parameters, summed over slownesses, from
    kh_misfit_hvp_partial's outputs over a basis (rows (P, np, nl, 6), grad
    (np, nl, 6); direction order[k] moves packed parameter k) plus the
    Gauss-Newton block G over the directions when given (kh_lm_assemble)."""
    dev = rows.device
    P, n_p = rows.shape[0], rows.shape[1]
    g = torch.empty((P,), dtype=torch.float64, device=dev)
    H = torch.empty((P, P), dtype=torch.float64, device=dev)
    order = order.to(device=dev, dtype=torch.int64).contiguous()
    rc = _abi.lib().kh_lm_assemble(nl, n_p, _abi.ptr(rows.contiguous()), _abi.ptr(grad),
                                   _abi.ptr(G) if G is not None else None, _abi.ptr(order),
                                   _abi.ptr(g), _abi.ptr(H), _abi.stream_handle(dev))
    _abi.check(rc, "kh_lm_assemble")
    return g, H


def gram(X: torch.Tensor, scale: float = 1.0, out: torch.Tensor | None = None) -> torch.Tensor:
    """scale * X X^T (P, P) float64 on the FP64 matrix cores (kh_gram); X (P, K)."""
    check_precision(X, "X")
    X = X.detach().to(torch.float64).contiguous()
    P, K = X.shape
    dev = X.device
    if out is None:
        out = torch.empty((P, P), dtype=torch.float64, device=dev)
    lib = _abi.lib()
    nbytes = int(lib.kh_gram_workspace_bytes(P, K))
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
    rc = lib.kh_gram(_abi.ptr(X), P, K, float(scale), _abi.ptr(out), _abi.ptr(ws), ws.numel(),
                     _abi.stream_handle(dev))
    _abi.check(rc, "kh_gram")
    return out
