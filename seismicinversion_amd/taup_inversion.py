"""Drop-in for the solver part of Kennett_Reflectivity/taup_inversion.py
(taup_inversion.py:56-470, 527-571): Newton-Levenberg-Marquardt in
log-parameter space with the exact Hessian.

Per iteration the device evaluates, for all slownesses in one launch, the
misfit chi = sum_p sum_w |R - R_obs|^2, its gradient (adjoint sweep) and the
exact Hessian in the packed parameters:

* hessian_mode "gram" (default): the P Jacobian columns and the second-order
  term 2 Re(sum conj(R - R_obs) d^2R[e_k]) come from one tangent-over-adjoint
  launch (kh_misfit_hvp, second_only); the Gauss-Newton block 2 Re(J^H J) is
  a Gram matrix on the FP64 matrix cores (kh_gram).
* hessian_mode "hvp": the P exact Hessian-vector products H e_k directly.

The rest of the iteration follows taup_inversion.py:389-439: the P x P
damped solve, the damping schedule (lambda_0 = 0.01 max|diag H|, x5 on
rejection, /2 on acceptance, at most 10 trials) and the convergence tests.
invert_taup runs it on the device by default (_DeviceLoop: all trials of
an iteration solved and evaluated at once, the first improving one
accepted), or on the host one trial at a time as the reference does
(device_loop=False, with np.linalg.solve).
"""

from __future__ import annotations

import logging
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi, derivatives, engine
from .kennett_torch import _pack_params, _unpack_params, model_to_tensors, packed_index
from .layer_model import LayerModel

__all__ = [
    "InversionResult",
    "taup_misfit_grad_hessian",
    "newton_lm_step",
    "invert_taup",
    "compute_taup_traces",
    "write_model_table_latex",
    "write_model_profiles_tikz",
    "plot_taup_traces",
    "plot_convergence_curves",
    "main",
]

logger = logging.getLogger(__name__)
_F = torch.float64
_C = torch.complex128


@dataclass
class InversionResult:
    """Convergence history and models of a tau-p inversion (taup_inversion.py:56-67)."""

    true_model: LayerModel
    initial_model: LayerModel
    recovered_model: LayerModel
    misfit_history: list[float] = field(default_factory=list)
    grad_norm_history: list[float] = field(default_factory=list)
    param_error_history: list[float] = field(default_factory=list)
    n_iterations: int = 0
    converged: bool = False


def _capture(fn, dev):
    """(graph, fn()) with fn's device work captured as a HIP graph
    (torch.cuda.CUDAGraph), not executed.  Capture runs on a side stream:
    unlike the torch.cuda.graph context it neither synchronises the whole
    device nor empties the allocator cache, and it is thread-local, so other
    threads' work is not held up."""
    cur = torch.cuda.current_stream(dev)
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(cur)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        graph.capture_begin(capture_error_mode="thread_local")
        try:
            out = fn()
        finally:
            graph.capture_end()
    cur.wait_stream(side)
    return graph, out


class _Graphed:
    """A device computation of a (nl, 6) layer table, captured once as a HIP
    graph and replayed for every new table.

    The caller runs the computation eagerly once first (its allocations and
    any lazy set-up are then settled); the capture itself executes nothing,
    so no computation runs more often than it is asked for.  The table is
    copied into a static input before each replay; the outputs are static
    tensors overwritten by the next replay, so callers copy them out first.
    The captured work is this library's kernels and small torch reductions
    along short axes; none of it records a memset node."""

    def __init__(self, fn, layers: torch.Tensor):
        self.inp = layers.detach().clone()
        self.graph, self.out = _capture(lambda: fn(self.inp), layers.device)

    def __call__(self, layers: torch.Tensor):
        self.inp.copy_(layers)
        self.graph.replay()
        return self.out


class _Problem:
    """Device-resident inversion data: reference layer table, slownesses,
    frequencies, observed reflectivity and the packed-parameter map.

    With ``use_graphs`` the two fixed-shape device computations of an LM
    iteration — misfit, gradient and Hessian of a model, and the misfit of a
    trial model — are captured as HIP graphs on first use and replayed; the
    host keeps the data-dependent control (solve, accept/reject, damping).
    Replays run exactly the kernels of the eager path, so both give
    identical numbers."""

    def __init__(self, alpha_ref, beta_ref, rho_ref, thickness_ref, Q_alpha, Q_beta, p_values,
                 omega, R_obs, free_surface=False, use_graphs: bool = False):
        self.dev = _abi.device()
        cols = [torch.as_tensor(x, dtype=_F).detach().cpu() for x in
                (alpha_ref, beta_ref, rho_ref, thickness_ref, Q_alpha, Q_beta)]
        self.refs = cols
        self.nl = cols[0].shape[0]
        self.p_values = [float(p) for p in p_values]
        self.p = torch.tensor(self.p_values, dtype=_F, device=self.dev)
        self.omega = torch.as_tensor(omega).detach().to(_C).to(self.dev)
        if isinstance(R_obs, dict):
            R_obs = torch.stack([torch.as_tensor(R_obs[p]) for p in p_values])
        self.R_obs = torch.as_tensor(R_obs).to(_C).reshape(len(self.p_values), -1).to(self.dev)
        self.free_surface = free_surface
        idx = packed_index(self.nl)
        self.P = len(idx)
        # the basis directions deepest layer first: with partial reverse
        # sweeps their cost falls with the layer, so the launch starts its
        # longest workgroups first (the first one, whose sweep also gives the
        # gradient, is complete either way)
        order = sorted(range(self.P), key=lambda k: -idx[k][0])
        basis = torch.zeros((self.P, self.nl, 6), dtype=_F)
        for k, kk in enumerate(order):
            l, c = idx[kk]
            basis[k, l, c] = 1.0
        self.basis = basis.to(self.dev)
        # packed parameter k -> its basis row
        self.unorder = torch.tensor(np.argsort(order), dtype=torch.int64, device=self.dev)
        self.use_graphs = use_graphs
        self._graphs: dict[str, _Graphed] = {}
        self._eager_once: set[str] = set()

    def table(self, phys: np.ndarray) -> torch.Tensor:
        a, b, r, h = _unpack_params(torch.as_tensor(phys, dtype=_F), *self.refs[:4])
        h = torch.where(h < 1e30, h, torch.full_like(h, float("inf")))
        tab = torch.stack([a, b, r, h, self.refs[4], self.refs[5]], dim=1)
        return engine.layers_tensor(tab, self.dev)

    def _run(self, key, fn, layers):
        """fn(layers); with graphs: eagerly on the first call, captured on
        the second and replayed from then on."""
        if not self.use_graphs:
            return fn(layers)
        g = self._graphs.get(key)
        if g is None:
            if key not in self._eager_once:
                self._eager_once.add(key)
                return fn(layers)
            g = self._graphs[key] = _Graphed(fn, layers)
        return g(layers)

    def _misfit_rows(self, layers: torch.Tensor) -> torch.Tensor:
        R = engine.forward(layers, self.p, self.omega, free_surface=self.free_surface)
        return engine.residual_norms(R, self.R_obs)

    def misfit(self, phys: np.ndarray) -> float:
        chi = self._run("misfit", self._misfit_rows, self.table(phys))
        return float(np.sum(chi.cpu().numpy()))

    def _grad_hessian_dev(self, layers: torch.Tensor, mode: str):
        # rows of the basis directions down to one layer below each
        # direction's own (the reverse sweep stops there); kh_lm_assemble
        # sums them over slowness and takes the rest of the symmetric matrix
        # from the deeper layers' rows
        if mode == "hvp":
            chi, grad, rows, _ = derivatives.misfit_hvp(layers, self.basis, self.p, self.omega,
                                                        self.R_obs, self.free_surface,
                                                        partial=True)
            G = None
        elif mode == "gram":
            chi, grad, rows, dR = derivatives.misfit_hvp(layers, self.basis, self.p, self.omega,
                                                         self.R_obs, self.free_surface,
                                                         second_only=True, want_dR=True,
                                                         partial=True)
            G = engine.gram(torch.view_as_real(dR).reshape(self.P, -1), 2.0)
        else:
            raise ValueError(f"unknown hessian mode {mode!r}")
        g, H = engine.lm_assemble(self.nl, rows, grad, G, self.unorder)
        return chi, g, H

    def grad_hessian(self, phys: np.ndarray, mode: str = "gram"):
        if mode not in ("hvp", "gram"):
            raise ValueError(f"unknown hessian mode {mode!r}")
        chi, g, H = self._run("gh_" + mode, lambda L: self._grad_hessian_dev(L, mode),
                              self.table(phys))
        return float(np.sum(chi.cpu().numpy())), g.cpu().numpy(), H.cpu().numpy()


class _DeviceLoop:
    """The LM loop of taup_inversion.py:389-439 resident on the device.

    One iteration is a fixed sequence of launches on the current stream,
    with no host round trip:
      kh_lm_tables (the iterate's layer table) -> the misfit / gradient /
      exact Hessian evaluation (kh_lm_assemble sums it) -> kh_lm_prepare
      (log transform, history, first damping, convergence tests) ->
      kh_lm_trials, kh_lm_trial_misfits, kh_lm_decide for trial 0 of the
      damping schedule -> the same three for trials 1..max_trials-1 at once
      (one LU solve per workgroup), which return at once when trial 0 was
      accepted.  The first trial below the current misfit is accepted, as
      the sequential loop would.
    The loop state, history and buffers are static tensors, so with
    ``use_graphs`` the second iteration is captured as one HIP graph and
    replayed for the rest (the first runs eagerly and settles every
    allocation).  The host only reads the 8-byte loop status, one iteration
    behind (so the device never waits for it); at most one evaluation past
    the stop is computed and changes nothing."""

    def __init__(self, prob: _Problem, true_params: np.ndarray, init_params: np.ndarray,
                 max_iter: int, tol: float, hessian_mode: str, max_trials: int):
        if hessian_mode not in ("hvp", "gram"):
            raise ValueError(f"unknown hessian mode {hessian_mode!r}")
        dev = self.dev = prob.dev
        self.prob, self.mode = prob, hessian_mode
        self.nl, self.P, self.np = prob.nl, prob.P, len(prob.p_values)
        self.ntr, self.max_iter, self.tol = int(max_trials), int(max_iter), float(tol)
        f64 = dict(dtype=_F, device=dev)
        self.ref = prob.table(np.asarray(true_params, dtype=np.float64))
        self.x_true = torch.tensor(true_params, **f64)
        self.log_x = torch.tensor(np.log(init_params), **f64)
        self.state = torch.tensor([-1.0, 0.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0], **f64)
        self.hist = torch.zeros((max(self.max_iter, 1), 3), **f64)
        self.table = torch.empty((self.nl, 6), **f64)
        self.g_log = torch.empty((self.P,), **f64)
        self.H_log = torch.empty((self.P, self.P), **f64)
        self.trial_x = torch.empty((max(self.ntr, 1), self.P), **f64)
        self.trial_tab = torch.empty((max(self.ntr, 1), self.nl, 6), **f64)
        self.trial_chi = torch.empty((max(self.ntr, 1), self.np), **f64)
        self.ws_bytes = int(_abi.lib().kh_lm_trials_workspace_bytes(self.nl, self.ntr))
        self.ws = torch.empty((max(self.ws_bytes, 1),), dtype=torch.uint8, device=dev)
        self.status = torch.zeros((max(self.max_iter, 1),), dtype=_F, pin_memory=True)

    def step(self) -> None:
        """Enqueue one iteration on the current stream."""
        L, ptr, prob = _abi.lib(), _abi.ptr, self.prob
        s = _abi.stream_handle(self.dev)
        _abi.check(L.kh_lm_tables(ptr(self.ref), self.nl, ptr(self.log_x), 1, ptr(self.table),
                                  None, s), "kh_lm_tables")
        chi, g, H = prob._grad_hessian_dev(self.table, self.mode)
        g, H = g.contiguous(), H.contiguous()
        _abi.check(L.kh_lm_prepare(self.nl, self.np, chi.data_ptr(), chi.stride(0), ptr(g),
                                   ptr(H), ptr(self.log_x), ptr(self.x_true), self.tol,
                                   self.max_iter, ptr(self.state), ptr(self.g_log),
                                   ptr(self.H_log), ptr(self.hist), s), "kh_lm_prepare")
        # trial 0 first; the rest of the schedule only when it was rejected
        # (their kernels return at once otherwise)
        batches = [(0, 1), (1, self.ntr - 1)] if self.ntr > 1 else [(0, self.ntr)]
        for first, n in batches:
            last = int(first + n == self.ntr)
            if n > 0:
                _abi.check(L.kh_lm_trials(self.nl, first, n, ptr(self.H_log), ptr(self.g_log),
                                          ptr(self.log_x), ptr(self.ref), ptr(self.state),
                                          ptr(self.trial_x), ptr(self.trial_tab), ptr(self.ws),
                                          self.ws.numel(), s), "kh_lm_trials")
                _abi.check(L.kh_lm_trial_misfits(ptr(self.trial_tab), self.nl, first, n,
                                                 ptr(prob.p), self.np, ptr(prob.omega),
                                                 prob.omega.numel(), int(bool(prob.free_surface)),
                                                 ptr(prob.R_obs), ptr(self.state),
                                                 ptr(self.trial_chi), s), "kh_lm_trial_misfits")
            _abi.check(L.kh_lm_decide(self.nl, first, n, last, self.np, ptr(self.trial_chi),
                                      ptr(self.trial_x), ptr(self.state), ptr(self.log_x), s),
                       "kh_lm_decide")

    def run(self, use_graphs: bool = False):
        """Run up to max_iter iterations; returns (history (n, 3), final log_x,
        status) on the host."""
        graph = None
        events = []
        for it in range(self.max_iter):
            if use_graphs and it > 0:
                if graph is None:
                    graph, _ = _capture(self.step, self.dev)
                graph.replay()
            else:
                self.step()
            self.status[it:it + 1].copy_(self.state[4:5], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            events.append(ev)
            if it > 0:
                events[it - 1].synchronize()
                if self.status[it - 1] != 0.0:
                    break
        torch.cuda.current_stream(self.dev).synchronize()
        state = self.state.cpu().numpy()
        n = int(state[3])
        return self.hist[:n].cpu().numpy(), self.log_x.cpu().numpy(), int(state[4])


# Largest packed-parameter count for which invert_taup takes the device loop
# by default: the damped P x P systems then fit in LDS (kh_lm_trials_workspace_bytes
# is 0 up to there).
DEVICE_LOOP_MAX_P = 83


def _log_transform(m, g, H):
    """g_log = m g, H_log = m m^T H + diag(g m) (taup_inversion.py:180-185)."""
    return m * g, np.outer(m, m) * H + np.diag(g * m)


def taup_misfit_grad_hessian(log_params, alpha_ref, beta_ref, rho_ref, thickness_ref, Q_alpha,
                             Q_beta, p_values, omega, R_obs, hessian_mode: str = "gram"):
    """(chi, g_log, H_log) summed over the slownesses (taup_inversion.py:112-187)."""
    prob = _Problem(alpha_ref, beta_ref, rho_ref, thickness_ref, Q_alpha, Q_beta, p_values, omega,
                    R_obs)
    m = np.exp(np.asarray(log_params, dtype=np.float64))
    chi, g, H = prob.grad_hessian(m, hessian_mode)
    g_log, H_log = _log_transform(m, g, H)
    return chi, g_log, H_log


def _eval_misfit_only(log_params, alpha_ref, beta_ref, rho_ref, thickness_ref, Q_alpha, Q_beta,
                      p_values, omega, R_obs) -> float:
    """Misfit without derivatives (taup_inversion.py:190-230)."""
    prob = _Problem(alpha_ref, beta_ref, rho_ref, thickness_ref, Q_alpha, Q_beta, p_values, omega,
                    R_obs)
    return prob.misfit(np.exp(np.asarray(log_params, dtype=np.float64)))


def newton_lm_step(g: np.ndarray, H: np.ndarray, lam: float) -> np.ndarray:
    """Solve (H + lam I) dx = -g (taup_inversion.py:238-254)."""
    return np.linalg.solve(H + lam * np.eye(len(g)), -g)


def _to_model(tab_cols, ref_model: LayerModel) -> LayerModel:
    a, b, r, h = (t.detach().cpu().numpy() for t in tab_cols)
    h = h.copy()
    h[h >= 1e29] = np.inf
    return LayerModel.from_arrays(alpha=a, beta=b, rho=r, thickness=h,
                                  Q_alpha=ref_model.Q_alpha, Q_beta=ref_model.Q_beta)


def invert_taup(true_model: LayerModel, p_values: list[float], nfreq: int = 64,
                perturbation: float = 0.15, max_iter: int = 100, seed: int = 42,
                tol: float = 1e-8, hessian_mode: str = "gram",
                max_trials: int = 10, use_graphs: bool = False,
                device_loop: bool | None = None) -> InversionResult:
    """Newton-LM inversion from a perturbed start model (taup_inversion.py:262-470).

    ``device_loop`` True runs the whole iteration on the device: the damped
    solves of all trials, their misfits and the accept / reject decision
    included (_DeviceLoop); False solves and decides on the host as the
    reference does, one trial at a time.  The default (None) takes the
    device loop while the damped systems fit in LDS (DEVICE_LOOP_MAX_P
    packed parameters, 22 layers), where it is the faster of the two; for
    deeper models the host's LAPACK solve is faster than one workgroup's.
    ``use_graphs`` replays the iterations from a captured HIP graph (same
    numbers, fewer launches)."""
    t = model_to_tensors(true_model)
    refs = (t["alpha"], t["beta"], t["rho"], t["thickness"], t["Q_alpha"], t["Q_beta"])
    dw = 2.0 * np.pi / 64.0
    omega = torch.arange(1, nfreq + 1, dtype=_F) * dw
    dev = _abi.device()
    obs = engine.forward(engine.layers_tensor(true_model, dev),
                         torch.tensor([float(p) for p in p_values], dtype=_F), omega.to(_C))
    prob = _Problem(*refs, p_values, omega, obs, use_graphs=use_graphs)

    true_params = _pack_params(*refs[:4]).numpy()
    n_params = len(true_params)
    n_sub = true_model.n_layers - 1
    rng = np.random.default_rng(seed)
    perturb = np.ones(n_params)
    n_vel = 2 * n_sub
    perturb[:n_vel] = rng.uniform(1.0 - perturbation, 1.0 + perturbation, n_vel)
    perturb[n_vel:n_vel + n_sub] = rng.uniform(1.0 - perturbation * 0.67,
                                               1.0 + perturbation * 0.67, n_sub)
    perturb[n_vel + n_sub:] = rng.uniform(1.0 - perturbation * 1.33,
                                          1.0 + perturbation * 1.33, n_sub - 1)
    init_params = true_params * perturb
    initial_model = _to_model(_unpack_params(torch.tensor(init_params), *refs[:4]), true_model)

    if device_loop is None:
        device_loop = len(true_params) <= DEVICE_LOOP_MAX_P
    if device_loop:
        loop = _DeviceLoop(prob, true_params, init_params, max_iter, tol, hessian_mode, max_trials)
        hist, log_x, status = loop.run(use_graphs)
        misfits = [float(v) for v in hist[:, 0]]
        gnorms = [float(v) for v in hist[:, 1]]
        perrs = [float(v) for v in hist[:, 2]]
        for iteration, (chi, gn, pe) in enumerate(hist):
            logger.info("iter %3d  chi=%.4e  ||g||=%.4e  err=%.4e", iteration, chi, gn, pe)
        if status == 2:
            logger.warning("Line search failed at iteration %d", len(misfits) - 1)
        return _result(true_model, initial_model, refs, log_x, misfits, gnorms, perrs)

    log_x = np.log(init_params)
    misfits = []
    gnorms = []
    perrs = []
    lam = None
    for iteration in range(max_iter):
        m = np.exp(log_x)
        chi, g_phys, H_phys = prob.grad_hessian(m, hessian_mode)
        g, H = _log_transform(m, g_phys, H_phys)
        grad_norm = float(np.linalg.norm(g))
        param_err = float(np.linalg.norm(m - true_params) / np.linalg.norm(true_params))
        misfits.append(chi)
        gnorms.append(grad_norm)
        perrs.append(param_err)
        logger.info("iter %3d  chi=%.4e  ||g||=%.4e  err=%.4e", iteration, chi, grad_norm,
                    param_err)
        if lam is None:
            lam = 0.01 * max(float(np.abs(np.diag(H)).max()), 1e-10)
        if grad_norm < tol:
            break
        if len(misfits) > 1 and abs(misfits[-2] - chi) / max(abs(chi), 1e-30) < tol:
            break
        accepted = False
        for _ in range(max_trials):
            trial = log_x + newton_lm_step(g, H, lam)
            if prob.misfit(np.exp(trial)) < chi:
                log_x = trial
                lam = max(lam / 2.0, 1e-15)
                accepted = True
                break
            lam *= 5.0
        if not accepted:
            logger.warning("Line search failed at iteration %d", iteration)
            break

    return _result(true_model, initial_model, refs, log_x, misfits, gnorms, perrs)


def _result(true_model, initial_model, refs, log_x, misfits, gnorms, perrs) -> InversionResult:
    """The recovered model and history (taup_inversion.py:441-470)."""
    recovered = _to_model(_unpack_params(torch.tensor(np.exp(log_x)), *refs[:4]), true_model)
    return InversionResult(true_model=true_model, initial_model=initial_model,
                           recovered_model=recovered, misfit_history=misfits,
                           grad_norm_history=gnorms, param_error_history=perrs,
                           n_iterations=len(misfits),
                           converged=bool(perrs and perrs[-1] < 0.01))


def compute_taup_traces(model: LayerModel, p_values: list[float], T: float = 64.0,
                        nw: int = 1024):
    """Ricker-convolved tau-p traces (taup_inversion.py:527-571): R * S through
    the Hermitian inverse FFT, all slownesses in one forward launch and one
    trace launch."""
    from .source import ricker_spectrum

    dev = _abi.device()
    dw = 2.0 * np.pi / T
    omega = np.arange(1, nw, dtype=np.float64) * dw
    S = torch.from_numpy(ricker_spectrum(omega, nw * dw)).to(dev)
    R = engine.forward_grid(engine.layers_tensor(model, dev),
                            torch.tensor([float(p) for p in p_values], dtype=_F), 1.0, dw, 0.0,
                            nw - 1)
    tr = engine.hermitian_traces(R * S[None, :], nw, 0.0, T).cpu().numpy()
    time = np.arange(2 * nw, dtype=np.float64) * (T / float(2 * nw))
    return time, {p: tr[i] for i, p in enumerate(p_values)}


def __getattr__(name):  # the report writers live in taup_report
    if name in ("write_model_table_latex", "write_model_profiles_tikz", "plot_taup_traces",
                "plot_convergence_curves"):
        from . import taup_report

        return getattr(taup_report, name)
    raise AttributeError(name)


def main(argv: list[str] | None = None) -> None:
    """Inversion CLI (taup_inversion.py:1082-1198): a YAML config (default: the
    5-layer ocean-crust config shipped with the package), the same override
    flags, the device LM, then inversion_config.yaml and the configured
    report formats in the output directory."""
    import argparse
    from pathlib import Path

    from . import taup_report
    from .inversion_config import DEFAULT_CONFIG, load_config, save_config

    ap = argparse.ArgumentParser(description="Tau-p inversion with Newton-Levenberg-Marquardt "
                                             "on the GPU")
    ap.add_argument("--config", type=Path, default=None,
                    help="YAML config (default: the built-in 5-layer ocean-crust model)")
    ap.add_argument("--max-iter", type=int, default=None, help="Override max Newton iterations")
    ap.add_argument("--perturbation", type=float, default=None,
                    help="Override initial model perturbation fraction")
    ap.add_argument("--seed", type=int, default=None, help="Override random seed")
    ap.add_argument("--output-dir", type=Path, default=None, help="Override output directory")
    ap.add_argument("--hessian-mode", choices=("gram", "hvp"), default="gram",
                    help="exact Hessian as MFMA Gram + second-order term, or as HVPs")
    ap.add_argument("--graphs", action="store_true",
                    help="replay each iteration's device work from captured HIP graphs")
    args = ap.parse_args(argv)
    logging.basicConfig(level=logging.INFO, format="%(message)s")

    cfg = load_config(args.config if args.config is not None else DEFAULT_CONFIG)
    if args.max_iter is not None:
        cfg.max_iter = args.max_iter
    if args.perturbation is not None:
        cfg.perturbation = args.perturbation
    if args.seed is not None:
        cfg.seed = args.seed
    if args.output_dir is not None:
        cfg.output.directory = args.output_dir

    result = invert_taup(cfg.true_model, cfg.p_values, nfreq=cfg.nfreq,
                         perturbation=cfg.perturbation, max_iter=cfg.max_iter, seed=cfg.seed,
                         tol=cfg.tol, hessian_mode=args.hessian_mode, use_graphs=args.graphs)
    logger.info("Converged: %s", result.converged)
    logger.info("Iterations: %d", result.n_iterations)
    logger.info("Final misfit: %.2e", result.misfit_history[-1])
    logger.info("Final param error: %.4e", result.param_error_history[-1])

    out = Path(cfg.output.directory)
    out.mkdir(parents=True, exist_ok=True)
    save_config(cfg, out / "inversion_config.yaml")
    fmts = cfg.output.formats
    if "table" in fmts:
        taup_report.write_model_table_latex(result, out / "taup_model_parameters.tex",
                                            layer_names=cfg.layer_names)
    if "profiles" in fmts:
        taup_report.write_model_profiles_tikz(result, out / "taup_model_profiles.tex")
    if "traces" in fmts:
        td = cfg.output.trace_display
        taup_report.plot_taup_traces(result, out / "taup_trace_comparison.pdf", T=64.0,
                                     nw=td.nw, t_max=td.t_max)
    if "convergence" in fmts:
        taup_report.plot_convergence_curves(result, out / "taup_convergence.pdf")
    logger.info("Outputs in %s", out)


if __name__ == "__main__":
    main()
