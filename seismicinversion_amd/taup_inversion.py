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

The host keeps what the reference keeps there: the P x P damped solve
(np.linalg.solve), the damping schedule (lambda_0 = 0.01 max|diag H|,
x5 on rejection, /2 on acceptance, at most 10 trials) and the convergence
tests, so the iteration sequence follows taup_inversion.py:389-439.
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


class _Graphed:
    """A device computation of a (nl, 6) layer table, captured once as a HIP
    graph (torch.cuda.CUDAGraph) and replayed for every new table.

    The caller runs the computation eagerly once first (its allocations and
    any lazy set-up are then settled); the capture itself executes nothing,
    so no computation runs more often than it is asked for.  The table is
    copied into a static input before each replay; the outputs are static
    tensors overwritten by the next replay, so callers copy them out first.
    Capture is thread-local (other threads may use the device meanwhile).
    The captured work is this library's kernels and small torch reductions
    along short axes; none of it records a memset node."""

    def __init__(self, fn, layers: torch.Tensor):
        dev = layers.device
        self.inp = layers.detach().clone()
        cur = torch.cuda.current_stream(dev)
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(cur)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.stream(side):
            # capture on the side stream; unlike the torch.cuda.graph context
            # this neither synchronises the whole device nor empties the
            # allocator cache, so other threads' work is not held up
            self.graph.capture_begin(capture_error_mode="thread_local")
            try:
                self.out = fn(self.inp)
            finally:
                self.graph.capture_end()
        cur.wait_stream(side)

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
        self.li = torch.tensor([l for l, _ in idx], device=self.dev)
        self.ci = torch.tensor([c for _, c in idx], device=self.dev)
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
        self.unorder = torch.tensor(np.argsort(order), device=self.dev)  # packed k -> row
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
        # direction's own (the reverse sweep stops there); the rest of the
        # symmetric matrix comes from the deeper layers' rows
        if mode == "hvp":
            chi, grad, hvp, _ = derivatives.misfit_hvp(layers, self.basis, self.p, self.omega,
                                                       self.R_obs, self.free_surface,
                                                       partial=True)
            rows = hvp.sum(1)[self.unorder][:, self.li, self.ci]
            H = derivatives.symmetric_from_partial(rows, self.li)
        elif mode == "gram":
            chi, grad, S, dR = derivatives.misfit_hvp(layers, self.basis, self.p, self.omega,
                                                      self.R_obs, self.free_surface,
                                                      second_only=True, want_dR=True,
                                                      partial=True)
            X = torch.view_as_real(dR[self.unorder]).reshape(self.P, -1)
            H = engine.gram(X, 2.0) + derivatives.symmetric_from_partial(
                S.sum(1)[self.unorder][:, self.li, self.ci], self.li)
        else:
            raise ValueError(f"unknown hessian mode {mode!r}")
        g = grad.sum(0)[self.li, self.ci]
        return chi, g, H

    def grad_hessian(self, phys: np.ndarray, mode: str = "gram"):
        if mode not in ("hvp", "gram"):
            raise ValueError(f"unknown hessian mode {mode!r}")
        chi, g, H = self._run("gh_" + mode, lambda L: self._grad_hessian_dev(L, mode),
                              self.table(phys))
        return float(np.sum(chi.cpu().numpy())), g.cpu().numpy(), H.cpu().numpy()


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
                max_trials: int = 10, use_graphs: bool = False) -> InversionResult:
    """Newton-LM inversion from a perturbed start model (taup_inversion.py:262-470).
    ``use_graphs`` replays the iteration's device work from captured HIP
    graphs (same numbers, fewer launches from the host)."""
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

    log_x = np.log(init_params)
    misfits: list[float] = []
    gnorms: list[float] = []
    perrs: list[float] = []
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
