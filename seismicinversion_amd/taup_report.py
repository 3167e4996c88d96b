"""Report outputs of a tau-p inversion run — the four ``output.formats`` of the
inversion CLI (Kennett_Reflectivity/taup_inversion.py:579-1045):

* ``table``:       ``write_model_table_latex`` — booktabs/siunitx table of the
                   true, initial and recovered Vp, Vs, rho, h with the
                   recovered model's largest relative error per layer;
* ``profiles``:    ``write_model_profiles_tikz`` — TikZ figure, three panels
                   (Vp, Vs, rho) of step-function depth profiles;
* ``traces``:      ``plot_taup_traces`` — tau-p traces of the true and
                   recovered models (compute_taup_traces on the device) and
                   their residuals;
* ``convergence``: ``plot_convergence_curves`` — misfit, gradient norm and
                   relative parameter error per iteration.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

__all__ = ["write_model_table_latex", "write_model_profiles_tikz", "plot_taup_traces",
           "plot_convergence_curves", "depth_profile"]

LAYER_NAMES = ["Ocean", "Sediment", "Crust", "Upper mantle", "Half-space"]


def _models(result):
    return (("True", result.true_model), ("Initial", result.initial_model),
            ("Recovered", result.recovered_model))


def _layer_error(model, true, i) -> float:
    """Largest relative error of layer i over the parameters the inversion
    moves (Vp, rho; Vs where the true layer is elastic; h where finite)."""
    pairs = [(model.alpha[i], true.alpha[i]), (model.rho[i], true.rho[i])]
    if true.beta[i] > 0:
        pairs.append((model.beta[i], true.beta[i]))
    if np.isfinite(true.thickness[i]):
        pairs.append((model.thickness[i], true.thickness[i]))
    return max(abs(v - t) / max(abs(t), 1e-30) for v, t in pairs)


def write_model_table_latex(result, output, layer_names: list[str] | None = None) -> None:
    """LaTeX table of the three models (needs booktabs and siunitx)."""
    names = list(layer_names) if layer_names is not None else LAYER_NAMES
    rows = [r"\begin{table}[htbp]", r"\centering",
            r"\caption{True, initial and recovered model parameters.}",
            r"\label{tab:taup-inversion-parameters}",
            r"\sisetup{round-mode=places, round-precision=3}",
            r"\begin{tabular}{l S[table-format=1.3] S[table-format=1.3]"
            r" S[table-format=1.3] S[table-format=1.3] r}",
            r"\toprule",
            r"Layer & {$V_P$} & {$V_S$} & {$\rho$} & {$h$} & {Error} \\",
            r"      & {(km/s)} & {(km/s)} & {(g/cm$^3$)} & {(km)} & {(\%)} \\",
            r"\midrule"]
    for k, (tag, m) in enumerate(_models(result)):
        rows.append(rf"\multicolumn{{6}}{{l}}{{\textit{{{tag} model}}}} \\")
        for i in range(m.n_layers):
            name = names[i] if i < len(names) else f"Layer {i}"
            h = m.thickness[i]
            hs = r"\infty" if np.isinf(h) else f"{h:.3f}"
            err = f"{100.0 * _layer_error(m, result.true_model, i):.1e}" if tag == "Recovered" else ""
            rows.append(rf"\quad {name} & {m.alpha[i]:.3f} & {m.beta[i]:.3f} & {m.rho[i]:.3f}"
                        rf" & {hs} & {err} \\")
        rows.append(r"\midrule" if k < 2 else r"\bottomrule")
    rows += [r"\end{tabular}", r"\end{table}"]
    Path(output).write_text("\n".join(rows) + "\n", encoding="utf-8")


def depth_profile(model, param: str, halfspace_ext: float = 2.0):
    """(depths, values) of a step-function profile; the half-space is drawn
    ``halfspace_ext`` km thick."""
    vals = getattr(model, param)
    z, depths, values = 0.0, [], []
    for i in range(model.n_layers):
        h = float(model.thickness[i])
        h = halfspace_ext if np.isinf(h) else h
        depths += [z, z + h]
        values += [float(vals[i])] * 2
        z += h
    return np.array(depths), np.array(values)


def write_model_profiles_tikz(result, output, halfspace_ext: float = 2.0) -> None:
    r"""TikZ figure: Vp, Vs and rho against depth for the three models (needs
    \usepackage{tikz})."""
    panels = (("alpha", r"$V_P$ (km\,s$^{-1}$)"), ("beta", r"$V_S$ (km\,s$^{-1}$)"),
              ("rho", r"$\rho$ (g\,cm$^{-3}$)"))
    styles = ("black, line width=1.2pt", "blue!70!black, dashed, line width=0.8pt",
              "red!70!black, line width=1.0pt")
    zmax = max(depth_profile(m, "alpha", halfspace_ext)[0][-1] for _, m in _models(result))
    width, height, gap = 3.8, 7.0, 2.0
    ys = height / zmax
    out = [r"% needs \usepackage{tikz}", r"\begin{figure}[htbp]", r"\centering",
           r"\begin{tikzpicture}"]
    for k, (param, label) in enumerate(panels):
        x0 = k * (width + gap)
        vals = np.concatenate([depth_profile(m, param, halfspace_ext)[1] for _, m in _models(result)])
        lo, hi = float(vals.min()), float(vals.max())
        pad = 0.1 * (hi - lo) if hi > lo else 0.5
        lo, hi = max(0.0, lo - pad), hi + pad
        xs = width / (hi - lo)
        out.append(rf"\draw[thin] ({x0:.3f},0) rectangle ({x0 + width:.3f},{-height:.3f});")
        out.append(rf"\node[below] at ({x0 + width / 2:.3f},{-height - 0.4:.3f}) {{{label}}};")
        for t in np.linspace(lo, hi, 4):
            xt = x0 + (t - lo) * xs
            out.append(rf"\draw ({xt:.3f},{-height:.3f}) -- ++(0,0.1) "
                       rf"node[below=2pt, font=\scriptsize] {{{t:.2f}}};")
        if k == 0:
            for zt in np.linspace(0.0, zmax, 5):
                out.append(rf"\draw ({x0:.3f},{-zt * ys:.3f}) -- ++(0.1,0) "
                           rf"node[left=4pt, font=\scriptsize] {{{zt:.1f}}};")
            out.append(rf"\node[rotate=90] at ({x0 - 1.1:.3f},{-height / 2:.3f}) {{Depth (km)}};")
        for (tag, m), style in zip(_models(result), styles):
            d, v = depth_profile(m, param, halfspace_ext)
            pts = " -- ".join(f"({x0 + (vv - lo) * xs:.3f},{-dd * ys:.3f})" for dd, vv in zip(d, v))
            out.append(rf"\draw[{style}] {pts};")
    lx = 3 * width + 2 * gap + 0.3
    for j, ((tag, _), style) in enumerate(zip(_models(result), styles)):
        y = -0.4 - 0.5 * j
        out.append(rf"\draw[{style}] ({lx:.3f},{y:.3f}) -- ++(0.6,0) node[right, black] {{{tag}}};")
    out += [r"\end{tikzpicture}",
            r"\caption{Depth profiles of the true, initial and recovered models.}",
            r"\label{fig:taup-model-profiles}", r"\end{figure}"]
    Path(output).write_text("\n".join(out) + "\n", encoding="utf-8")


def plot_taup_traces(result, output, p_values: list[float] | None = None, T: float = 64.0,
                     nw: int = 1024, t_max: float = 15.0) -> None:
    """Tau-p traces of the true (black) and recovered (red) models side by
    side, and their residuals (blue, each trace scaled to its own peak)."""
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    from .taup_inversion import compute_taup_traces

    if p_values is None:
        p_values = [0.10, 0.20, 0.30, 0.40, 0.50, 0.60]
    time, tr_true = compute_taup_traces(result.true_model, p_values, T, nw)
    _, tr_rec = compute_taup_traces(result.recovered_model, p_values, T, nw)
    keep = time <= t_max
    t = time[keep]
    peak = max(float(np.abs(tr_true[p][keep]).max()) for p in p_values) or 1.0
    step = (p_values[1] - p_values[0]) if len(p_values) > 1 else 0.1
    fig, (ax, axr) = plt.subplots(1, 2, figsize=(14, 8), sharey=True)
    for p in p_values:
        a = tr_true[p][keep] / peak * 0.8 * step
        b = tr_rec[p][keep] / peak * 0.8 * step
        ax.plot(p + a, t, "k-", linewidth=0.6)
        ax.fill_betweenx(t, p, p + a, where=a > 0, color="black", alpha=0.8)
        ax.plot(p + b, t, "r-", linewidth=0.6)
        r = tr_rec[p][keep] - tr_true[p][keep]
        rp = float(np.abs(r).max())
        axr.plot(p + (r / rp * 0.8 * step if rp > 0 else r), t, "b-", linewidth=0.6)
    for a_, title in ((ax, "True (black) and recovered (red)"), (axr, "Residual (self-normalised)")):
        a_.set_xlabel("Slowness p (s/km)")
        a_.set_title(title)
    ax.set_ylabel("Time (s)")
    ax.invert_yaxis()
    fig.tight_layout()
    fig.savefig(str(output), bbox_inches="tight")
    plt.close(fig)


def plot_convergence_curves(result, output) -> None:
    """Misfit, gradient norm and relative parameter error per iteration."""
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    it = np.arange(1, len(result.misfit_history) + 1)
    fig, axes = plt.subplots(1, 3, figsize=(15, 5))
    for ax, hist, style, ylabel, title in (
            (axes[0], result.misfit_history, "b-o", r"Misfit $\chi$", "Misfit"),
            (axes[1], result.grad_norm_history, "r-o", r"$\|\nabla\chi\|$", "Gradient norm"),
            (axes[2], result.param_error_history, "k-o",
             r"$\|m - m_{true}\| / \|m_{true}\|$", "Relative parameter error")):
        ax.semilogy(it, hist, style, markersize=3)
        ax.set_xlabel("Iteration")
        ax.set_ylabel(ylabel)
        ax.set_title(title)
        ax.grid(True, alpha=0.3)
    fig.tight_layout()
    fig.savefig(str(output), bbox_inches="tight")
    plt.close(fig)
