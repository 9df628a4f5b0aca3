"""Figures of the res.dat tables: chi(0), xi_L/L, e and m against beta.

The same figure set as the reference's make-plots.py (make-plots.py:137-260):
L = 16 at H = 0 ("0-*"), L = 16 at H = 0, 1/16, 2/16, 3/16 ("H-*", and m/H
against the paramagnetic line m/H = beta, "H-mH"), and H = 0 at L = 16, 14, 12
("L-*"), each a curve with a band of one standard error, beta on a log2 axis.
Written without LaTeX or pandas (matplotlib's own mathtext, numpy).  With
`overlay`, a second set of tables (for example the reference's published ones
beside this engine's) is drawn dashed in the same colours, so the two can be
compared curve by curve.

Usage: python plots.py SIM_DIR [--out PLOTS_DIR] [--overlay OTHER_SIM_DIR]
       [--format pdf|png]
"""
from __future__ import annotations

import argparse
import os

import matplotlib

matplotlib.use("Agg")
from matplotlib import pyplot as plt  # noqa: E402
import numpy as np  # noqa: E402

COLOURS = ["#169f62", "#194bb2", "#a31cc5", "#d81e2c"]
# res.dat columns: beta, then (value, error) of chi(0), xi_L/L, e, m
VARS = [("chi_0", 1, r"$\chi(0)$"), ("xi_LL", 3, r"$\xi_L/L$"), ("e", 5, r"$e$"),
        ("m", 7, r"$m$")]


def res_path(sim_dir: str, L: int, H: float) -> str:
    return os.path.join(sim_dir, f"{L:02}", f"{H:06.4f}-res.dat")


def read_res(sim_dir: str, L: int, H: float) -> np.ndarray | None:
    path = res_path(sim_dir, L, H)
    return np.loadtxt(path, comments="#") if os.path.exists(path) else None


def _band(ax, t, col, colour, label, dashed=False, scale=1.0):
    x, y, e = t[:, 0], t[:, col] * scale, t[:, col + 1] * scale
    ax.plot(x, y, color=colour, label=label, linestyle="--" if dashed else "-",
            linewidth=1.0)
    if not dashed:
        ax.fill_between(x, y - e, y + e, color=colour, alpha=0.5, linewidth=0.0)


def _axes(n=4):
    figs, axs = [], []
    for _ in range(n):
        f, a = plt.subplots(figsize=(12 / 2.54, 8 / 2.54), constrained_layout=True)
        figs.append(f)
        axs.append(a)
    for a, (_, _, name) in zip(axs, VARS):
        a.set_xscale("log", base=2)
        a.set_xlabel(r"$\beta$")
        a.set_ylabel(name)
        if name != r"$e$":
            a.axhline(0.0, color="#000000", alpha=0.25, linestyle="--", linewidth=1.0)
    return figs, axs


def _save(figs, names, out_dir, fmt):
    paths = []
    for f, n in zip(figs, names):
        p = os.path.join(out_dir, f"{n}.{fmt}")
        f.savefig(p)
        plt.close(f)
        paths.append(p)
    return paths


def _series(axs, sim_dir, overlay, cases, label_fmt):
    drawn = 0
    for i, (L, H) in enumerate(cases):
        for d, dashed in ((sim_dir, False), (overlay, True)):
            if d is None:
                continue
            t = read_res(d, L, H)
            if t is None:
                continue
            label = None if dashed else label_fmt(L, H)
            for a, (_, col, _) in zip(axs, VARS):
                _band(a, t, col, COLOURS[i % len(COLOURS)], label, dashed)
            drawn += not dashed
    return drawn


def make_plots(sim_dir: str, out_dir: str, overlay: str | None = None, fmt: str = "pdf"):
    """Writes every figure whose tables exist under sim_dir; returns the paths."""
    os.makedirs(out_dir, exist_ok=True)
    paths = []
    names = [v[0] for v in VARS]

    figs, axs = _axes()
    if _series(axs, sim_dir, overlay, [(16, 0.0)], lambda L, H: None):
        paths += _save(figs, [f"0-{n}" for n in names], out_dir, fmt)
    else:
        for f in figs:
            plt.close(f)

    hs = [(16, i / 16) for i in range(4)]
    figs, axs = _axes()
    if _series(axs, sim_dir, overlay, hs, lambda L, H: f"$H$ = {H:06.4f}"):
        for a in axs:
            a.legend(frameon=False)
        paths += _save(figs, [f"H-{n}" for n in names], out_dir, fmt)
    else:
        for f in figs:
            plt.close(f)

    # m/H against the paramagnetic line m/H = beta
    f, a = plt.subplots(figsize=(12 / 2.54, 8 / 2.54), constrained_layout=True)
    a.set_xscale("log", base=2)
    a.set_xlabel(r"$\beta$")
    a.set_ylabel(r"$m/H$")
    drawn = 0
    for i, (L, H) in enumerate(hs):
        if H == 0.0:
            continue
        for d, dashed in ((sim_dir, False), (overlay, True)):
            t = read_res(d, L, H) if d else None
            if t is None:
                continue
            _band(a, t, 7, COLOURS[i], None if dashed else f"$H$ = {H:06.4f}", dashed,
                  scale=1.0 / H)
            drawn += not dashed
    if drawn:
        b = np.geomspace(0.125, 2.0, 50)
        a.plot(b, b, color=COLOURS[0], alpha=0.5, linestyle=":", linewidth=1.0,
               label=r"$m/H = \beta$")
        a.legend(frameon=False)
        paths += _save([f], ["H-mH"], out_dir, fmt)
    else:
        plt.close(f)

    figs, axs = _axes()
    ls = [(16, 0.0), (14, 0.0), (12, 0.0)]
    if _series(axs, sim_dir, overlay, ls, lambda L, H: f"$L$ = {L:02}"):
        for a in axs:
            a.legend(frameon=False)
        paths += _save(figs, [f"L-{n}" for n in names], out_dir, fmt)
    else:
        for f in figs:
            plt.close(f)
    return paths


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("sim_dir")
    ap.add_argument("--out", default="Plots")
    ap.add_argument("--overlay", default=None)
    ap.add_argument("--format", default="pdf", choices=["pdf", "png"])
    a = ap.parse_args()
    for p in make_plots(a.sim_dir, a.out, a.overlay, a.format):
        print(p)


if __name__ == "__main__":
    main()
