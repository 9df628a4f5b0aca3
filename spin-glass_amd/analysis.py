"""obs.bin -> res.dat, the reference's offline estimators.

Follows the reference's analyze-all.py (analyze_obs, lines 29-88): the first
half of the file is discarded as thermalisation, time averages are taken per
sample, then disorder averages with sigma = sqrt(var_samples / NDIS):
  chi(0)  = [<q0^2>]                                   (lines 38-41)
  chi(k)  = mean over x,y,z of [<Re q1^2>] + [<Im q1^2>] (lines 43-50)
  xi_L/L  = sqrt(chi(0)/chi(k) - 1) / (2 sin(pi/L)) / L  (lines 52-61)
  e, m    = mean over time and copies, then samples      (lines 63-71)
and writes the same file: a "# tot_MC_steps : {:,}" header and one row of
nine "%16.12f" columns per temperature (lines 73-88).  The arithmetic runs on
the float32 records with the same numpy reductions, so the same obs.bin gives
the same res.dat.

Error bars.  The 64 samples of one engine word share the random number of
every site update, so their thermal fluctuations are correlated and the
per-sample formula above understates the error where thermal noise
dominates.  With --jackknife the script also writes <H>-res.jk.dat: the
errors of chi(0), xi_L/L, e and m from a leave-one-group-out jackknife over
the 64-sample groups (the independent units of this engine).

Usage: python analysis.py <sim_dir> <L> <H> [n_samples] [n_temps] [--jackknife]
"""
from __future__ import annotations

import os
import sys

import numpy as np

OBS_DTYPE = np.dtype([("e", "f4", (2,)), ("m", "f4", (2,)), ("q_0", "f4"),
                      ("q_1_r", "f4", (3,)), ("q_1_i", "f4", (3,))])


def ladder(n_temps: int = 24) -> np.ndarray:
    # the reference ladder as analyze-all.py:24 writes it
    return np.logspace(1, -3, num=n_temps, base=2.0)


def time_means(data: np.ndarray):
    """Per-sample time means of structured obs records [measurement][sample]
    [temp]: <q0^2>, <Re q1^2>, <Im q1^2> ([s][t][3]), <e> and <m> averaged
    over copies ([s][t])."""
    chi_0 = np.average(data["q_0"] ** 2, axis=0)
    re = np.average(data["q_1_r"] ** 2, axis=0)
    im = np.average(data["q_1_i"] ** 2, axis=0)
    e = np.average(data["e"], axis=(0, 3))
    m = np.average(data["m"], axis=(0, 3))
    return chi_0, re, im, e, m


def estimators(data: np.ndarray, L: int):
    """data: structured obs records [measurement][sample][temp]."""
    return estimators_from_time_means(*time_means(data), L)


def estimators_from_time_means(chi_0, re, im, e, m, L: int):
    """Disorder averages from per-sample time means: chi_0 = <q0^2> [s][t],
    re/im = <Re/Im q1^2> [s][t][3], e/m = <e>/<m> averaged over copies
    [s][t].  Returns chi0, sigma, xi_L/L, sigma, e, sigma, m, sigma ([t])."""
    n_dis = chi_0.shape[0]
    chi_0_m = np.average(chi_0, axis=0)
    chi_0_v = np.var(chi_0, axis=0)
    chi_0_e = np.sqrt(chi_0_v / n_dis)

    re_m, re_v = np.average(re, axis=0), np.var(re, axis=0)
    im_m, im_v = np.average(im, axis=0), np.var(im, axis=0)
    chi_1_m = np.average(re_m + im_m, axis=1)
    chi_1_v = np.average(re_v + im_v, axis=1)

    rel_0 = chi_0_v / chi_0_m ** 2
    rel_1 = chi_1_v / chi_1_m ** 2
    ratio = chi_0_m / chi_1_m
    ratio_v = ratio ** 2 * (rel_0 + rel_1)
    xi = (np.sqrt(ratio - 1) / (2 * np.sin(np.pi / L))) / L
    xi_v = (xi / (2 * (ratio - 1))) ** 2 * ratio_v
    xi_e = np.sqrt(xi_v / n_dis)

    e_m, e_e = np.average(e, axis=0), np.sqrt(np.var(e, axis=0) / n_dis)
    m_m, m_e = np.average(m, axis=0), np.sqrt(np.var(m, axis=0) / n_dis)
    return chi_0_m, chi_0_e, xi, xi_e, e_m, e_e, m_m, m_e


def group_jackknife_errors(means, L: int, block: int = 64):
    """Errors of chi0, xi_L/L, e, m ([4][T]) from per-sample time means
    (time_means' order), with the `block`-sample groups as the independent
    units: a leave-one-group-out jackknife, sqrt((G-1)/G sum_g (th_g -
    mean th)^2) over the G groups.  For the linear columns (chi0, e, m) this
    is the standard error of the G group means."""
    n = means[0].shape[0]
    G = n // block
    if G < 2:
        raise ValueError("the jackknife needs at least two groups")
    th = []
    for g in range(G):
        keep = np.r_[0:g * block, (g + 1) * block:G * block]
        c = estimators_from_time_means(*[a[keep] for a in means], L)
        th.append([c[0], c[2], c[4], c[6]])
    th = np.asarray(th)
    return np.sqrt((G - 1) / G * ((th - th.mean(axis=0)) ** 2).sum(axis=0))


def jk_path(res_path: str) -> str:
    return res_path[:-len("-res.dat")] + "-res.jk.dat"


def write_jk(path: str, beta, jk) -> None:
    np.savetxt(path, np.column_stack([beta, jk.T]), fmt="%16.12f",
               header="beta, sigma of chi0, xi_L/L, e, m: leave-one-lane-group-out jackknife")


def analyze_obs(sim_dir: str, L: int, H: float, n_samples: int = 2048, n_temps: int = 24,
                betas: np.ndarray | None = None, sweeps_per_record: int = 1024,
                jackknife: bool = False) -> str:
    """Reads <sim_dir>/<LL>/<H:06.4f>-obs.bin and writes ...-res.dat next to
    it (and ...-res.jk.dat with `jackknife`); returns the res.dat path."""
    sub = os.path.join(sim_dir, f"{L:02}")
    path = os.path.join(sub, f"{H:06.4f}-obs.bin")
    therm = os.path.getsize(path) // 2
    data = np.memmap(path, dtype=OBS_DTYPE, mode="r", offset=therm)
    data = data.reshape(-1, n_samples, n_temps)
    means = time_means(data)
    cols = estimators_from_time_means(*means, L)
    beta = ladder(n_temps) if betas is None else np.asarray(betas)
    out = os.path.join(sub, f"{H:06.4f}-res.dat")
    write_res(out, beta, cols, 2 * data.shape[0] * sweeps_per_record)
    if jackknife:
        write_jk(jk_path(out), beta, group_jackknife_errors(means, L))
    return out


def write_res(path: str, beta, cols, tot_sweeps: int) -> None:
    with open(path, "w") as fh:
        fh.write(f"# tot_MC_steps : {tot_sweeps:,}\n")
        for t in range(len(beta)):
            fh.write(f"{beta[t]:16.12f}")
            for c in cols:
                fh.write(f"{c[t]:16.12f}")
            fh.write("\n")


if __name__ == "__main__":
    a = [v for v in sys.argv[1:] if v != "--jackknife"]
    if len(a) < 3:
        sys.exit(__doc__)
    print(analyze_obs(a[0], int(a[1]), float(a[2]), *(int(v) for v in a[3:5]),
                      jackknife="--jackknife" in sys.argv[1:]))
