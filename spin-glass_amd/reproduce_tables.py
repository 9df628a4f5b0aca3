"""Runs the reference's full protocol on this engine and writes res.dat.

The reference's published tables (Simulations/<L>/<H>-res.dat) come from 16
invocations of `csg` of 262,144 sweeps each: 2048 samples, 2 copies, its 24-
temperature ladder, PT every 16 sweeps, an obs record every 1024 sweeps, and
analyze-all.py over the second half of the records.  This script runs the
same protocol in one call, streams the records of the second half into the
per-sample time means (instead of writing an 8.8 GB obs.bin), applies the
same disorder averages (analysis.estimators_from_time_means) and writes
<out>/<LL>/<H>-res.dat.  With --compare it prints z-scores against a table.

Usage: python reproduce_tables.py --L 16 --H 0 [--samples 2048]
       [--sweeps 4194304] [--out DIR] [--compare path/to/res.dat]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

import analysis
import sg


def run(L, H, samples, sweeps, out_dir, seed0=0x5EED0201, progress=True):
    sim = sg.Sim(L=L, n_samples=samples, n_replicas=2, betas=sg.REF_BETAS, H=H, pt_every=16,
                 meas_every=1024, form=sg.SG_FORM_BATCHED, coupling_seed=seed0,
                 spin_seed=seed0 + 1)
    sim.keep_records = False
    sim.init_lattice()
    n_rec = sweeps // 1024
    first_kept = n_rec - n_rec // 2
    T = 24
    acc = dict(q0=np.zeros((samples, T)), re=np.zeros((samples, T, 3)),
               im=np.zeros((samples, T, 3)), e=np.zeros((samples, T)), m=np.zeros((samples, T)))
    count = [0, 0]
    t0 = time.time()

    def sink(sweep, rec):
        i = count[0]
        count[0] += 1
        if i >= first_kept:
            o = sg.obs_records(rec, L, 2, H, samples)
            acc["q0"] += o["q_0"].astype(np.float64) ** 2
            acc["re"] += o["q_1_r"].astype(np.float64) ** 2
            acc["im"] += o["q_1_i"].astype(np.float64) ** 2
            acc["e"] += o["e"].astype(np.float64).mean(axis=-1)
            acc["m"] += o["m"].astype(np.float64).mean(axis=-1)
            count[1] += 1
        if progress and count[0] % 256 == 0:
            print(f"  {count[0]}/{n_rec} records, {time.time() - t0:.0f} s", flush=True)

    chunk = 1024 * 256
    for done in range(0, sweeps, chunk):
        sim.run(min(chunk, sweeps - done), sink=sink)
    n = count[1]
    means = [acc[k] / n for k in ("q0", "re", "im", "e", "m")]
    cols = analysis.estimators_from_time_means(*means, L)
    sub = os.path.join(out_dir, f"{L:02}")
    os.makedirs(sub, exist_ok=True)
    path = os.path.join(sub, f"{H:06.4f}-res.dat")
    analysis.write_res(path, analysis.ladder(T), cols, 2 * n * 1024)
    if samples % 64 == 0 and samples >= 128:
        analysis.write_jk(analysis.jk_path(path), analysis.ladder(T),
                          analysis.group_jackknife_errors(means, L))
    ms, flips = sim.timing()
    return path, flips / (ms * 1e6)


def compare(ours, ref):
    """z = (ours - ref) / hypot(sigma_ours, sigma_ref) per temperature for
    chi0, xi_L/L, e, m, with the table's own errors and, where the jackknife
    file exists, with the lane-group jackknife errors for ours (zb)."""
    a, b = np.loadtxt(ours), np.loadtxt(ref)
    jkp = analysis.jk_path(ours)
    jk = np.loadtxt(jkp) if os.path.exists(jkp) else None
    print(" beta      z(chi0)  z(xi/L)  z(e)    z(m)   |  zb(chi0) zb(xi/L) zb(e)   zb(m)")
    z = np.zeros((len(a), 4))
    zb = np.zeros((len(a), 4))
    for t in range(len(a)):
        for k, c in enumerate((1, 3, 5, 7)):
            s = np.hypot(a[t, c + 1], b[t, c + 1])
            z[t, k] = (a[t, c] - b[t, c]) / s if s > 0 else 0.0
            if jk is not None:
                sb = np.hypot(max(jk[t, 1 + k], a[t, c + 1]), b[t, c + 1])
                zb[t, k] = (a[t, c] - b[t, c]) / sb if sb > 0 else 0.0
        line = f" {a[t, 0]:.4f}  " + "  ".join(f"{v:+6.2f}" for v in z[t])
        if jk is not None:
            line += "  |  " + "  ".join(f"{v:+6.2f}" for v in zb[t])
        print(line)
    print(f"max |z| = {np.abs(z).max():.2f}" +
          (f", max |zb| = {np.abs(zb).max():.2f}" if jk is not None else ""))
    return z, zb


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=16)
    ap.add_argument("--H", type=float, default=0.0)
    ap.add_argument("--samples", type=int, default=2048)
    ap.add_argument("--sweeps", type=int, default=4194304)
    ap.add_argument("--out", default="reproduced")
    ap.add_argument("--compare", default=None)
    a = ap.parse_args()
    path, rate = run(a.L, a.H, a.samples, a.sweeps, a.out)
    print(f"wrote {path} ({rate:.0f} spin flips/ns over the run)")
    if a.compare:
        compare(path, a.compare)


if __name__ == "__main__":
    sys.exit(main())
