"""Shared test helpers: the FP64 observable reductions computed independently
in numpy from integer measurement records (the checker of sg_get_results)."""
from __future__ import annotations

import math

import numpy as np


def geometric_betas(tmin, tmax, n):
    if n == 1:
        return np.array([1.0 / tmin])
    f = np.arange(n) / (n - 1)
    return 1.0 / (tmin * (tmax / tmin) ** f)


def results_from_records(records, L, R, H, n_samples, betas):
    """records: list of int32 arrays [group][temp][lane][field].
    Returns [T][11] with the columns of sg.RES_COLUMNS.

    Integer observables are totalled exactly (Python integers) over
    measurements, samples and replicas and divided once; the k_min overlap is
    formed per lane and measurement as an FP64 sum over planes c = 0..L-1 in
    order (elementwise, no fused multiply-add), totalled per lane over the
    measurements in order and over samples in ascending order.  That is the
    arithmetic the product specifies, so the comparison can be exact."""
    N = L ** 3
    Nd = float(N)
    T = len(betas)
    cw = [math.cos(2.0 * math.pi * c / L) for c in range(L)]
    sw = [math.sin(2.0 * math.pi * c / L) for c in range(L)]
    nm = len(records)
    sU = [0] * T; sK = [0] * T; sAQ = [0] * T; sQ2 = [0] * T; sQ4 = [0] * T
    sq1 = np.zeros((n_samples, T))
    for rec in records:
        G = rec.shape[0]
        r = rec.transpose(0, 2, 1, 3).reshape(G * 64, T, -1)[:n_samples].astype(np.int64)
        for t in range(T):
            sU[t] += int(r[:, t, :R].sum())
            sK[t] += int(r[:, t, R:2 * R].sum())
        if R >= 2:
            Q = N - 2 * r[:, :, 2 * R]
            for t in range(T):
                q = [int(v) for v in Q[:, t]]
                sAQ[t] += sum(abs(v) for v in q)
                sQ2[t] += sum(v * v for v in q)
                sQ4[t] += sum(v ** 4 for v in q)
            q1 = np.zeros((n_samples, T))
            for d in range(3):
                W = r[:, :, 2 * R + 1 + d * L: 2 * R + 1 + (d + 1) * L].astype(np.float64)
                re = np.zeros((n_samples, T))
                im = np.zeros((n_samples, T))
                for c in range(L):
                    re = re + cw[c] * W[:, :, c]
                    im = im + sw[c] * W[:, :, c]
                re = re * (-2.0 / N)
                im = im * (-2.0 / N)
                q1 = q1 + (re * re + im * im)
            sq1 = sq1 + q1 / 3.0
    out = np.zeros((T, 11))
    nS = float(nm * n_samples)
    nRS = float(nm * R * n_samples)
    for t in range(T):
        out[t, 0] = betas[t]
        out[t, 10] = nm
        out[t, 6] = out[t, 9] = np.nan
        if nm == 0:
            continue
        U = float(sU[t]) / nRS
        K = float(sK[t]) / nRS
        M = 2.0 * K - Nd
        q2 = float(sQ2[t]) / nS / (Nd * Nd)
        q4 = float(sQ4[t]) / nS / (Nd * Nd * Nd * Nd)
        c = 0.0
        for s_ in range(n_samples):
            c += float(sq1[s_, t])
        c1 = c / nS
        out[t, 1] = ((2.0 * U - 3.0 * Nd) - H * M) / Nd
        out[t, 2] = M / Nd
        out[t, 3] = float(sAQ[t]) / nS / Nd
        out[t, 4] = q2
        out[t, 5] = q4
        out[t, 6] = 0.5 * (3.0 - q4 / (q2 * q2)) if q2 > 0 else np.nan
        out[t, 7] = q2
        out[t, 8] = c1
        out[t, 9] = (math.sqrt(q2 / c1 - 1.0) / (2.0 * math.sin(math.pi / L)) / L
                     if c1 > 0 and q2 / c1 > 1.0 else np.nan)
    return out
