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
    Returns [T][11] with the columns of sg.RES_COLUMNS."""
    N = L ** 3
    T = len(betas)
    e_s = np.zeros((n_samples, T)); m_s = np.zeros((n_samples, T))
    aq_s = np.zeros((n_samples, T)); q1_s = np.zeros((n_samples, T))
    Q2_s = [[0] * T for _ in range(n_samples)]
    Q4_s = [[0] * T for _ in range(n_samples)]
    cw = np.array([math.cos(2.0 * math.pi * c / L) for c in range(L)])
    sw = np.array([math.sin(2.0 * math.pi * c / L) for c in range(L)])
    nm = len(records)
    for rec in records:
        G = rec.shape[0]
        r = rec.transpose(0, 2, 1, 3).reshape(G * 64, T, -1)[:n_samples].astype(np.int64)
        for s in range(n_samples):
            for t in range(T):
                f = r[s, t]
                se = sm = 0.0
                for rho in range(R):
                    M = float(2 * f[R + rho] - N)
                    se += (float(2 * f[rho] - 3 * N) - H * M) / N
                    sm += M / N
                e_s[s, t] += se / R
                m_s[s, t] += sm / R
                if R >= 2:
                    Q = N - 2 * int(f[2 * R])
                    Q2_s[s][t] += Q * Q
                    Q4_s[s][t] += Q ** 4
                    aq_s[s, t] += abs(Q) / N
                    q1 = 0.0
                    for d in range(3):
                        W = f[2 * R + 1 + d * L: 2 * R + 1 + (d + 1) * L].astype(np.float64)
                        re = -2.0 / N * float(cw @ W)
                        im = -2.0 / N * float(sw @ W)
                        q1 += re * re + im * im
                    q1_s[s, t] += q1 / 3.0
    out = np.zeros((T, 11))
    for t in range(T):
        e = np.mean(e_s[:, t] / nm)
        m = np.mean(m_s[:, t] / nm)
        aq = np.mean(aq_s[:, t] / nm)
        q2 = np.mean([Q2_s[s][t] / nm / N ** 2 for s in range(n_samples)])
        q4 = np.mean([Q4_s[s][t] / nm / N ** 4 for s in range(n_samples)])
        c1 = np.mean(q1_s[:, t] / nm)
        out[t] = [betas[t], e, m, aq, q2, q4,
                  0.5 * (3 - q4 / q2 ** 2) if q2 > 0 else np.nan, q2, c1,
                  math.sqrt(q2 / c1 - 1) / (2 * math.sin(math.pi / L)) / L
                  if c1 > 0 and q2 / c1 > 1 else np.nan, nm]
    return out
