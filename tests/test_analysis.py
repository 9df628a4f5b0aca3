"""The res.dat writer against hand-computed values (the reference's
analyze-all.py formulas evaluated in plain Python on a tiny input), and the
obs-record conversion against direct sums over spins."""
import math
import os

import numpy as np

import analysis
import sg


def _tiny(tmp_path):
    # 4 measurements x 2 samples x 2 temperatures; the first 2 are discarded
    d = np.zeros((4, 2, 2), dtype=analysis.OBS_DTYPE)
    rng = np.random.default_rng(5)
    for f, shape in (("e", (2,)), ("m", (2,)), ("q_1_r", (3,)), ("q_1_i", (3,))):
        d[f] = rng.uniform(-0.5, 0.5, size=(4, 2, 2) + shape).astype(np.float32)
    d["q_0"] = rng.uniform(-0.9, 0.9, size=(4, 2, 2)).astype(np.float32)
    d["q_0"][:2] = 99.0       # must be cut away as thermalisation
    sub = tmp_path / "08"
    sub.mkdir()
    (sub / "0.1250-obs.bin").write_bytes(d.tobytes())
    return d[2:].astype(analysis.OBS_DTYPE)


def test_res_dat_matches_hand_computation(tmp_path):
    kept = _tiny(tmp_path)
    out = analysis.analyze_obs(str(tmp_path), 8, 0.125, n_samples=2, n_temps=2,
                               betas=[1.5, 0.5])
    lines = open(out).read().splitlines()
    assert lines[0] == "# tot_MC_steps : 4,096"
    rows = np.array([[float(line[i:i + 16]) for i in range(0, 9 * 16, 16)] for line in lines[1:]])
    L = 8
    for t in range(2):
        def col(field, comp=None):
            v = kept[field][:, :, t] if comp is None else kept[field][:, :, t, comp]
            return [[float(v[m, s]) for m in range(2)] for s in range(2)]

        def mean(a):
            return sum(a) / len(a)

        def var(a):
            mu = mean(a)
            return sum((x - mu) ** 2 for x in a) / len(a)

        chi0_s = [mean([q * q for q in col("q_0")[s]]) for s in range(2)]
        chi0, chi0_v = mean(chi0_s), var(chi0_s)
        re_s = [[mean([q * q for q in col("q_1_r", d)[s]]) for s in range(2)] for d in range(3)]
        im_s = [[mean([q * q for q in col("q_1_i", d)[s]]) for s in range(2)] for d in range(3)]
        chi1 = mean([mean(re_s[d]) + mean(im_s[d]) for d in range(3)])
        chi1_v = mean([var(re_s[d]) + var(im_s[d]) for d in range(3)])
        ratio = chi0 / chi1
        ratio_v = ratio ** 2 * (chi0_v / chi0 ** 2 + chi1_v / chi1 ** 2)
        xi = math.sqrt(ratio - 1) / (2 * math.sin(math.pi / L)) / L if ratio > 1 else float("nan")
        xi_e = math.sqrt((xi / (2 * (ratio - 1))) ** 2 * ratio_v / 2) if ratio > 1 else float("nan")
        e_s = [mean([float(kept["e"][m, s, t, c]) for m in range(2) for c in range(2)])
               for s in range(2)]
        m_s = [mean([float(kept["m"][m, s, t, c]) for m in range(2) for c in range(2)])
               for s in range(2)]
        expect = [[1.5, 0.5][t], chi0, math.sqrt(chi0_v / 2), xi, xi_e, mean(e_s),
                  math.sqrt(var(e_s) / 2), mean(m_s), math.sqrt(var(m_s) / 2)]
        for k, (a, b) in enumerate(zip(rows[t], expect)):
            if math.isnan(b):
                assert math.isnan(a)
            else:
                assert abs(a - b) <= 2e-6 * max(1.0, abs(b)) + 1e-11, (t, k, a, b)


def test_obs_records_from_integer_fields():
    """sg.obs_records turns U, K, W and plane counts into the reference's
    float32 e, m, q_0, q_1 (eamsim.cu:320-325, 451-511)."""
    L, R, H, N = 4, 2, 0.25, 64
    rng = np.random.default_rng(9)
    s = rng.choice([-1, 1], size=(R, L, L, L))     # [rho][z][y][x]
    J = rng.choice([-1, 1], size=(3, L, L, L))
    rec = np.zeros((1, 1, 64, 2 * R + 1 + 3 * L), dtype=np.int32)
    for rho in range(R):
        E = 0
        for d, ax in enumerate((2, 1, 0)):
            E -= (J[d] * s[rho] * np.roll(s[rho], -1, axis=ax)).sum()
        U = (E + 3 * N) // 2
        rec[0, 0, 0, rho] = U
        rec[0, 0, 0, R + rho] = (s[rho] > 0).sum()
    w = (s[0] != s[1]).astype(int)
    rec[0, 0, 0, 2 * R] = w.sum()
    rec[0, 0, 0, 2 * R + 1:2 * R + 1 + L] = w.sum(axis=(0, 1))
    rec[0, 0, 0, 2 * R + 1 + L:2 * R + 1 + 2 * L] = w.sum(axis=(0, 2))
    rec[0, 0, 0, 2 * R + 1 + 2 * L:] = w.sum(axis=(1, 2))
    o = sg.obs_records(rec, L, R, H, 1)[0, 0]
    q = s[0] * s[1]
    k = 2 * math.pi / L
    z, y, x = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    for rho in range(R):
        E = -sum((J[d] * s[rho] * np.roll(s[rho], -1, axis=ax)).sum()
                 for d, ax in enumerate((2, 1, 0))) - H * s[rho].sum()
        assert abs(o["e"][rho] - E / N) < 1e-6
        assert abs(o["m"][rho] - s[rho].sum() / N) < 1e-6
    assert abs(o["q_0"] - q.sum() / N) < 1e-6
    for d, c in enumerate((x, y, z)):
        assert abs(o["q_1_r"][d] - (q * np.cos(k * c)).sum() / N) < 1e-6
        assert abs(o["q_1_i"][d] - (q * np.sin(k * c)).sum() / N) < 1e-6


def test_group_jackknife_errors():
    """The lane-group jackknife agrees with the per-sample formula when
    samples are independent and picks up a fluctuation shared by the 64
    samples of a group, which the per-sample formula misses."""
    rng = np.random.default_rng(7)
    S, T, L = 8192, 2, 8
    means = [rng.random((S, T)) + 1.0, rng.random((S, T, 3)) * 0.1 + 0.05,
             rng.random((S, T, 3)) * 0.1 + 0.05, -rng.random((S, T)), rng.random((S, T)) * 0.01]
    jk = analysis.group_jackknife_errors(means, L)
    naive = analysis.estimators_from_time_means(*means, L)
    for k, col in enumerate((1, 3, 5, 7)):
        np.testing.assert_allclose(jk[k], naive[col], rtol=0.35)
    shared = np.repeat(rng.normal(0.0, 0.05, (S // 64, T)), 64, axis=0)
    means[3] = means[3] + shared
    jk = analysis.group_jackknife_errors(means, L)
    naive = analysis.estimators_from_time_means(*means, L)
    expected = math.sqrt((1 / 12) / S + 0.05 ** 2 / (S // 64))
    np.testing.assert_allclose(jk[2], expected, rtol=0.35)
    assert np.all(jk[2] > 1.4 * naive[5])


def test_group_jackknife_of_linear_columns_is_the_group_standard_error():
    """For chi0, e and m (plain means over samples) the leave-one-group-out
    jackknife equals the standard error of the group means,
    sqrt(sum_g (x_g - x)^2 / (G (G - 1))): a closed form that a wrong
    weight, a dropped group or an off-by-one in the blocks would miss."""
    rng = np.random.default_rng(11)
    S, T, L, B = 640, 3, 8, 64
    means = [rng.random((S, T)) + 1.0, rng.random((S, T, 3)) * 0.1 + 0.05,
             rng.random((S, T, 3)) * 0.1 + 0.05, -rng.random((S, T)), rng.random((S, T)) * 0.01]
    jk = analysis.group_jackknife_errors(means, L, block=B)
    G = S // B
    for k, a in ((0, means[0]), (2, means[3]), (3, means[4])):
        xg = a.reshape(G, B, T).mean(axis=1)
        se = np.sqrt(((xg - xg.mean(axis=0)) ** 2).sum(axis=0) / (G * (G - 1)))
        np.testing.assert_allclose(jk[k], se, rtol=1e-10)
    assert np.all(np.isfinite(jk[1])) and np.all(jk[1] > 0)


def test_analyze_obs_writes_the_jackknife_file(tmp_path):
    """analysis.py --jackknife: the res.jk.dat beside res.dat holds, per
    temperature, beta and the jackknife errors computed from the second half
    of obs.bin."""
    rng = np.random.default_rng(5)
    L, S, T, n_rec = 8, 128, 4, 6
    rec = np.zeros((n_rec, S, T), dtype=analysis.OBS_DTYPE)
    for f in ("e", "m"):
        rec[f] = rng.normal(0.0, 0.3, rec[f].shape)
    for f in ("q_1_r", "q_1_i"):                     # chi(k) below chi(0): xi is defined
        rec[f] = rng.normal(0.0, 0.1, rec[f].shape)
    rec["q_0"] = rng.normal(0.0, 0.3, rec["q_0"].shape)
    os.makedirs(tmp_path / "08")
    rec.tofile(tmp_path / "08" / "0.0000-obs.bin")
    out = analysis.analyze_obs(str(tmp_path), L, 0.0, S, T, jackknife=True)
    jk = np.loadtxt(analysis.jk_path(out))
    assert jk.shape == (T, 5)
    np.testing.assert_allclose(jk[:, 0], analysis.ladder(T), rtol=1e-11)
    means = analysis.time_means(rec[n_rec // 2:])
    assert np.all(np.isfinite(jk)) and np.all(jk[:, 1:] > 0)
    np.testing.assert_allclose(jk[:, 1:].T, analysis.group_jackknife_errors(means, L),
                               rtol=1e-9, atol=1e-12)


def test_plots_of_the_reproduced_tables(tmp_path):
    """The figure set of the reference's make-plots.py (L = 16 at H = 0, the
    H series with m/H, the L series), drawn from this engine's reproduced
    tables with the published ones overlaid."""
    import plots
    here = os.path.dirname(os.path.abspath(__file__))
    paths = plots.make_plots(os.path.join(here, "golden", "engine"), str(tmp_path),
                             overlay=os.path.join(here, "golden", "res"), fmt="png")
    names = sorted(os.path.basename(p) for p in paths)
    expect = sorted([f"{k}-{v}.png" for k in ("0", "H", "L") for v in ("chi_0", "xi_LL", "e", "m")]
                    + ["H-mH.png"])
    assert names == expect
    assert all(os.path.getsize(p) > 1000 for p in paths)
