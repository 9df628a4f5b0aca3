"""Pins the CPU oracle (oracle/sg_oracle.c) against known answers, closed
forms, invariants and exact enumeration.  Each check is chosen so that a
plausible mistake (wrong sign of dE, swapped threshold index, dropped bond,
wrong PT acceptance sign, wrong Fourier sign) fails it."""
import itertools
import math

import numpy as np
import pytest

import oracle
from helpers import geometric_betas

TWO32 = 2 ** 32


def bits(words):
    """uint64 array (...) -> bool array (..., 64), bit l at index l."""
    w = np.ascontiguousarray(words, dtype=np.uint64)
    b = np.unpackbits(w.view(np.uint8).reshape(*w.shape, 8), axis=-1, bitorder="little")
    return b.astype(bool)


def site(L, x, y, z):
    return x + L * (y + L * z)


def bond_energy_numpy(L, S, J):
    """Per-lane unsatisfied-bond count U and up count K, by direct loops."""
    sb, jb = bits(S), bits(J)
    U = np.zeros(64, dtype=np.int64)
    for z, y, x in itertools.product(range(L), repeat=3):
        i = site(L, x, y, z)
        for d, j in enumerate((site(L, (x + 1) % L, y, z), site(L, x, (y + 1) % L, z),
                               site(L, x, y, (z + 1) % L))):
            U += sb[i] ^ sb[j] ^ jb[d, i]
    return U, sb.sum(axis=0)


# ----------------------------------------------------------------- Philox
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expect):
    # Random123 known-answer vectors for philox4x32-10 (Salmon et al., SC'11)
    assert [int(v) for v in oracle.philox(ctr, key)] == expect


# -------------------------------------------------------- deterministic exp
def test_det_exp_matches_libm():
    xs = np.concatenate([np.linspace(-40.0, 0.0, 20001), -np.logspace(-12, 1.6, 2001)])
    worst = max(abs(oracle.det_exp(x) - math.exp(x)) / math.exp(x) for x in xs)
    assert worst < 4e-16
    assert oracle.det_exp(0.0) == 1.0
    assert oracle.det_exp(-41.0) == 0.0
    vals = [oracle.det_exp(x) for x in np.linspace(-30, 0, 5001)]
    assert all(a <= b for a, b in zip(vals, vals[1:]))


# ------------------------------------------------------------- thresholds
@pytest.mark.parametrize("beta,H", [(0.125, 0.0), (0.9, 0.0), (2.0, 0.0), (0.5, 0.0625),
                                    (1.3, 0.1875), (0.7, 2.5)])
def test_thresholds_closed_form(beta, H):
    T = oracle.thresholds(beta, H)
    for sig in (0, 1):
        for u in range(7):
            dE = 12 - 4 * u + 2 * H * (1 if sig else -1)
            if dE <= 0:
                assert T[sig, u] == TWO32
            else:
                exact = TWO32 * math.exp(-beta * dE)
                assert abs(int(T[sig, u]) - exact) <= 1.0 + exact * 1e-15
        # acceptance rises with the number of unsatisfied bonds
        assert all(T[sig, u] <= T[sig, u + 1] for u in range(6))
    if H > 0:   # a spin along the field (bit 1) is harder to flip
        assert all(T[1, u] <= T[0, u] for u in range(7))
    else:
        assert (T[0] == T[1]).all()


# ------------------------------------------------------------------ sweeps
@pytest.mark.parametrize("L,H", [(2, 0.0), (4, 0.0), (6, 0.3), (8, 0.0), (8, 1.5)])
def test_bitsliced_equals_scalar(L, H):
    b = geometric_betas(0.4, 2.5, 3)
    a = oracle.OracleSim(L, 2, 2, b, H=H, pt_every=1, meas_every=2, scalar=False)
    s = oracle.OracleSim(L, 2, 2, b, H=H, pt_every=1, meas_every=2, scalar=True)
    for o in (a, s):
        o.init()
        o.run(6)
    np.testing.assert_array_equal(a.spins(), s.spins())
    for i in range(a.n_meas()):
        np.testing.assert_array_equal(a.meas(i), s.meas(i))


def _one_lattice(L, seed=7):
    o = oracle.OracleSim(L, 1, 1, [1.0], coupling_seed=seed, spin_seed=seed + 1, pt_every=0)
    o.init()
    return o.spins()[0, 0, 0].copy(), o.couplings()[0].copy()


@pytest.mark.parametrize("H", [0.0, 0.7])
def test_beta_zero_flips_every_spin(H):
    L = 6
    S, J = _one_lattice(L)
    T = oracle.thresholds(0.0, H)
    assert (T == TWO32).all()
    out = oracle.sweep_lattice(L, S.copy(), J, T, 99, 0, 0, 0, 0)
    np.testing.assert_array_equal(out, ~S)


def test_zero_temperature_never_raises_energy():
    L = 8
    S, J = _one_lattice(L)
    T = oracle.thresholds(60.0, 0.0)    # every dE > 0 move has threshold 0
    assert (T[:, :3] == 0).all()
    U_prev, _ = oracle.lattice_energy(L, S, J)
    for n in range(10):
        S = oracle.sweep_lattice(L, S, J, T, 5, 0, 0, 0, n)
        U, _ = oracle.lattice_energy(L, S, J)
        assert (U <= U_prev).all()
        U_prev = U


def test_energy_counts_match_direct_loop():
    for L in (2, 4, 6):
        S, J = _one_lattice(L, seed=L)
        U, K = oracle.lattice_energy(L, S, J)
        Un, Kn = bond_energy_numpy(L, S, J)
        np.testing.assert_array_equal(U, Un)
        np.testing.assert_array_equal(K, Kn)


def test_gauge_invariance():
    """J_ij -> J_ij ^ e_i ^ e_j, S -> S ^ e leaves every bond bit, hence every
    decision, unchanged (H = 0): the gauged run is the original run ^ e."""
    L = 6
    S, J = _one_lattice(L, seed=3)
    rng = np.random.default_rng(1)
    e = rng.integers(0, 2 ** 63, size=L ** 3, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    e ^= rng.integers(0, 2 ** 63, size=L ** 3, dtype=np.uint64)
    Jg = J.copy()
    for z, y, x in itertools.product(range(L), repeat=3):
        i = site(L, x, y, z)
        for d, j in enumerate((site(L, (x + 1) % L, y, z), site(L, x, (y + 1) % L, z),
                               site(L, x, y, (z + 1) % L))):
            Jg[d, i] = J[d, i] ^ e[i] ^ e[j]
    T = oracle.thresholds(0.8, 0.0)
    A, B = S.copy(), S ^ e
    for n in range(8):
        A = oracle.sweep_lattice(L, A, J, T, 11, 0, 0, 0, n)
        B = oracle.sweep_lattice(L, B, Jg, T, 11, 0, 0, 0, n)
    np.testing.assert_array_equal(B, A ^ e)
    np.testing.assert_array_equal(oracle.lattice_energy(L, A, J)[0],
                                  oracle.lattice_energy(L, B, Jg)[0])


def test_lanes_are_independent():
    """Changing the couplings of one lane changes no other lane, with PT."""
    L = 4
    b = geometric_betas(0.5, 2.0, 4)
    a = oracle.OracleSim(L, 1, 2, b, pt_every=1)
    c = oracle.OracleSim(L, 1, 2, b, pt_every=1)
    a.init()
    c.init()
    J = c.couplings()
    J ^= np.uint64(1 << 37)   # flip every bond of lane 37
    c.set_couplings(J)
    a.run(30)
    c.run(30)
    mask = np.uint64(~(1 << 37) & (2 ** 64 - 1))
    np.testing.assert_array_equal(a.spins() & mask, c.spins() & mask)
    assert ((a.spins() ^ c.spins()) != 0).any()


def test_single_site_acceptance_frequency():
    """Black sites: the fraction of lanes flipped with u unsatisfied bonds
    equals T[u] / 2^32 (u is fixed during the black half-sweep)."""
    L, beta = 8, 0.35
    T = oracle.thresholds(beta, 0.0)
    flips = np.zeros(7)
    counts = np.zeros(7)
    S, J = _one_lattice(L, seed=21)
    N = L ** 3
    black = [site(L, x, y, z) for z in range(L) for y in range(L) for x in range(L)
             if (x + y + z) % 2 == 0]
    nbrs = []
    for i in black:
        x, y, z = i % L, (i // L) % L, i // (L * L)
        nbrs.append(((site(L, (x + 1) % L, y, z), J[0, i]),
                     (site(L, (x - 1) % L, y, z), J[0, site(L, (x - 1) % L, y, z)]),
                     (site(L, x, (y + 1) % L, z), J[1, i]),
                     (site(L, x, (y - 1) % L, z), J[1, site(L, x, (y - 1) % L, z)]),
                     (site(L, x, y, (z + 1) % L), J[2, i]),
                     (site(L, x, y, (z - 1) % L), J[2, site(L, x, y, (z - 1) % L)])))
    for n in range(60):
        pre = S.copy()
        S = oracle.sweep_lattice(L, S, J, T, 1234, 0, 0, 0, n)
        u = np.zeros((len(black), 64), dtype=np.int64)
        for k, i in enumerate(black):
            for j, jw in nbrs[k]:
                u[k] += bits(np.array([pre[i] ^ pre[j] ^ jw]))[0]
        fl = bits(pre[black] ^ S[black])
        for v in range(7):
            sel = u == v
            flips[v] += fl[sel].sum()
            counts[v] += sel.sum()
    assert N > 0
    for v in range(7):
        if counts[v] < 2000:
            continue
        p = min(1.0, int(T[0, v]) / TWO32)
        f = flips[v] / counts[v]
        # words share one random number, so trials are grouped by 64: widen
        se = math.sqrt(max(p * (1 - p), 1e-6) / counts[v]) * 8
        assert abs(f - p) < 5 * se + 1e-9, (v, f, p)


# ------------------------------------------------------ exact enumeration
def _exact_l2(J_lane, beta, H):
    """Boltzmann <E>, <M> and the two-replica <q^2> of one L=2 sample."""
    L, N = 2, 8
    bonds = []
    for z, y, x in itertools.product(range(L), repeat=3):
        i = site(L, x, y, z)
        for d, j in enumerate((site(L, (x + 1) % L, y, z), site(L, x, (y + 1) % L, z),
                               site(L, x, y, (z + 1) % L))):
            bonds.append((i, j, -1 if J_lane[d, i] else 1))
    states = np.array(list(itertools.product((-1, 1), repeat=N)))
    E = np.zeros(len(states))
    for i, j, Jv in bonds:
        E -= Jv * states[:, i] * states[:, j]
    M = states.sum(axis=1)
    E = E - H * M
    w = np.exp(-beta * (E - E.min()))
    p = w / w.sum()
    q = (states @ states.T) / N
    return float(p @ E) / N, float(p @ M) / N, float(p @ (q ** 2) @ p)


@pytest.mark.parametrize("H", [0.0, 0.4])
def test_exact_enumeration_l2_with_pt(H):
    """Long PT runs at L=2 reproduce the exact Boltzmann averages of every
    lane at every temperature (a wrong PT acceptance sign biases them)."""
    betas = np.array([1.1, 0.7, 0.4])
    o = oracle.OracleSim(2, 1, 2, betas, H=H, pt_every=1, meas_every=1, coupling_seed=77,
                         spin_seed=78)
    o.init()
    J = o.couplings()[0]
    jb = bits(J)                       # (3, 8, 64)
    n = 6000
    o.run(n)
    recs = np.stack([o.meas(i)[0] for i in range(o.n_meas())])   # (n, T, 64, nf)
    N = 8
    U = recs[..., 0:2].astype(float)
    K = recs[..., 2:4].astype(float)
    E = ((2 * U - 3 * N) - H * (2 * K - N)) / N
    Wq = recs[..., 4].astype(float)
    q2 = ((N - 2 * Wq) / N) ** 2
    for t, beta in enumerate(betas):
        ex = np.array([_exact_l2(jb[:, :, l], beta, H) for l in range(64)])
        e_mc = E[:, t].mean(axis=(0, 2))           # per lane
        q2_mc = q2[:, t].mean(axis=0)
        # batch-means error over 20 blocks, pooled over lanes
        blocks = E[:, t].mean(axis=2).reshape(20, -1, 64).mean(axis=1)
        se = blocks.std(axis=0, ddof=1) / math.sqrt(20)
        z = (e_mc - ex[:, 0]) / np.maximum(se, 1e-3)
        assert abs(z.mean()) < 4 / math.sqrt(64) * 1.5 + 0.3, (t, z.mean())
        assert np.mean(np.abs(z) < 4) > 0.95
        assert abs(np.mean(q2_mc - ex[:, 2])) < 0.01


def test_equal_betas_always_swap():
    b = np.array([0.8, 0.8, 0.8, 0.8])
    o = oracle.OracleSim(4, 1, 1, b, pt_every=1)
    o.init()
    o.run(5)
    acc, att = o.swap_counts()
    np.testing.assert_array_equal(acc, att)
    assert att.sum() > 0


def test_pt_step_follows_the_acceptance_rule():
    """One PT step restated in Python: pairs (t, t+1) with t = b&1, b&1+2,
    ...; lane l accepts iff x = (b_t - b_t1)(E_t - E_t1) >= 0 or
    r < floor(2^32 exp(x)), r = Philox((l>>2, b, g, PT|rho|t), seed)[l&3]."""
    L, H, seed = 4, 0.3, 0xABCDEF12345
    betas = np.array([1.4, 1.0, 0.75, 0.55, 0.4])
    for b in (0, 1, 2, 7):
        o = oracle.OracleSim(L, 2, 2, betas, H=H, pt_every=1, spin_seed=seed)
        o.init()
        o.run(3)
        S0 = o.spins()
        U, K = o.energies()
        oracle.pt_step(o, b)
        S1 = o.spins()
        N = L ** 3
        expect = S0.copy()
        for g in range(2):
            for rho in range(2):
                for t in range(b & 1, len(betas) - 1, 2):
                    m = 0
                    for l in range(64):
                        ctr = [l >> 2, b, g, (4 << 28) | (rho << 16) | t]
                        r = int(oracle.philox(ctr, [seed & 0xFFFFFFFF, seed >> 32])[l & 3])
                        Ea = 2.0 * U[g, rho, t, l] - 3 * N - H * (2.0 * K[g, rho, t, l] - N)
                        Eb = 2.0 * U[g, rho, t + 1, l] - 3 * N - H * (2.0 * K[g, rho, t + 1, l] - N)
                        x = (betas[t] - betas[t + 1]) * (Ea - Eb)
                        # det_exp is pinned against libm by its own test
                        if x >= 0 or r < int(oracle.det_exp(x) * TWO32):
                            m |= 1 << l
                    m = np.uint64(m)
                    d = (S0[g, rho, t] ^ S0[g, rho, t + 1]) & m
                    expect[g, rho, t] ^= d
                    expect[g, rho, t + 1] ^= d
        np.testing.assert_array_equal(S1, expect)


# --------------------------------------------------------------- measuring
def test_measurement_fields_match_direct_count():
    L = 4
    b = geometric_betas(0.5, 2.0, 2)
    o = oracle.OracleSim(L, 1, 2, b, pt_every=1, meas_every=3)
    o.init()
    o.run(3)
    rec = o.meas(0)[0]          # (T, 64, nf)
    S = o.spins()[0]            # (R, T, N)
    U, K = o.energies()
    for t in range(2):
        w = bits(S[0, t] ^ S[1, t])              # (N, 64)
        np.testing.assert_array_equal(rec[t, :, 4], w.sum(axis=0))
        wx = w.reshape(L, L, L, 64)              # (z, y, x, lane)
        np.testing.assert_array_equal(rec[t, :, 5:5 + L], wx.sum(axis=(0, 1)).T)
        np.testing.assert_array_equal(rec[t, :, 5 + L:5 + 2 * L], wx.sum(axis=(0, 2)).T)
        np.testing.assert_array_equal(rec[t, :, 5 + 2 * L:5 + 3 * L], wx.sum(axis=(1, 2)).T)
        np.testing.assert_array_equal(rec[t, :, 0:2].T, U[0, :, t])
        np.testing.assert_array_equal(rec[t, :, 2:4].T, K[0, :, t])


def test_fourier_overlap_formula():
    """-(2/N) sum_c W_d[c] e^{2 pi i c/L} equals N^-1 sum_s s1 s2 e^{i k x_s}
    (reference eamsim.cu:486-500), checked on random states."""
    L, N = 6, 216
    rng = np.random.default_rng(3)
    s1 = rng.choice([-1, 1], size=(L, L, L))
    s2 = rng.choice([-1, 1], size=(L, L, L))
    qloc = s1 * s2                               # [z, y, x]
    w = (s1 != s2).astype(int)
    k = 2 * math.pi / L
    c = np.arange(L)
    for axis_sum, coord in (((0, 1), "x"), ((0, 2), "y"), ((1, 2), "z")):
        Wd = w.sum(axis=axis_sum)
        direct = sum(qloc[z, y, x] * np.exp(1j * k * {"x": x, "y": y, "z": z}[coord])
                     for z in range(L) for y in range(L) for x in range(L)) / N
        via_w = -2.0 / N * complex(np.cos(k * c) @ Wd, np.sin(k * c) @ Wd)
        assert abs(direct - via_w) < 1e-12
