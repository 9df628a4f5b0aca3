"""The sweep kernels' Metropolis network against the acceptance rule it
replaces, by exhaustive enumeration (CPU).

The kernels (metropolis32 / metropolis32_field in
spin-glass_amd/csrc/sg_kernels.hip) do not compare u with a threshold table
per lane.  With thresholds rising in u, "flip iff u >= nfin or r < T[sigma][u]"
equals "u >= k" for a per-site k, and u >= k is the carry out of
u + (8 - k) >= 8: the bias 8 - k enters the carry-save tree of the six bond
bits and only carries are formed.  These tests restate that network bit by
bit (one lane; the kernels run it on 32 lanes at once) and compare it with
the rule on every bond pattern, spin and threshold count.  The rule is the
one the reference applies per spin (Program/src/eamsim.cu:184-198) and the
oracle evaluates directly (oracle/sg_oracle.c).  The GPU parity tests then
check the kernels themselves against the oracle.
"""
import itertools


def maj(a, b, c):
    return (a & b) | (a & c) | (b & c)


def sel(a, b, c):
    return b if a else c


def rule(u, s, R):
    """Flip iff u >= 5, else iff r is below T[s][u]; R[s][u] says whether it
    is (thresholds past nfin count as 2^32, always below)."""
    return 1 if u >= 5 else R[s][u]


def carry_network_h0(e, B0, B1):
    """metropolis32: 15 ops per half word (the 6 bond XORs give e); B0 = the
    threshold count n is odd, B1 = n >= 2."""
    s1, k1 = e[0] ^ e[1] ^ e[2], maj(e[0], e[1], e[2])
    s2, k2 = e[3] ^ e[4] ^ e[5], maj(e[3], e[4], e[5])
    k3 = maj(s1, s2, 1 - B0)
    t, m1 = k1 ^ k2 ^ k3, maj(k1, k2, k3)
    g = maj(t, B0, B1)
    return m1 | g


def carry_network_field(e, s, C):
    """metropolis32_field: 20 ops per half word; C[sigma] = (bit 0, bit 1,
    bit 2 | bit 3, bit 3) of the bias 8 - k for spin bit sigma."""
    s1, k1 = e[0] ^ e[1] ^ e[2], maj(e[0], e[1], e[2])
    s2, k2 = e[3] ^ e[4] ^ e[5], maj(e[3], e[4], e[5])
    c0 = sel(s, C[1][0], C[0][0])
    k3 = maj(s1, s2, c0)
    t, m1 = k1 ^ k2 ^ k3, maj(k1, k2, k3)
    m2 = t & sel(s, C[1][1], C[0][1])
    w8 = maj(m1, m2, sel(s, C[1][2], C[0][2]))
    return w8 | sel(s, C[1][3], C[0][3])


def test_h0_network_equals_the_threshold_rule():
    checked = 0
    for e in itertools.product((0, 1), repeat=6):
        u = sum(e)
        # H = 0: thresholds for u = 0, 1, 2 (dE = 12, 8, 4), T(0) <= T(1) <= T(2);
        # n = how many of them r is below (below T(0) implies below T(1), T(2))
        for n in range(4):
            a0, a1, a2 = int(n >= 3), int(n >= 2), int(n >= 1)
            R_rule = [[a0, a1, a2, 1, 1]] * 2        # u = 3, 4 always flip
            B0 = a0 ^ a1 ^ a2                        # the kernel's masks
            B1 = a1
            for s in (0, 1):
                assert carry_network_h0(e, B0, B1) == rule(u, s, R_rule), (e, n)
                checked += 1
    assert checked == 64 * 4 * 2


def test_field_network_equals_the_threshold_rule():
    checked = 0
    for e in itertools.product((0, 1), repeat=6):
        u = sum(e)
        # per spin bit: a[u] (u = 0..4) rises with u (a thermometer code);
        # m = a[0] + ... + a[4] in 0..5 covers every table with nfin <= 5
        for m0, m1 in itertools.product(range(6), repeat=2):
            a = [[int(v >= 5 - m) for v in range(5)] for m in (m0, m1)]
            C = []
            for g in range(2):
                x = a[g]
                C.append((1 - (x[0] ^ x[1] ^ x[2] ^ x[3] ^ x[4]),
                          int(x[4] == x[2] and not x[0]),
                          x[4],
                          x[0]))
            for s in (0, 1):
                assert carry_network_field(e, s, C) == rule(u, s, a), (e, m0, m1, s)
                checked += 1
    assert checked == 64 * 36 * 2


def test_a_dropped_bias_bit_is_caught():
    """The enumeration is sharp: the H = 0 network with its weight-1 bias
    bit inverted (B0 taken for n even) disagrees with the rule somewhere."""
    bad = 0
    for e in itertools.product((0, 1), repeat=6):
        for n in range(4):
            a0, a1, a2 = int(n >= 3), int(n >= 2), int(n >= 1)
            R_rule = [[a0, a1, a2, 1, 1]] * 2
            bad += carry_network_h0(e, 1 - (a0 ^ a1 ^ a2), a1) != rule(sum(e), 0, R_rule)
    assert bad > 0
