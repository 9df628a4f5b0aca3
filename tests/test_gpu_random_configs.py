"""Randomised run configurations against the CPU oracle, bit for bit.

Each case draws, from its own seed, a lattice size, a sample count (partial
groups included), replicas, a temperature ladder, a field, the PT and
measurement cadences, the form, a kernel shape and tile the library offers
at that size (the tuner's pick where none is asked for), and a few run
lengths that end inside and across PT intervals and measurements.  After
every run the spins, per-lane energies, swap counts and measurement records
must equal the oracle's, and the FP64 results those of the oracle's records.
The fixed-shape tests elsewhere cover each kernel at chosen points; these
cases cover the combinations between them."""
import numpy as np
import pytest

import oracle
import sg
from helpers import geometric_betas, results_from_records

pytestmark = pytest.mark.gpu

SIZES = [2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 36, 40, 48]
KERNELS = [sg.SG_KERNEL_AUTO, sg.SG_KERNEL_RESIDENT, sg.SG_KERNEL_STREAMED,
           sg.SG_KERNEL_RESIDENT_SINGLE, sg.SG_KERNEL_SPLIT]


def _draw(i):
    rng = np.random.default_rng(0x5EED0100 + i)
    L = int(rng.choice(SIZES))
    big = L >= 28
    n_samples = int(rng.integers(1, 65 if big else 140))
    R = int(rng.integers(1, 3 if big else 4))
    T = int(rng.integers(1, 5 if big else 11))
    tmin = float(rng.uniform(0.2, 0.8))
    betas = geometric_betas(tmin, float(rng.uniform(tmin + 0.1, 2.5)), T)
    H = float(rng.choice([0.0, 0.0, 0.0625, 0.3, 1.0, 2.5]))
    pt_every = int(rng.integers(0, 9))
    meas_every = int(rng.choice([0, int(rng.integers(1, 13))]))
    form = int(rng.choice([sg.SG_FORM_PLAIN, sg.SG_FORM_BATCHED]))
    kernel = int(rng.choice(KERNELS))
    tile = int(rng.integers(0, 3))
    runs = [int(x) for x in rng.integers(1, 14, size=int(rng.integers(2, 4)))]
    seeds = [int(x) for x in rng.integers(1, 2**63, size=2)]
    return dict(L=L, n_samples=n_samples, R=R, betas=betas, H=H, pt_every=pt_every,
                meas_every=meas_every, form=form, kernel=kernel, tile=tile, runs=runs,
                seeds=seeds)


def _make(c):
    """The case's handle, with its kernel and tile if the library has them at
    this L and size, else the tuner's pick (what an out-of-range request gets
    is tested in test_gpu_parity.py)."""
    kw = dict(L=c["L"], n_samples=c["n_samples"], n_replicas=c["R"], betas=c["betas"],
              H=c["H"], coupling_seed=c["seeds"][0], spin_seed=c["seeds"][1],
              pt_every=c["pt_every"], meas_every=c["meas_every"], form=c["form"])
    for kernel, tile in ((c["kernel"], c["tile"]), (c["kernel"], 0), (sg.SG_KERNEL_AUTO, 0)):
        try:
            return sg.Sim(kernel=kernel, tile=tile, **kw)
        except sg.SGError:
            continue
    raise AssertionError("the tuner's pick was refused")


@pytest.mark.parametrize("i", range(48))
def test_random_configuration_matches_oracle(i):
    c = _draw(i)
    sim = _make(c)
    print(f"case {i}: L={c['L']} samples={c['n_samples']} R={c['R']} T={len(c['betas'])} "
          f"H={c['H']} pt={c['pt_every']} meas={c['meas_every']} form={c['form']} "
          f"shape={sim.shape} runs={c['runs']}", flush=True)
    sim.init_lattice()
    orc = oracle.OracleSim(c["L"], sim.G, c["R"], c["betas"], H=c["H"],
                           coupling_seed=c["seeds"][0], spin_seed=c["seeds"][1],
                           pt_every=c["pt_every"], meas_every=c["meas_every"])
    orc.init()
    for n in c["runs"]:
        sim.run(n)
        orc.run(n)
        assert sim.sweep == orc.sweep
        np.testing.assert_array_equal(sim.spins(), orc.spins())
        U, K = sim.energies()
        Uo, Ko = orc.energies()
        np.testing.assert_array_equal(U, Uo)
        np.testing.assert_array_equal(K, Ko)
        a, b = sim.swap_counts()
        ao, bo = orc.swap_counts()
        np.testing.assert_array_equal(a, ao)
        np.testing.assert_array_equal(b, bo)
    assert len(sim.records) == orc.n_meas()
    for k, (_, rec) in enumerate(sim.records):
        np.testing.assert_array_equal(rec, orc.meas(k))
    if orc.n_meas():
        ref = results_from_records([orc.meas(k) for k in range(orc.n_meas())], c["L"], c["R"],
                                   c["H"], c["n_samples"], c["betas"])
        np.testing.assert_array_equal(sim.results(), ref)
