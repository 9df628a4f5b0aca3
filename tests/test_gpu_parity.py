"""Bit-exact parity of libsg (HIP, through the C-ABI) with the CPU oracle.

Every spin word, per-lane energy, measurement record and swap count must be
identical; the FP64 results must agree within 1e-12 relative."""
import numpy as np
import pytest

import oracle
import sg
from helpers import geometric_betas, results_from_records

pytestmark = pytest.mark.gpu

CSEED, SSEED = 0x5EED0001, 0x5EED0002


def _pair(L, n_samples, R, betas, H=0.0, pt_every=1, meas_every=0, form=sg.SG_FORM_PLAIN,
          cseed=CSEED, sseed=SSEED):
    sim = sg.Sim(L=L, n_samples=n_samples, n_replicas=R, betas=betas, H=H,
                 coupling_seed=cseed, spin_seed=sseed, pt_every=pt_every,
                 meas_every=meas_every, form=form)
    sim.init_lattice()
    orc = oracle.OracleSim(L, sim.G, R, betas, H=H, coupling_seed=cseed, spin_seed=sseed,
                           pt_every=pt_every, meas_every=meas_every)
    orc.init()
    return sim, orc


def _assert_same(sim, orc):
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


@pytest.mark.parametrize("L", [2, 4, 8, 16])
def test_init_matches_oracle(L):
    sim, orc = _pair(L, 100, 2, geometric_betas(0.5, 2.0, 3))
    np.testing.assert_array_equal(sim.couplings(), orc.couplings())
    np.testing.assert_array_equal(sim.spins(), orc.spins())


@pytest.mark.parametrize("L", [2, 4, 6, 8, 10, 12, 14, 16])
def test_sweeps_pt_measurements_match_oracle(L):
    betas = geometric_betas(0.5, 2.0, 4)
    sim, orc = _pair(L, 80, 2, betas, pt_every=2, meas_every=3)
    sim.run(5)
    orc.run(5)
    _assert_same(sim, orc)
    sim.run(7)
    orc.run(7)
    _assert_same(sim, orc)
    assert len(sim.records) == orc.n_meas() == 4
    for i, (sweep, rec) in enumerate(sim.records):
        assert sweep == 3 * (i + 1)
        np.testing.assert_array_equal(rec, orc.meas(i))
    ref = results_from_records([orc.meas(i) for i in range(orc.n_meas())], L, 2, 0.0, 80, betas)
    np.testing.assert_allclose(sim.results(), ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("H", [0.0625, 0.5, 2.5])
def test_field_matches_oracle(H):
    betas = geometric_betas(0.3, 2.0, 5)
    sim, orc = _pair(8, 64, 2, betas, H=H, pt_every=1, meas_every=4)
    sim.run(9)
    orc.run(9)
    _assert_same(sim, orc)
    for i, (_, rec) in enumerate(sim.records):
        np.testing.assert_array_equal(rec, orc.meas(i))


def test_batched_form_is_bit_identical():
    betas = geometric_betas(0.2, 2.0, 6)
    plain, orc = _pair(12, 192, 2, betas, pt_every=4, meas_every=8)
    batched = sg.Sim(L=12, n_samples=192, n_replicas=2, betas=betas, coupling_seed=CSEED,
                     spin_seed=SSEED, pt_every=4, meas_every=8, form=sg.SG_FORM_BATCHED)
    batched.init_lattice()
    for s in (plain, batched):
        s.run(21)
    orc.run(21)
    _assert_same(plain, orc)
    _assert_same(batched, orc)
    for (_, a), (_, b) in zip(plain.records, batched.records):
        np.testing.assert_array_equal(a, b)


def test_reference_ladder_l16_matches_oracle():
    sim, orc = _pair(16, 64, 2, sg.REF_BETAS, pt_every=16, meas_every=32)
    sim.run(40)
    orc.run(40)
    _assert_same(sim, orc)
    np.testing.assert_array_equal(sim.records[0][1], orc.meas(0))


def test_checkpoint_resume_is_bit_exact():
    betas = geometric_betas(0.5, 2.0, 4)
    a = sg.Sim(L=8, n_samples=64, betas=betas, pt_every=3, meas_every=5)
    a.init_lattice()
    a.run(10)
    blob = a.save_checkpoint_bytes()
    a.run(10)
    b = sg.Sim(L=8, n_samples=64, betas=betas, pt_every=3, meas_every=5)
    b.load_checkpoint_bytes(blob)
    b.run(10)
    np.testing.assert_array_equal(a.spins(), b.spins())
    np.testing.assert_array_equal(a.records[-1][1], b.records[-1][1])
    np.testing.assert_allclose(a.results(), b.results(), rtol=0, atol=0)
    bad = bytearray(blob)
    bad[len(bad) // 2] ^= 1
    with pytest.raises(sg.SGError, match="checksum"):
        b.load_checkpoint_bytes(bytes(bad))
    with pytest.raises(sg.SGError, match="size"):
        b.load_checkpoint_bytes(blob[:-8])
