"""Bit-exact parity of libsg (HIP, through the C-ABI) with the CPU oracle.

Every spin word, per-lane energy, measurement record and swap count must be
identical, and so must the FP64 results (exact integer totals and one
fixed-order FP64 sum, restated by tests/helpers.py), well inside the 1e-12
relative tolerance the results are specified to."""
import numpy as np
import pytest

import oracle
import sg
from helpers import geometric_betas, results_from_records

pytestmark = pytest.mark.gpu

CSEED, SSEED = 0x5EED0001, 0x5EED0002


def _pair(L, n_samples, R, betas, H=0.0, pt_every=1, meas_every=0, form=sg.SG_FORM_PLAIN,
          cseed=CSEED, sseed=SSEED, kernel=sg.SG_KERNEL_AUTO, tile=0):
    sim = sg.Sim(L=L, n_samples=n_samples, n_replicas=R, betas=betas, H=H,
                 coupling_seed=cseed, spin_seed=sseed, pt_every=pt_every,
                 meas_every=meas_every, form=form, kernel=kernel, tile=tile)
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
    np.testing.assert_array_equal(sim.results(), ref)


@pytest.mark.parametrize("form", [sg.SG_FORM_PLAIN, sg.SG_FORM_BATCHED])
@pytest.mark.parametrize("L", [12, 14])
def test_plain_dropin_shape_at_L12_14_matches_oracle(L, form):
    """The plain drop-in configuration at L = 12 and 14 with its full
    ladder: 64 samples, 2 replicas, 24 temperatures geometric in T = [0.2, 2],
    PT every 16 sweeps, a measurement every 48; three PT intervals in the
    plain and the batched form (the tuner's shape).  Spins, energies, swap
    counts, records and FP64 results equal the oracle's."""
    betas = geometric_betas(0.2, 2.0, 24)
    sim, orc = _pair(L, 64, 2, betas, pt_every=16, meas_every=48, form=form)
    sim.run(48)
    orc.run(48)
    _assert_same(sim, orc)
    assert len(sim.records) == orc.n_meas() == 1
    np.testing.assert_array_equal(sim.records[0][1], orc.meas(0))
    ref = results_from_records([orc.meas(0)], L, 2, 0.0, 64, betas)
    np.testing.assert_array_equal(sim.results(), ref)


@pytest.mark.parametrize("L", [18, 20, 22, 24, 26, 28, 32, 48])
def test_large_lattice_sweeps_match_oracle(L):
    """L > 16 runs the resident large-lattice kernels: two lattices in LDS
    with couplings from the site records for L = 18, 20 (site 1 of a quad has
    the opposite parity at L = 18, the same at L = 20); one lattice in LDS
    and PT in its own launch for L = 22, 24, 26 (uneven z-runs at 22 and 26);
    updated in place in global memory above; 80 samples = one full and one
    partial lane group; 4 temperatures give pair-only and mixed task lists."""
    betas = geometric_betas(0.3, 2.0, 4)
    sim, orc = _pair(L, 80, 2, betas, pt_every=2, meas_every=3, kernel=sg.SG_KERNEL_RESIDENT)
    sim.run(4)
    orc.run(4)
    _assert_same(sim, orc)
    sim.run(3)
    orc.run(3)
    _assert_same(sim, orc)
    assert len(sim.records) == orc.n_meas() == 2
    for i, (_, rec) in enumerate(sim.records):
        np.testing.assert_array_equal(rec, orc.meas(i))


@pytest.mark.parametrize("form", [sg.SG_FORM_PLAIN, sg.SG_FORM_BATCHED])
@pytest.mark.parametrize("H", [0.0, 0.5])
@pytest.mark.parametrize("L", [32, 48])
def test_split_kernel_matches_oracle(L, H, form):
    """A lattice held by four work-groups (slab pairs exchanging boundary
    planes every half-sweep; both colours in LDS at L = 32, one colour at a
    time at L = 48): 80 samples (a full and a partial group; the plain form
    launches each group's 16 lattices alone), PT every 4 sweeps, a
    measurement every 8, runs ending inside an interval."""
    betas = geometric_betas(0.3, 2.0, 8)
    sim, orc = _pair(L, 80, 2, betas, H=H, pt_every=4, meas_every=8, form=form,
                     kernel=sg.SG_KERNEL_SPLIT)
    for n in (4, 9, 11):
        sim.run(n)
        orc.run(n)
        _assert_same(sim, orc)
    assert len(sim.records) == orc.n_meas() == 3
    for i, (_, rec) in enumerate(sim.records):
        np.testing.assert_array_equal(rec, orc.meas(i))


@pytest.mark.parametrize("H", [0.5])
@pytest.mark.parametrize("form", [sg.SG_FORM_PLAIN, sg.SG_FORM_BATCHED])
def test_split_kernel_second_tile_matches_oracle(form, H):
    """The split kernel's second tile at L = 32 (512 threads per work-group,
    four quads each instead of two)."""
    betas = geometric_betas(0.3, 2.0, 8)
    sim, orc = _pair(32, 80, 2, betas, H=H, pt_every=4, meas_every=8, form=form,
                     kernel=sg.SG_KERNEL_SPLIT, tile=2)
    assert sim.shape == (sg.SG_KERNEL_SPLIT, 1)
    for n in (4, 9, 11):
        sim.run(n)
        orc.run(n)
        _assert_same(sim, orc)
    for i, (_, rec) in enumerate(sim.records):
        np.testing.assert_array_equal(rec, orc.meas(i))


@pytest.mark.parametrize("L,T,H,form", [(32, 24, 0.0, sg.SG_FORM_BATCHED),
                                         (48, 20, 0.5, sg.SG_FORM_BATCHED),
                                         (32, 40, 0.5, sg.SG_FORM_PLAIN)])
def test_split_kernel_beyond_one_launch_matches_oracle(L, T, H, form):
    """More lattices than one split launch holds (64: one work-group per CU
    for each of their four slab pairs): 128 samples, so 96 and 80 lattices in
    the batched form and 80 per group in the plain form, run as consecutive
    launches of at most 64."""
    betas = geometric_betas(0.3, 2.0, T)
    sim, orc = _pair(L, 128, 2, betas, H=H, pt_every=4, meas_every=8, form=form,
                     kernel=sg.SG_KERNEL_SPLIT)
    assert sim.shape[0] == sg.SG_KERNEL_SPLIT
    for n in (4, 9):
        sim.run(n)
        orc.run(n)
        _assert_same(sim, orc)
    for i, (_, rec) in enumerate(sim.records):
        np.testing.assert_array_equal(rec, orc.meas(i))


def test_missing_tile_is_rejected():
    betas = geometric_betas(0.3, 2.0, 4)
    for kw in (dict(L=48, kernel=sg.SG_KERNEL_SPLIT, tile=2),   # one tile at L = 48
               dict(L=16, kernel=sg.SG_KERNEL_STREAMED, tile=3),
               dict(L=16, kernel=sg.SG_KERNEL_RESIDENT, tile=2)):
        with pytest.raises(sg.SGError, match="tile"):
            sg.Sim(n_samples=64, n_replicas=2, betas=betas, **kw)


@pytest.mark.parametrize("kernel", [sg.SG_KERNEL_RESIDENT, sg.SG_KERNEL_STREAMED])
@pytest.mark.parametrize("H", [0.0, 0.5])
@pytest.mark.parametrize("L", [30, 34, 36, 38, 42, 44, 46])
def test_other_large_sizes_match_oracle(L, H, kernel):
    """Sizes the benchmarks do not use: k_big (resident, in L2) and k_phase
    (streamed colour launches) at L = 30, 34, 38, 42 and 46 (L = 2 mod 4: a
    colour row holds an odd number of words) and 36, 44, at H = 0 and
    H != 0.  With L = 28, 32, 40, 48 elsewhere, every even L above 26 that
    sg_create accepts is covered."""
    betas = geometric_betas(0.3, 2.0, 3)
    sim, orc = _pair(L, 64, 2, betas, H=H, pt_every=2, meas_every=3, kernel=kernel)
    sim.run(4)
    orc.run(4)
    _assert_same(sim, orc)
    assert len(sim.records) == orc.n_meas() == 1
    np.testing.assert_array_equal(sim.records[0][1], orc.meas(0))


@pytest.mark.parametrize("case", [
    # the reference's CPU case: L = 8, 32 samples (half a lane group), 16
    # temperatures geometric in [0.5, 2], PT every sweep
    dict(L=8, n=32, R=2, T=(0.5, 2.0, 16), pt=1, meas=4),
    dict(L=12, n=64, R=1, T=(0.2, 2.0, 5), pt=2, meas=3),            # one replica
    dict(L=16, n=64, R=2, T=(0.5, 0.5, 1), pt=4, meas=4),            # one temperature: no pairs
    dict(L=20, n=64, R=2, T=(0.2, 2.0, 4), pt=0, meas=0),            # no PT, no measurement
    dict(L=24, n=130, R=2, T=(0.3, 2.0, 3), pt=2, meas=3,            # three groups, k_one,
         form=sg.SG_FORM_BATCHED, kernel=sg.SG_KERNEL_RESIDENT),      # captured intervals
    dict(L=32, n=64, R=1, T=(0.3, 2.0, 2), pt=2, meas=0,             # slab sweeps, one replica
         kernel=sg.SG_KERNEL_STREAMED),
], ids=["reference_cpu_case_L8", "R1", "T1", "noPT", "L24_3groups_batched", "L32_slab_R1"])
def test_edge_shapes_match_oracle(case):
    tmin, tmax, nt = case["T"]
    betas = geometric_betas(tmin, tmax, nt)
    sim, orc = _pair(case["L"], case["n"], case["R"], betas, pt_every=case["pt"],
                     meas_every=case["meas"], form=case.get("form", sg.SG_FORM_PLAIN),
                     kernel=case.get("kernel", sg.SG_KERNEL_AUTO))
    for n in (5, 4, 8):
        sim.run(n)
        orc.run(n)
        _assert_same(sim, orc)
    assert len(sim.records) == orc.n_meas()
    for i, (_, rec) in enumerate(sim.records):
        np.testing.assert_array_equal(rec, orc.meas(i))


@pytest.mark.parametrize("form", [sg.SG_FORM_PLAIN, sg.SG_FORM_BATCHED])
@pytest.mark.parametrize("L,H", [(2, 0.0), (4, 0.0), (8, 0.0), (12, 0.5), (16, 0.0), (18, 0.0),
                                 (20, 0.25), (28, 0.0), (32, 0.0)])
def test_single_lattice_shape_matches_oracle(L, H, form):
    """One lattice per work-group with the two-lattice kernels (k_block,
    k_mid, k_big): every lattice its own task, PT in its own launch; 5
    temperatures (pairs of both parities leave different slots out), one
    full and one partial lane group, measurements between PT steps."""
    betas = geometric_betas(0.3, 2.0, 5)
    sim, orc = _pair(L, 80, 2, betas, H=H, pt_every=2, meas_every=3, form=form,
                     kernel=sg.SG_KERNEL_RESIDENT_SINGLE)
    for n in (4, 3, 5):
        sim.run(n)
        orc.run(n)
        _assert_same(sim, orc)
    assert len(sim.records) == orc.n_meas() == 4
    for i, (_, rec) in enumerate(sim.records):
        np.testing.assert_array_equal(rec, orc.meas(i))


@pytest.mark.parametrize("tile", [1, 2])
@pytest.mark.parametrize("L", [12, 16, 18, 24, 32, 40, 48])
def test_streamed_kernel_matches_oracle(L, tile):
    """One launch per colour over (lattice, chunk) tasks, PT in its own
    launch, energies accumulated by the last colour launch; at L = 32 pairs
    of sweeps run as slab launches (one per sweep, double-buffered spins) and
    odd leftovers as colour launches: segments of 2 and 1 sweeps here.  Both
    tiles: one or two quads (slab columns) per thread."""
    betas = geometric_betas(0.3, 2.0, 4)
    sim, orc = _pair(L, 80, 2, betas, pt_every=2, meas_every=3, kernel=sg.SG_KERNEL_STREAMED,
                     tile=tile)
    assert sim.shape == (sg.SG_KERNEL_STREAMED, tile - 1)
    sim.run(4)
    orc.run(4)
    _assert_same(sim, orc)
    sim.run(3)
    orc.run(3)
    _assert_same(sim, orc)
    for i, (_, rec) in enumerate(sim.records):
        np.testing.assert_array_equal(rec, orc.meas(i))


@pytest.mark.parametrize("L,H,pt", [(20, 0.25, 1), (20, 1.5, 1), (32, 0.5, 2), (48, 0.25, 4)])
def test_streamed_kernel_field_matches_oracle(L, H, pt):
    betas = geometric_betas(0.3, 2.0, 3)
    sim, orc = _pair(L, 64, 2, betas, H=H, pt_every=pt, meas_every=3,
                     kernel=sg.SG_KERNEL_STREAMED)
    sim.run(5)
    orc.run(5)
    _assert_same(sim, orc)
    np.testing.assert_array_equal(sim.records[0][1], orc.meas(0))


@pytest.mark.parametrize("L", [20, 32])
def test_resident_and_streamed_kernels_agree_in_both_forms(L):
    betas = geometric_betas(0.2, 2.0, 6)
    runs = []
    for kernel in (sg.SG_KERNEL_RESIDENT, sg.SG_KERNEL_STREAMED):
        for form in (sg.SG_FORM_PLAIN, sg.SG_FORM_BATCHED):
            s = sg.Sim(L=L, n_samples=128, n_replicas=2, betas=betas, pt_every=4, meas_every=8,
                       form=form, kernel=kernel)
            s.init_lattice()
            s.run(36)          # the batched form replays its captured intervals
            runs.append((s.spins(), s.energies(), s.swap_counts(), s.results()))
    for r in runs[1:]:
        np.testing.assert_array_equal(r[0], runs[0][0])
        np.testing.assert_array_equal(r[1][0], runs[0][1][0])
        np.testing.assert_array_equal(r[1][1], runs[0][1][1])
        np.testing.assert_array_equal(r[2][0], runs[0][2][0])
        np.testing.assert_array_equal(r[3], runs[0][3])


@pytest.mark.parametrize("L", [18, 20, 22, 24, 26])
def test_large_lattice_field_matches_oracle(L):
    betas = geometric_betas(0.3, 2.0, 3)
    sim, orc = _pair(L, 64, 2, betas, H=0.5, pt_every=1, meas_every=3,
                     kernel=sg.SG_KERNEL_RESIDENT)
    sim.run(5)
    orc.run(5)
    _assert_same(sim, orc)
    np.testing.assert_array_equal(sim.records[0][1], orc.meas(0))


@pytest.mark.parametrize("form", [sg.SG_FORM_PLAIN, sg.SG_FORM_BATCHED])
def test_maximum_replicas_and_many_temperatures_match_oracle(form):
    """The largest replica count accepted (8), 40 temperatures (19 and 20 PT
    pairs by parity), a partial second lane group, PT every sweep."""
    betas = geometric_betas(0.3, 2.0, 40)
    sim, orc = _pair(6, 70, 8, betas, pt_every=1, meas_every=2, form=form)
    for n in (3, 4):
        sim.run(n)
        orc.run(n)
        _assert_same(sim, orc)
    for i, (_, rec) in enumerate(sim.records):
        np.testing.assert_array_equal(rec, orc.meas(i))


@pytest.mark.parametrize("H", [0.0625, 0.5, 2.0, 2.5, 4.0])
def test_field_matches_oracle(H):
    """H != 0 thresholds by spin bit and unsatisfied bonds u: at H = 2 the
    u = 4 move of an up spin costs nothing (four finite thresholds for one
    spin value, five for the other), at H = 2.5 both have a finite u = 4
    threshold, at H = 4 (the largest H accepted) u = 5 costs nothing."""
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


@pytest.mark.parametrize("warm,steps", [(5, 20), (64, 64), (3, 7), (61, 4)])
def test_prepared_runs_instantiate_nothing_and_match_oracle(warm, steps):
    """sg_prepare builds every graph two split calls replay (measurement every
    1000 sweeps, PT every 16, so the replay lengths vary near each call's end
    and each measurement); the runs then add no graph and give the oracle's
    bits."""
    betas = geometric_betas(0.4, 2.0, 6)
    sim, orc = _pair(8, 64, 2, betas, pt_every=16, meas_every=1000, form=sg.SG_FORM_BATCHED)
    sim.prepare(warm * 16)
    sim.prepare(steps * 16, from_sweep=warm * 16)
    n = sim.graph_count
    assert n > 0
    sim.run(warm * 16)
    sim.run(steps * 16)
    assert sim.graph_count == n
    orc.run((warm + steps) * 16)
    _assert_same(sim, orc)


def test_reference_ladder_l16_matches_oracle():
    sim, orc = _pair(16, 64, 2, sg.REF_BETAS, pt_every=16, meas_every=32)
    sim.run(40)
    orc.run(40)
    _assert_same(sim, orc)
    np.testing.assert_array_equal(sim.records[0][1], orc.meas(0))


def test_benchmark_workload_at_full_size_matches_oracle():
    """The benchmark's default workload at its full size: L=16, 2048 samples
    (32 groups), 24 temperatures geometric in T = [0.2, 2], 2 replicas, PT
    every 16 sweeps, batched form with the tuner's work-group shape.  Three PT
    intervals and one measurement: every spin word, per-lane energy, swap
    count and record equal to the oracle's, FP64 results within 1e-12."""
    betas = geometric_betas(0.2, 2.0, 24)
    sim, orc = _pair(16, 2048, 2, betas, pt_every=16, meas_every=32, form=sg.SG_FORM_BATCHED)
    sim.run(48)
    orc.run(48)
    _assert_same(sim, orc)
    assert len(sim.records) == orc.n_meas() == 1
    np.testing.assert_array_equal(sim.records[0][1], orc.meas(0))
    ref = results_from_records([orc.meas(0)], 16, 2, 0.0, 2048, betas)
    np.testing.assert_array_equal(sim.results(), ref)


@pytest.mark.parametrize("L", [32, 48])
def test_large_lattice_benchmark_at_full_size_matches_oracle(L):
    """The large-lattice benchmark's shape at its full size: 64 samples, 32
    temperatures geometric in T = [0.3, 2], 2 replicas (64 lattices), PT every
    16, batched form with the tuner's shape and tile (the split kernel at
    both sizes on the MI355X).  Two PT intervals and a measurement, bit-exact."""
    betas = geometric_betas(0.3, 2.0, 32)
    sim, orc = _pair(L, 64, 2, betas, pt_every=16, meas_every=32, form=sg.SG_FORM_BATCHED)
    k, t, _ = sg.tune_tiles(L, 64)
    assert sim.shape == (k, t)
    sim.run(32)
    orc.run(32)
    _assert_same(sim, orc)
    assert len(sim.records) == orc.n_meas() == 1
    np.testing.assert_array_equal(sim.records[0][1], orc.meas(0))
    ref = results_from_records([orc.meas(0)], L, 2, 0.0, 64, betas)
    np.testing.assert_array_equal(sim.results(), ref)


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


def test_product_reproduces_canonical_digests():
    import hashlib
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "canonical.json")))
    for name, g in gold.items():
        p = g["params"]
        sim = sg.Sim(L=p["L"], n_samples=64 * p["groups"], n_replicas=p["replicas"],
                     betas=p["betas"], H=p["H"], coupling_seed=CSEED, spin_seed=SSEED,
                     pt_every=p["pt_every"], meas_every=p["meas_every"])
        sim.init_lattice()
        sim.run(p["sweeps"])
        dig = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
        assert dig(sim.couplings()) == g["couplings"], name
        assert dig(sim.spins()) == g["spins"], name
        assert [dig(r) for _, r in sim.records] == g["records"], name


def test_cli_resume_equals_straight_run(tmp_path):
    """csg <sim_dir> <H> twice (the second resumes from the checkpoint)
    appends exactly the obs.bin records of one run twice as long; the log
    file is removed on success; analysis.py turns obs.bin into res.dat."""
    import os
    import subprocess
    import analysis
    csg = os.path.join(os.path.dirname(sg.__file__), "bin", "csg")
    flags = ["--L", "8", "--samples", "96", "--meas-every", "64", "--temps", "6",
             "--tmin", "0.5", "--tmax", "3.0", "--pt-every", "4"]
    a, b = tmp_path / "a" / "08", tmp_path / "b" / "08"
    a.mkdir(parents=True)
    b.mkdir(parents=True)
    for _ in range(2):
        subprocess.run([csg, str(a), "0.25", "--sweeps", "512", *flags], check=True,
                       cwd=tmp_path, capture_output=True)
    subprocess.run([csg, str(b), "0.25", "--sweeps", "1024", *flags], check=True,
                   cwd=tmp_path, capture_output=True)
    ra = (a / "0.2500-obs.bin").read_bytes()
    rb = (b / "0.2500-obs.bin").read_bytes()
    assert len(ra) == 16 * 96 * 6 * 44
    assert ra == rb
    assert not list(tmp_path.glob("*.log"))
    out = analysis.analyze_obs(str(tmp_path / "a"), 8, 0.25, n_samples=96, n_temps=6,
                               betas=1.0 / (0.5 * 6.0 ** (np.arange(6) / 5)), sweeps_per_record=64)
    rows = np.loadtxt(out)
    assert rows.shape == (6, 9)
    assert np.all(rows[:, 5] < 0)            # e < 0 at every temperature
    assert np.all(np.diff(rows[:, 5]) > 0)   # and rises with T


CLI_FLAGS = ["--L", "8", "--samples", "96", "--meas-every", "64", "--temps", "6",
             "--tmin", "0.5", "--tmax", "3.0", "--pt-every", "4"]


def _csg():
    import os
    return os.path.join(os.path.dirname(sg.__file__), "bin", "csg")


def test_cli_obs_bin_equals_obs_records_of_the_oracle(tmp_path):
    """The csg writer's float32 obs_s records (eamsim.cuh:55-62) are byte-equal
    to sg.obs_records applied to the oracle's integer records of the same
    run (H from std::stof, the handle's own ladder and the CLI's default
    seeds)."""
    import subprocess
    d = tmp_path / "08"
    d.mkdir()
    subprocess.run([_csg(), str(d), "0.25", "--sweeps", "256", *CLI_FLAGS], check=True,
                   cwd=tmp_path, capture_output=True)
    got = (d / "0.2500-obs.bin").read_bytes()
    probe = sg.Sim(L=8, n_samples=96, n_replicas=2, T_range=(0.5, 3.0, 6))
    betas = probe.betas
    probe.close()
    orc = oracle.OracleSim(8, 2, 2, betas, H=0.25, coupling_seed=CSEED, spin_seed=SSEED,
                           pt_every=4, meas_every=64)
    orc.init()
    orc.run(256)
    assert orc.n_meas() == 4
    want = b"".join(sg.obs_records(orc.meas(i), 8, 2, 0.25, 96).tobytes()
                    for i in range(orc.n_meas()))
    assert len(got) == 4 * 96 * 6 * 44
    assert got == want


def test_cli_interrupted_run_resumes_to_the_uninterrupted_obs_bin(tmp_path):
    """A run stopped after 5 records (checkpoints every 2 measurements, so the
    last checkpoint holds 4) leaves 5 records in obs.bin; the next invocation
    cuts obs.bin back to the checkpoint's length, finishes the interrupted
    run, and leaves exactly the obs.bin of an uninterrupted run; a further
    invocation then appends the same records to both."""
    import subprocess
    a, b = tmp_path / "a" / "08", tmp_path / "b" / "08"
    a.mkdir(parents=True)
    b.mkdir(parents=True)
    run = ["--sweeps", "512", "--checkpoint-every", "2", *CLI_FLAGS]
    subprocess.run([_csg(), str(a), "0.25", *run], check=True, cwd=tmp_path,
                   capture_output=True)
    r = subprocess.run([_csg(), str(b), "0.25", *run, "--stop-after-records", "5"],
                       cwd=tmp_path, capture_output=True)
    assert r.returncode != 0
    rec = 96 * 6 * 44
    assert len((b / "0.2500-obs.bin").read_bytes()) == 5 * rec
    assert list(tmp_path.glob("*.log"))          # a failed run keeps its log
    r = subprocess.run([_csg(), str(b), "0.25", *run], check=True, cwd=tmp_path,
                       capture_output=True)
    assert b"resuming an interrupted run at sweep 256 of 512" in r.stdout
    ra = (a / "0.2500-obs.bin").read_bytes()
    assert len(ra) == 8 * rec
    assert (b / "0.2500-obs.bin").read_bytes() == ra
    for d in (a, b):
        subprocess.run([_csg(), str(d), "0.25", *run], check=True, cwd=tmp_path,
                       capture_output=True)
    assert (a / "0.2500-obs.bin").read_bytes() == (b / "0.2500-obs.bin").read_bytes()
    assert len((a / "0.2500-obs.bin").read_bytes()) == 16 * rec
