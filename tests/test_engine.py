"""The batch / serving engine (sg_batch_*): every job's results are the same
alone and in a mix, bit for bit — whatever else runs beside it, whatever
the admission order, arena size or launch form — and match the oracle."""
import numpy as np
import pytest

import oracle
import sg
from helpers import geometric_betas, results_from_records

PT = 4


def _mix():
    """Mixed L, ladders, replica counts, sample counts, seeds, fields and
    sweep counts (all multiples of the PT cadence)."""
    return [
        sg.Job(L=8, n_sweeps=24, n_samples=64, betas=geometric_betas(0.5, 2.0, 4),
               coupling_seed=11, spin_seed=12, meas_every=8),
        sg.Job(L=12, n_sweeps=16, n_samples=100, betas=geometric_betas(0.2, 2.0, 6),
               coupling_seed=21, spin_seed=22, meas_every=4),
        sg.Job(L=16, n_sweeps=12, n_samples=64, T_range=(0.3, 2.0, 5),
               coupling_seed=31, spin_seed=32, meas_every=12),
        sg.Job(L=8, n_sweeps=40, n_samples=130, betas=geometric_betas(0.4, 1.5, 3), H=0.25,
               coupling_seed=41, spin_seed=42, meas_every=8),
        sg.Job(L=20, n_sweeps=8, n_samples=64, betas=geometric_betas(0.3, 2.0, 4),
               coupling_seed=51, spin_seed=52, meas_every=4),
        sg.Job(L=24, n_sweeps=8, n_samples=64, n_replicas=3, betas=geometric_betas(0.3, 2.0, 3),
               coupling_seed=61, spin_seed=62, meas_every=8),
        sg.Job(L=12, n_sweeps=20, n_samples=64, betas=geometric_betas(0.2, 2.0, 7),
               coupling_seed=71, spin_seed=72, meas_every=20),
        sg.Job(L=4, n_sweeps=32, n_samples=64, betas=geometric_betas(0.5, 2.0, 2),
               coupling_seed=81, spin_seed=82, meas_every=4),
    ]


def _alone(job):
    sim = sg.Sim(L=job.L, n_samples=job.n_samples, n_replicas=job.n_replicas, betas=job.betas,
                 T_range=job.T_range, H=job.H, coupling_seed=job.coupling_seed,
                 spin_seed=job.spin_seed, pt_every=PT, meas_every=job.meas_every)
    sim.keep_records = False
    sim.init_lattice()
    sim.run(job.n_sweeps)
    res = sim.results()
    acc, _ = sim.swap_counts()
    sim.close()
    return res, acc


@pytest.fixture(scope="module")
def alone():
    return [_alone(j) for j in _mix()]


def _check_same(got, ref):
    for (r, s), (r0, s0) in zip(got, ref):
        np.testing.assert_array_equal(r, r0)
        np.testing.assert_array_equal(s, s0)


@pytest.mark.gpu
@pytest.mark.parametrize("form", [sg.SG_FORM_BATCHED, sg.SG_FORM_PLAIN])
@pytest.mark.parametrize("slice_steps", [1, 3])
def test_mix_matches_each_job_alone(alone, form, slice_steps):
    b = sg.Batch(pt_every=PT, slice_pt_steps=slice_steps, form=form, arena_bytes=256 << 20)
    jobs = _mix()
    ids = [b.submit(j) for j in jobs[:4]]
    b.step()                       # the rest arrive while the first run
    ids += [b.submit(j) for j in jobs[4:]]
    b.drain()
    assert all(b.poll(i) == sg.JOB_DONE for i in ids)
    _check_same([b.result(i) for i in ids], alone)
    assert b.stats()["arena_in_use"] == 0
    b.close()


def _long_mix():
    """Jobs long enough for the engine to capture and replay its intervals
    (a class's tables must serve several intervals unchanged first): runs of
    4, 2 and 1 intervals between measurements, two jobs of one class that
    finish at different sweeps (the class is rebuilt mid-run), resident,
    one-lattice and streamed shapes, H != 0."""
    b4 = geometric_betas(0.3, 2.0, 4)
    return [
        sg.Job(L=8, n_sweeps=160, n_samples=64, betas=b4, coupling_seed=101, spin_seed=102,
               meas_every=12),
        sg.Job(L=8, n_sweeps=120, n_samples=96, betas=geometric_betas(0.4, 1.5, 3), H=0.25,
               coupling_seed=111, spin_seed=112, meas_every=40),
        sg.Job(L=12, n_sweeps=96, n_samples=64, betas=b4, coupling_seed=121, spin_seed=122,
               meas_every=16),
        sg.Job(L=12, n_sweeps=200, n_samples=64, betas=geometric_betas(0.2, 2.0, 5),
               coupling_seed=131, spin_seed=132, meas_every=24),
        sg.Job(L=20, n_sweeps=96, n_samples=64, betas=b4, coupling_seed=141, spin_seed=142,
               meas_every=32),
        sg.Job(L=24, n_sweeps=80, n_samples=64, betas=geometric_betas(0.3, 2.0, 3),
               coupling_seed=151, spin_seed=152, meas_every=16),
        sg.Job(L=32, n_sweeps=64, n_samples=64, betas=geometric_betas(0.3, 2.0, 2),
               coupling_seed=161, spin_seed=162, meas_every=32, kernel=sg.SG_KERNEL_STREAMED),
        sg.Job(L=16, n_sweeps=96, n_samples=100, betas=geometric_betas(0.3, 2.0, 5),
               coupling_seed=171, spin_seed=172, meas_every=24,
               kernel=sg.SG_KERNEL_RESIDENT_SINGLE),
        sg.Job(L=20, n_sweeps=64, n_samples=64, betas=geometric_betas(0.3, 2.0, 3), H=0.5,
               coupling_seed=181, spin_seed=182, meas_every=32,
               kernel=sg.SG_KERNEL_RESIDENT_SINGLE),
    ]


@pytest.mark.gpu
@pytest.mark.parametrize("slice_steps", [4, 3])
def test_captured_intervals_match_each_job_alone(slice_steps):
    jobs = _long_mix()
    ref = [_alone(j) for j in jobs]
    b = sg.Batch(pt_every=PT, slice_pt_steps=slice_steps, arena_bytes=512 << 20)
    ids = [b.submit(j) for j in jobs]
    b.drain()
    _check_same([b.result(i) for i in ids], ref)
    assert b.stats()["arena_in_use"] == 0
    ls = b.launch_stats()
    # every class interval ran (a class runs while any of its jobs does),
    # most of them through replays
    classes = {}
    for j in jobs:
        key = (j.L, j.H != 0.0, j.kernel)
        classes[key] = max(classes.get(key, 0), j.n_sweeps // PT)
    assert ls["intervals_direct"] + ls["intervals_graphed"] == sum(classes.values())
    assert ls["intervals_graphed"] > ls["intervals_direct"] and ls["graphs_built"] > 0
    b.close()


@pytest.mark.gpu
def test_small_arena_queues_jobs_and_keeps_results(alone):
    jobs = _mix()
    biggest = max(j.footprint() for j in jobs)
    b = sg.Batch(pt_every=PT, arena_bytes=int(biggest * 1.5))
    ids = [b.submit(j) for j in jobs]
    running, queued = b.step()
    assert queued > 0               # admission is limited by the footprint
    b.drain()
    _check_same([b.result(i) for i in ids], alone)
    b.close()


@pytest.mark.gpu
def test_engine_job_matches_oracle():
    betas = geometric_betas(0.5, 2.0, 4)
    job = sg.Job(L=8, n_sweeps=24, n_samples=80, betas=betas, coupling_seed=5, spin_seed=6,
                 meas_every=8)
    b = sg.Batch(pt_every=PT, arena_bytes=64 << 20)
    other = b.submit(sg.Job(L=12, n_sweeps=32, n_samples=64, betas=betas, meas_every=8))
    jid = b.submit(job)
    b.drain()
    res, acc = b.result(jid)
    b.result(other)
    orc = oracle.OracleSim(8, 2, 2, betas, coupling_seed=5, spin_seed=6, pt_every=PT,
                           meas_every=8)
    orc.init()
    orc.run(24)
    ref = results_from_records([orc.meas(i) for i in range(orc.n_meas())], 8, 2, 0.0, 80, betas)
    np.testing.assert_array_equal(res, ref)
    np.testing.assert_array_equal(acc, orc.swap_counts()[0])
    b.close()


@pytest.mark.gpu
def test_serve_loop_completes_every_job(alone):
    jobs = _mix()
    arrivals = np.cumsum(np.full(len(jobs), 0.002))
    b = sg.Batch(pt_every=PT, arena_bytes=256 << 20)
    ids, wall = b.serve(jobs, arrivals)
    assert wall > arrivals[-1]
    for i, a in zip(ids, arrivals):
        t_sub, t_adm, t_done = b.times(i)
        assert t_sub <= t_adm <= t_done
    t0 = b.times(ids[0])[0] - arrivals[0]
    for i, a in zip(ids, arrivals):
        assert abs(b.times(i)[0] - (t0 + a)) < 1e-9
    _check_same([b.result(i) for i in ids], alone)
    b.close()


def test_engine_rejects_bad_parameters_before_any_device_call():
    with pytest.raises(sg.SGError, match="pt_every"):
        sg.Batch(pt_every=0)
    with pytest.raises(sg.SGError, match="slice_pt_steps"):
        sg.Batch(slice_pt_steps=0)
    with pytest.raises(sg.SGError, match="arena"):
        sg.Batch(arena_bytes=16)


@pytest.mark.gpu
def test_engine_validates_jobs():
    b = sg.Batch(pt_every=PT, arena_bytes=1 << 20)
    with pytest.raises(sg.SGError, match="multiple"):
        b.submit(sg.Job(L=8, n_sweeps=10, betas=[1.0, 0.5]))
    with pytest.raises(sg.SGError, match="exceeds"):
        b.submit(sg.Job(L=48, n_sweeps=4, n_samples=64, betas=geometric_betas(0.3, 2.0, 8)))
    with pytest.raises(sg.SGError, match="L must"):
        b.submit(sg.Job(L=7, n_sweeps=4, betas=[1.0]))
    b.close()


def test_job_footprint_counts_every_block():
    sg.load_library()
    j = sg.Job(L=16, n_sweeps=16, n_samples=128, betas=geometric_betas(0.2, 2.0, 24))
    N, G, R, T = 16 ** 3, 2, 2, 24
    spins = 8 * G * R * T * N
    J = 8 * G * 3 * N
    site_records = 8 * G * 6 * N          # couplings per site, checkerboard order
    major = spins + J + site_records
    assert major < j.footprint() < 1.1 * major


@pytest.mark.gpu
def test_autotuner_measures_every_shape_and_caches():
    P, S, O = sg.SG_KERNEL_RESIDENT, sg.SG_KERNEL_STREAMED, sg.SG_KERNEL_RESIDENT_SINGLE
    X = sg.SG_KERNEL_SPLIT
    k, ms = sg.tune_shapes(8, 64)
    assert ms[P] > 0 and ms[O] > 0 and ms[S] == 0 and ms[X] == 0  # no streamed kernel below L=12
    assert k == min((P, O), key=lambda m: ms[m])
    k, ms = sg.tune_shapes(24, 64)                               # k_one: one shape resident
    assert ms[P] > 0 and ms[O] == 0 and ms[S] > 0 and ms[X] == 0 and k in (P, S)
    k, ms = sg.tune_shapes(32, 64)                               # split: 4 x 64 work-groups fit
    assert ms[P] > 0 and ms[O] > 0 and ms[S] > 0 and ms[X] > 0
    assert k == min((P, O, S, X), key=lambda m: ms[m])
    assert sg.tune_shapes(32, 60) == (k, ms)                     # same power-of-two bucket
    assert sg.tune_kernel(32, 64) == (k, ms[P], ms[S])
    # tiles: streamed and split have two at L = 32; the choice is the fastest
    # (shape, tile) and each shape's time is its best tile
    kt, t, tms = sg.tune_tiles(32, 64)
    assert kt == k and {(S, 0), (S, 1), (X, 0), (X, 1), (P, 0), (O, 0)} <= set(tms)
    assert (kt, t) == min(tms, key=tms.get)
    for m in (P, O, S, X):
        assert ms[m] == min(v for (mm, _), v in tms.items() if mm == m)
    sim = sg.Sim(L=32, n_samples=64, n_replicas=2, T_range=(0.3, 2.0, 32),
                 form=sg.SG_FORM_BATCHED)                        # SG_KERNEL_AUTO: the tuner's pick
    assert sim.shape == (kt, t)
    sim.close()


@pytest.mark.gpu
def test_mixed_benchmark_batch_at_full_size_matches_oracle():
    """The mixed-L benchmark batch at its full size: eight 64-sample jobs with
    L per group [20, 8, 20, 8, 20, 12, 20, 20] (its seeded draw), 32
    temperatures geometric in T = [0.2, 2], 2 replicas, PT every 16, in one
    batched engine (three classes on three streams, most intervals replayed
    from captured graphs).  Every job's FP64 results and swap counts equal
    the oracle's run of the same job.  Results: 1e-12 relative, with an
    absolute floor of 1e-13 for means that nearly cancel (m, about 7e-4 from
    terms of order 1: 256 such terms summed in another order differ by up to
    256 * 2^-53 = 3e-14 of the largest)."""
    Ls = [20, 8, 20, 8, 20, 12, 20, 20]
    betas = geometric_betas(0.2, 2.0, 32)
    sweeps, meas = 256, 64
    jobs = [sg.Job(L=L, n_sweeps=sweeps, n_samples=64, betas=betas,
                   coupling_seed=0x5EED0001 + g, spin_seed=0x5EED1001 + g, meas_every=meas)
            for g, L in enumerate(Ls)]
    b = sg.Batch(pt_every=16)
    ids = [b.submit(j) for j in jobs]
    b.drain()
    st = b.launch_stats()
    assert st["intervals_graphed"] > 0 and st["intervals_direct"] > 0
    for j, i in zip(jobs, ids):
        res, acc = b.result(i)
        orc = oracle.OracleSim(j.L, 1, 2, betas, coupling_seed=j.coupling_seed,
                               spin_seed=j.spin_seed, pt_every=16, meas_every=meas)
        orc.init()
        orc.run(sweeps)
        ref = results_from_records([orc.meas(k) for k in range(orc.n_meas())], j.L, 2, 0.0, 64,
                                   betas)
        np.testing.assert_array_equal(res, ref, err_msg=f"L={j.L}")
        np.testing.assert_array_equal(acc, orc.swap_counts()[0], err_msg=f"L={j.L}")
    b.close()


def _serving_stream(n, rate, sweeps_lo, sweeps_hi, seed):
    """A job stream shaped like the serving workload (Poisson arrivals at
    `rate`, L ~ U{8,12,16,20,24}, 16..32 temperatures geometric in [0.2, 2],
    its own coupling and spin seeds, sweeps log-uniform in [lo, hi] rounded to
    the PT cadence, a measurement every 1024 sweeps)."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
    arrivals -= arrivals[0]
    jobs = []
    for i in range(n):
        L = int(rng.choice((8, 12, 16, 20, 24)))
        sw = float(np.exp(rng.uniform(np.log(sweeps_lo), np.log(sweeps_hi))))
        sw = max(PT, int(round(sw / PT)) * PT)
        T = int(rng.integers(16, 33))
        jobs.append(sg.Job(L=L, n_sweeps=sw, n_samples=64, T_range=(0.2, 2.0, T),
                           coupling_seed=0x5EED2000 + i, spin_seed=0x5EED3000 + i,
                           meas_every=1024))
    return jobs, arrivals


@pytest.mark.gpu
def test_second_wave_of_jobs_allocates_nothing():
    """Twelve serving-shaped jobs submitted at once, served to completion, then
    the same twelve again on the same engine: the second wave grows no task
    table and allocates no pinned block (retired jobs gave theirs back), and
    its results and swap counts equal the first wave's bit for bit."""
    jobs, _ = _serving_stream(12, 200.0, 64, 256, 0x5EED0006)
    b = sg.Batch(pt_every=PT, slice_pt_steps=2, arena_bytes=4 << 30)
    ids, _ = b.serve(jobs, [0.0] * len(jobs))
    first = [b.result(i) for i in ids]
    a1 = b.alloc_stats()
    assert a1["pinned_blocks"] > 0 and a1["table_growths"] > 0
    ids, _ = b.serve(jobs, [0.0] * len(jobs))
    second = [b.result(i) for i in ids]
    assert b.alloc_stats() == a1
    b.close()
    for j, (r1, c1), (r2, c2) in zip(jobs, first, second):
        np.testing.assert_array_equal(r1, r2, err_msg=f"L={j.L}")
        np.testing.assert_array_equal(c1, c2, err_msg=f"L={j.L}")


@pytest.mark.gpu
def test_classes_are_paced_by_their_own_queues():
    """An L = 24 job and an L = 8 job of the same sweep count submitted
    together: each class's stream is refilled as its own slices finish, so
    the L = 8 job (intervals of a few tens of microseconds) is done long
    before the L = 24 one (about 150 us per interval) instead of moving at
    its pace; both results equal the jobs run alone."""
    def jobs(n):
        return [sg.Job(L=24, n_sweeps=n, n_samples=64, T_range=(0.3, 2.0, 32),
                       coupling_seed=91, spin_seed=92),
                sg.Job(L=8, n_sweeps=n, n_samples=64, T_range=(0.3, 2.0, 16),
                       coupling_seed=93, spin_seed=94)]

    def run(n):
        b = sg.Batch(pt_every=PT, slice_pt_steps=4, max_inflight=4, arena_bytes=1 << 30)
        ids = [b.submit(j) for j in jobs(n)]
        b.drain()
        took = [b.times(i)[2] - b.times(i)[0] for i in ids]
        got = [b.result(i) for i in ids]
        b.close()
        return took, got

    run(256)   # first launches of both classes' kernels and graphs
    took, got = run(4096)
    print(f"L=24 {took[0]:.4f} s, L=8 {took[1]:.4f} s", flush=True)
    assert took[1] < 0.7 * took[0]     # in step with the L = 24 class: about 1
    _check_same(got, [_alone(j) for j in jobs(4096)])


@pytest.mark.gpu
def test_serving_workload_jobs_equal_their_alone_runs():
    """The serving configuration, scaled to test time: 24 jobs arriving as a
    Poisson stream at 200/s (seed 0x5EED0005), L from {8..24}, 16-32
    temperatures, 1e3-4e3 sweeps, own seeds, served by sg_serve with continuous
    batching.  Every job's FP64 results and accepted swap counts are
    bit-identical to the same parameters run alone through sg_run; the
    smallest-L job also matches the oracle; the loop runs the tuner zero
    times."""
    jobs, arrivals = _serving_stream(24, 200.0, 1e3, 4e3, 0x5EED0005)
    assert len({j.L for j in jobs}) >= 4
    b = sg.Batch(pt_every=PT, slice_pt_steps=2, arena_bytes=4 << 30)
    tunes = sg.tune_count()
    ids, wall = b.serve(jobs, arrivals)
    assert sg.tune_count() == tunes
    lat = [b.times(i)[2] - b.times(i)[0] for i in ids]
    served = [b.result(i) for i in ids]     # releases each job
    print(f"serve: {len(jobs)} jobs in {wall:.3f} s, p50 {np.percentile(lat, 50):.4f} s, "
          f"p99 {np.percentile(lat, 99):.4f} s", flush=True)
    b.close()
    for j, (res, acc) in zip(jobs, served):
        s = sg.Sim(L=j.L, n_samples=64, n_replicas=2, T_range=j.T_range,
                   coupling_seed=j.coupling_seed, spin_seed=j.spin_seed, pt_every=PT,
                   meas_every=j.meas_every)
        s.keep_records = False
        s.init_lattice()
        s.run(j.n_sweeps)
        np.testing.assert_array_equal(res, s.results(), err_msg=f"L={j.L} sweeps={j.n_sweeps}")
        np.testing.assert_array_equal(acc, s.swap_counts()[0], err_msg=f"L={j.L}")
        s.close()
    k = int(np.argmin([j.L * 100000 + j.n_sweeps for j in jobs]))
    j = jobs[k]
    probe = sg.Sim(L=j.L, n_samples=64, n_replicas=2, T_range=j.T_range)
    betas = probe.betas
    probe.close()
    oracle.set_threads(8)
    orc = oracle.OracleSim(j.L, 1, 2, betas, coupling_seed=j.coupling_seed,
                           spin_seed=j.spin_seed, pt_every=PT, meas_every=j.meas_every)
    orc.init()
    orc.run(j.n_sweeps)
    ref = results_from_records([orc.meas(i) for i in range(orc.n_meas())], j.L, 2, 0.0, 64,
                               betas)
    np.testing.assert_array_equal(served[k][0], ref)
    np.testing.assert_array_equal(served[k][1], orc.swap_counts()[0])


def _random_mix(seed):
    """A seeded mix: 6-10 jobs of random L (2-40), sample counts (partial
    groups), replicas, ladders, fields, sweep counts, measurement cadences and
    requested shapes."""
    rng = np.random.default_rng(seed)
    sizes = [2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 32, 40]
    kernels = [sg.SG_KERNEL_AUTO, sg.SG_KERNEL_AUTO, sg.SG_KERNEL_RESIDENT,
               sg.SG_KERNEL_STREAMED, sg.SG_KERNEL_RESIDENT_SINGLE]
    jobs = []
    for _ in range(int(rng.integers(6, 11))):
        L = int(rng.choice(sizes))
        big = L >= 28
        T = int(rng.integers(1, 5 if big else 9))
        tmin = float(rng.uniform(0.2, 0.8))
        jobs.append(sg.Job(
            L=L, n_sweeps=PT * int(rng.integers(1, 8 if big else 40)),
            n_samples=int(rng.integers(1, 65 if big else 200)),
            n_replicas=int(rng.integers(1, 4)),
            betas=geometric_betas(tmin, float(rng.uniform(tmin + 0.1, 2.5)), T),
            H=float(rng.choice([0.0, 0.0, 0.125, 0.7, 2.0])),
            coupling_seed=int(rng.integers(1, 2**63)), spin_seed=int(rng.integers(1, 2**63)),
            meas_every=PT * int(rng.choice([0, 1, 2, 3, 5])),
            kernel=int(rng.choice(kernels))))
    return jobs, rng


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(6))
def test_random_mix_matches_each_job_alone(seed):
    """Seeded random mixes through the engine: random slice length, form and
    arena size (from all jobs at once to barely the largest, so admission
    queues), half the jobs submitted while the first half runs.  Every job's
    results and swap counts equal its alone run, bit for bit."""
    jobs, rng = _random_mix(0x5EED0200 + seed)
    foot = [j.footprint() for j in jobs]
    arena = int(max(foot) * 1.5) if rng.integers(0, 2) else int(sum(foot) * 1.25) + (1 << 20)
    b = sg.Batch(pt_every=PT, slice_pt_steps=int(rng.integers(1, 5)),
                 form=int(rng.choice([sg.SG_FORM_BATCHED, sg.SG_FORM_PLAIN])), arena_bytes=arena)
    ids = []
    for k, j in enumerate(jobs):
        try:
            ids.append(b.submit(j))
        except sg.SGError as e:       # a shape the engine lacks at this L: the tuner's
            assert "kernel" in str(e)
            j.kernel = sg.SG_KERNEL_AUTO
            ids.append(b.submit(j))
        if k == len(jobs) // 2:
            b.step()
    print(f"mix {seed}: " + ", ".join(f"L={j.L}/{j.n_samples}x{j.n_replicas}x{j.n_temps}"
                                      f"/{j.n_sweeps}" for j in jobs), flush=True)
    b.drain()
    got = [b.result(i) for i in ids]
    assert b.stats()["arena_in_use"] == 0
    b.close()
    _check_same(got, [_alone(j) for j in jobs])
