"""Non-default bench.py workloads: mixed-L batch, large lattices, serving loop.

Each returns one dict with bench.py's keys (metric, value, unit, n_gpus,
steps, warmup, ms_per_step, higher_is_better, scaling, vs_baseline, dtype,
data, config, roofline).  Inputs are synthetic: seeded Philox couplings and
hot-start spins, as in the default workload.
"""
from __future__ import annotations

import time

import numpy as np

import sg

PT_EVERY = 16
VALU_PEAK_TOPS = 39.3
HBM_PEAK_GBS = 8000.0
W_ALG = 1.22
# Memory-side bytes per launch of the large-lattice kernels at 64 lattices (FETCH_SIZE +
# WRITE_SIZE, scaled by a copy of known size), keyed by (L, shape, tile): the split kernel per
# 16-sweep launch (most of it the boundary-plane exchange); profiles/k_split_L32_L48_64lattices.txt
TRAFFIC_LARGE = {(32, sg.SG_KERNEL_SPLIT, 0): 286.8e6, (48, sg.SG_KERNEL_SPLIT, 0): 771.3e6}
LARGE_KERNELS = {(32, sg.SG_KERNEL_SPLIT, 0): "k_split<32,4,1024,false,false> (16 sweeps per launch)",
                 (32, sg.SG_KERNEL_SPLIT, 1): "k_split<32,4,512,false,false> (16 sweeps per launch)",
                 (48, sg.SG_KERNEL_SPLIT, 0): "k_split<48,4,768,true,false> (16 sweeps per launch)",
                 (32, sg.SG_KERNEL_STREAMED, 0): "k_slab<32,false,false,1> (one sweep per launch)",
                 (32, sg.SG_KERNEL_STREAMED, 1): "k_slab<32,false,false,2> (one sweep per launch)",
                 (48, sg.SG_KERNEL_STREAMED, 0): "k_phase<48,false,false,1> (one colour per launch)",
                 (48, sg.SG_KERNEL_STREAMED, 1): "k_phase<48,false,false,2> (one colour per launch)"}
L_CHOICES = (8, 12, 16, 20, 24)


def _base(metric, value, unit, wall_s, config, roofline, steps=0, warmup=0, higher=True,
          scaling="strong"):
    return {
        "metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1,
        "steps": steps, "warmup": warmup,
        "ms_per_step": round(wall_s * 1e3 / max(steps, 1), 4),
        "higher_is_better": higher, "scaling": scaling, "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded Philox couplings and hot-start spins)",
        "config": config, "roofline": roofline,
    }


def _valu_roofline(flips_per_ns, kernel):
    achieved = W_ALG * flips_per_ns * 1e9 / 1e12
    return {"bound": "valu", "kernel": kernel, "achieved": round(achieved, 3),
            "peak": VALU_PEAK_TOPS,
            "unit": "TOP/s (int32 VALU, W_alg = 1.22 ops per flip attempt; whole-job rate)",
            "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": None}


def _run_batch(jobs, form, arena=8 << 30):
    b = sg.Batch(pt_every=PT_EVERY, slice_pt_steps=4, form=form, arena_bytes=arena)
    t0 = time.perf_counter()
    ids = [b.submit(j) for j in jobs]
    b.drain()
    dt = time.perf_counter() - t0
    for i in ids:
        b.result(i)
    b.close()
    return dt


# ------------------------------------------------------------------ mixed L
def mixed(args):
    rng = np.random.default_rng(0x5EED0003)
    Ls = [int(x) for x in rng.choice(L_CHOICES, size=8)]
    sweeps = args.sweeps or 100000
    sweeps -= sweeps % PT_EVERY

    def jobs(n):
        return [sg.Job(L=L, n_sweeps=n, n_samples=64, T_range=(0.2, 2.0, 32),
                       coupling_seed=0x5EED0001 + g, spin_seed=0x5EED1001 + g, meas_every=1024)
                for g, L in enumerate(Ls)]

    _run_batch(jobs(64), sg.SG_FORM_BATCHED)            # warm-up: code paths, arena
    flips = sum(j.flips() for j in jobs(sweeps))
    dt = _run_batch(jobs(sweeps), sg.SG_FORM_BATCHED)
    value = flips / dt / 1e9
    plain = None
    if not args.no_plain:
        ps = max(PT_EVERY, (sweeps // 10) - (sweeps // 10) % PT_EVERY)
        pflips = sum(j.flips() for j in jobs(ps))
        pdt = _run_batch(jobs(ps), sg.SG_FORM_PLAIN)
        plain = pflips / pdt / 1e9
    cfg = {"workload": f"8 groups x 64 samples, L per group {Ls} (drawn from {list(L_CHOICES)}, "
                       f"seed 0x5EED0003), 32 temperatures geometric in T=[0.2,2.0], 2 replicas, "
                       f"PT every {PT_EVERY}, measurement every 1024, {sweeps} sweeps per group",
           "form": "batched engine: one launch per (L, field) class and PT interval, classes on "
                   "separate streams, intervals of a stable class replayed from hipGraphs",
           "plain_flips_per_ns": None if plain is None else round(plain, 3),
           "plain_workload": "same jobs, one launch per job and interval (sweeps / 10)",
           "batched_over_plain": None if plain is None else round(value / plain, 3),
           "parallelism": "1 GPU"}
    return _base("spin flips/ns, mixed-L batch", value, "spin flips/ns", dt, cfg,
                 _valu_roofline(value, "k_block<8|12|16>, k_mid<20>, k_one<24> (per tuner)"),
                 steps=sweeps // PT_EVERY)


# ------------------------------------------------------------------ large L
def large(args):
    sweeps = args.sweeps or 1024
    sweeps -= sweeps % PT_EVERY
    per_l = {}
    shapes = {}
    tot_flips = tot_s = 0.0
    for L in (32, 48):
        s = sg.Sim(L=L, n_samples=64, n_replicas=2, T_range=(0.3, 2.0, 32), pt_every=PT_EVERY,
                   meas_every=1024, form=sg.SG_FORM_BATCHED)
        shapes[L] = s.shape          # the (shape, tile) SG_KERNEL_AUTO picked (the tuner's)
        s.keep_records = False
        s.init_lattice()
        s.run(2 * PT_EVERY)
        t0 = time.perf_counter()
        s.run(sweeps)                # batched form: each PT interval replays one hipGraph
        dt = time.perf_counter() - t0
        s.close()
        flips = float(L) ** 3 * 64 * 32 * 2 * sweeps
        fpn = flips / dt / 1e9
        # algorithmic HBM bytes: each spin word read and written once per sweep
        # (16 B per 64 flips), couplings 24 B per site per group and sweep
        # shared by the 64 lattices of the group
        bytes_per_flip = 16.0 / 64 + 24.0 / 64 / 64
        gbs = fpn * bytes_per_flip
        per_l[L] = {"flips_per_ns": round(fpn, 3),
                    "ms_per_interval": round(dt * 1e3 * PT_EVERY / sweeps, 4),
                    "hbm_alg_GBs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        tot_flips += flips
        tot_s += dt
    value = tot_flips / tot_s / 1e9
    cfg = {"workload": f"L=32 and L=48 (separately), 64 samples, 32 temperatures geometric in "
                       f"T=[0.3,2.0], 2 replicas, PT every {PT_EVERY}, {sweeps} sweeps each",
           "per_L": per_l, "form": "batched", "parallelism": "1 GPU"}
    r = _valu_roofline(value, ", ".join(LARGE_KERNELS.get((L, *shapes[L]), f"L={L} shape "
                                                           f"{shapes[L]}") for L in per_l))
    r["hbm"] = {L: v["hbm_frac"] for L, v in per_l.items()}
    r["traffic"] = {L: TRAFFIC_LARGE.get((L, *shapes[L])) for L in per_l}
    return _base("spin flips/ns, L=32 and 48", value, "spin flips/ns", tot_s, cfg, r,
                 steps=2 * sweeps // PT_EVERY)


# ------------------------------------------------------------------ serving
def serve_jobs(n, rate, seed=0x5EED0005):
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
    arrivals -= arrivals[0]
    jobs = []
    for i in range(n):
        L = int(rng.choice(L_CHOICES))
        sw = float(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
        sw = max(PT_EVERY, int(round(sw / PT_EVERY)) * PT_EVERY)
        T = int(rng.integers(16, 33))
        jobs.append(sg.Job(L=L, n_sweeps=sw, n_samples=64, T_range=(0.2, 2.0, T),
                           coupling_seed=0x5EED2000 + i, spin_seed=0x5EED3000 + i,
                           meas_every=1024))
    return jobs, arrivals


def pretune(field=False, max_lattices=8192):
    """Fills the tuner's cache for the serving mix (every L of L_CHOICES,
    power-of-two lattice counts from 32 up to max_lattices) before a serve
    run, so that the engine's classes take measured shapes; the engine itself
    never tunes."""
    for L in L_CHOICES:
        n = 32
        while n <= max_lattices:
            sg.tune_shapes(L, n, field=field)
            n *= 2


# serving engine: PT intervals per step and per class slice, slices queued per
# class ahead of the host
SERVE_SLICE, SERVE_INFLIGHT = 4, 4


def serve_run(n_jobs, rate, seed=0x5EED0005):
    """One run of the native serving loop over the seeded job stream: jobs
    arrive (Poisson, `rate` per second), are admitted by footprint, run in
    continuous batches and are collected; returns the latency percentiles and
    throughput."""
    jobs, arrivals = serve_jobs(n_jobs, rate, seed)
    b = sg.Batch(pt_every=PT_EVERY, slice_pt_steps=SERVE_SLICE, form=sg.SG_FORM_BATCHED,
                 arena_bytes=16 << 30, max_inflight=SERVE_INFLIGHT)
    w = [b.submit(sg.Job(L=L, n_sweeps=64, T_range=(0.2, 2.0, 16))) for L in L_CHOICES]
    b.drain()
    for i in w:
        b.result(i)
    tunes = sg.tune_count()
    ids, wall = b.serve(jobs, arrivals)
    if sg.tune_count() != tunes:
        raise RuntimeError("the tuner ran inside the serving loop")
    lat = []
    for i in ids:
        t_sub, t_adm, t_done = b.times(int(i))
        lat.append(t_done - t_sub)
        b.result(int(i))
    b.close()
    lat = np.array(lat)
    flips = sum(j.flips() for j in jobs)
    offered = flips / arrivals[-1] / 1e9 if arrivals[-1] > 0 else float("nan")
    # the floor under the tail: the stream's longest job (most sweeps x L^3) run
    # alone on an idle engine, submission to result
    big = max(range(n_jobs), key=lambda i: jobs[i].n_sweeps * jobs[i].L ** 3)
    j = jobs[big]
    b = sg.Batch(pt_every=PT_EVERY, slice_pt_steps=SERVE_SLICE, form=sg.SG_FORM_BATCHED,
                 arena_bytes=16 << 30, max_inflight=SERVE_INFLIGHT)
    i = b.submit(sg.Job(L=j.L, n_sweeps=j.n_sweeps, n_samples=64, T_range=j.T_range,
                        coupling_seed=j.coupling_seed, spin_seed=j.spin_seed,
                        meas_every=1024))
    b.drain()
    t_sub, _, t_done = b.times(int(i))
    b.result(i)
    b.close()
    return {"jobs": n_jobs, "rate_per_s": rate,
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "mean_latency_s": round(float(lat.mean()), 4),
            "served_flips_per_ns": round(flips / wall / 1e9, 1),
            "offered_flips_per_ns": round(offered, 1),
            "arrival_window_s": round(float(arrivals[-1]), 3),
            "wall_s": round(wall, 3),
            "longest_job_alone_s": round(t_done - t_sub, 4),
            "longest_job": f"L={j.L}, {j.n_sweeps} sweeps, {j.T_range[2]} T, 64 samples"}


def serve(args):
    pretune()
    r = serve_run(args.jobs, args.rate)
    cfg = {"workload": f"{args.jobs} jobs, Poisson arrivals at {args.rate:g}/s (seed 0x5EED0005), "
                       f"64 samples each, L ~ U{list(L_CHOICES)}, sweeps log-uniform in [1e3,1e5], "
                       f"T ~ U{{16..32}} geometric in [0.2,2.0], 2 replicas, PT every {PT_EVERY}",
           "form": "native serve loop: admission by footprint, continuous batching per L class",
           "parallelism": "1 GPU", **{k: v for k, v in r.items() if k not in ("jobs", "rate_per_s")}}
    return _base("spin flips/ns and p50/p99 job latency, serving loop", r["served_flips_per_ns"],
                 "spin flips/ns", r["wall_s"], cfg,
                 _valu_roofline(r["served_flips_per_ns"],
                                "k_block<8|12|16>, k_mid<20>, k_one<24> (per tuner)"),
                 steps=args.jobs)


# ------------------------------------------------------------------ field
def field(args):
    """The default line's shape at H != 0 (the reference publishes L = 16 at
    H = 1/16, 2/16, 3/16): L=16, 2048 samples, its 24-temperature ladder,
    2 replicas, PT every 16, batched form."""
    H = args.field_h
    sweeps = args.sweeps or 1024 * 4
    sweeps -= sweeps % PT_EVERY
    s = sg.Sim(L=16, n_samples=2048, n_replicas=2, betas=sg.REF_BETAS, H=H, pt_every=PT_EVERY,
               meas_every=1024, form=sg.SG_FORM_BATCHED)
    s.keep_records = False
    s.init_lattice()
    s.run(64 * PT_EVERY)
    t0 = time.perf_counter()
    s.run(sweeps)
    dt = time.perf_counter() - t0
    s.close()
    flips = 16.0 ** 3 * 64 * 32 * 2 * 24 * sweeps
    value = flips / dt / 1e9
    cfg = {"workload": f"L=16, 2048 samples, the reference's 24 temperatures, 2 replicas, H={H}, "
                       f"PT every {PT_EVERY}, measurement every 1024, {sweeps} sweeps",
           "form": "batched", "parallelism": "1 GPU"}
    return _base(f"spin flips/ns, L=16 at H={H}", value, "spin flips/ns", dt, cfg,
                 _valu_roofline(value, "k_block<16,true>"), steps=sweeps // PT_EVERY)


def run(args):
    return {"mixed": mixed, "large": large, "serve": serve, "field": field}[args.workload](args)
