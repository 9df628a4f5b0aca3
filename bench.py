"""Spin-flip throughput of the parallel-tempering engine on one MI355X.

Default workload (--workload l16; per step = one PT interval of 16 sweeps
over the whole batch):
  3D Edwards-Anderson +-J, L=16, 24 temperatures geometric in T = [0.2, 2.0],
  2 replicas per sample, PT attempt every 16 sweeps (alternating even/odd
  pairs), a measurement every 1000 sweeps, hot start from seeded Philox.
  batched form: 2048 samples (32 groups of 64, one uint64 word per site) per
  launch, the interval captured as one launch sequence;
  plain form:   64 samples (one group), the per-group drop-in launches.
value = batched spin-flip attempts per ns (whole job, inputs resident in HBM).

Other workloads (one JSON line each, same keys):
  mixed  8 groups of 64 samples, L per group drawn from {8,12,16,20,24}
         (seed 0x5EED0003), 32 temperatures in [0.2, 2.0], 2 replicas, 1e5
         sweeps each, all in one batch engine (one launch per L class and PT
         interval) vs the same jobs launched one by one;
  large  L = 32 and 48, 64 samples, 32 temperatures in [0.3, 2.0], 2 replicas;
  serve  Poisson arrivals (200 jobs/s by default) of 64-sample jobs with
         L ~ U{8,12,16,20,24}, sweeps log-uniform in [1e3, 1e5], 16..32
         temperatures in [0.2, 2.0], own seeds; p50/p99 arrival-to-result
         latency and throughput of the native serving loop.

Prints one JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "spin-glass_amd"))

import numpy as np  # noqa: E402

import sg  # noqa: E402

L, T, R = 16, 24, 2
T_MIN, T_MAX = 0.2, 2.0
PT_EVERY, MEAS_EVERY = 16, 1000
SAMPLES_BATCHED, SAMPLES_PLAIN = 2048, 64
W_ALG = 1.22          # int32 VALU ops per spin-flip attempt (algorithmic)
# int32 VALU peak: 256 CU x 4 SIMD x 32 lanes x 2.4 GHz.  A v_bitop3_b32
# throughput probe on the MI355X sustains 70.6 T lane-ops/s (90 %), so the
# 16-lane-per-SIMD figure (39.3) understates the hardware by 2x.
VALU_PEAK_TOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes per k_block<16,false> launch of the l16 workload (768 two-lattice tasks,
# 16 sweeps): rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of that kernel, each
# scaled by what the counter reads for a device copy of known size (the
# --copy-probe run); profiles/k_block_L16_batched_2048_traffic.txt.
TRAFFIC_L16 = 109.0e6


def make(samples, form):
    s = sg.Sim(L=L, n_samples=samples, n_replicas=R, T_range=(T_MIN, T_MAX, T),
               coupling_seed=0x5EED0001, spin_seed=0x5EED0002, pt_every=PT_EVERY,
               meas_every=MEAS_EVERY, form=form)
    s.keep_records = False
    s.init_lattice()
    return s


SHAPE_NAMES = {sg.SG_KERNEL_RESIDENT: "lattice pairs", sg.SG_KERNEL_RESIDENT_SINGLE: "single lattices",
               sg.SG_KERNEL_STREAMED: "streamed chunks"}


def time_form(samples, form, steps, warmup, profile=False):
    """Times `steps` PT intervals.  profile=False: the whole job as it runs
    (batched: captured intervals replayed).  profile=True: every sweep-kernel
    launch bracketed by HIP events on the handle's stream (direct launches),
    for the kernel's own duration."""
    s = make(samples, form)
    s.run(warmup * PT_EVERY)
    s.set_profiling(profile)
    t0 = time.perf_counter()
    s.run(steps * PT_EVERY)          # sg_run synchronises the stream on return
    t1 = time.perf_counter()
    n, ms, fl = s.kernel_stats(0)
    flips = float(steps * PT_EVERY) * L ** 3 * 64 * s.G * R * T
    out = dict(wall_s=t1 - t0, flips=flips, launches=n, sweep_kernel_ms=ms,
               sweep_kernel_flips=fl)
    s.close()
    return out


def cpu_baseline(budget_s=10.0):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    cores = len(os.sched_getaffinity(0))
    cores = min(cores, 16)
    oracle.set_threads(cores)
    betas = 1.0 / (T_MIN * (T_MAX / T_MIN) ** (np.arange(T) / (T - 1)))
    o = oracle.OracleSim(L, 1, R, betas, pt_every=PT_EVERY, meas_every=MEAS_EVERY)
    o.init()
    o.run(PT_EVERY)
    sweeps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.run(PT_EVERY)
        sweeps += PT_EVERY
    dt = time.perf_counter() - t0
    flips = float(sweeps) * L ** 3 * 64 * R * T
    return dict(value=flips / dt / 1e9, unit="spin flips/ns", cores=cores, kind="port",
                sample=f"oracle/sg_oracle.c bit-sliced path, L=16, 64 samples x {T} T x {R} R, "
                       f"{sweeps} sweeps with PT every {PT_EVERY}, {cores} OpenMP threads, "
                       f"{dt:.1f} s")


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB (a known byte count for scaling
    the FETCH_SIZE / WRITE_SIZE counters); outside every timed region."""
    import torch
    n = mib << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=64,
                    help="untimed PT intervals first (every captured interval shape is "
                         "instantiated before the timed region)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--workload", default="l16", choices=["l16", "mixed", "large", "serve", "field"])
    ap.add_argument("--sweeps", type=int, default=0, help="sweeps per job (mixed, large, field)")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: arrivals per second")
    ap.add_argument("--jobs", type=int, default=400, help="serve: number of jobs")
    ap.add_argument("--field-h", type=float, default=0.0625, help="field: H")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (samples shard across GPUs with no "
              "exchange); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
    if args.workload != "l16":
        import bench_workloads
        line = bench_workloads.run(args)
        print(json.dumps(line), flush=True)
        return

    b = time_form(SAMPLES_BATCHED, sg.SG_FORM_BATCHED, args.steps, args.warmup)
    value = b["flips"] / b["wall_s"] / 1e9
    # the work-group shape the autotuner picked for each form (cached per lattice count)
    shape_b = sg.tune_shapes(L, (SAMPLES_BATCHED // 64) * R * T)
    shape_p = sg.tune_shapes(L, (SAMPLES_PLAIN // 64) * R * T)
    k = time_form(SAMPLES_BATCHED, sg.SG_FORM_BATCHED, args.steps, args.warmup, profile=True)
    plain = None
    if not args.no_plain:
        p = time_form(SAMPLES_PLAIN, sg.SG_FORM_PLAIN, args.steps, args.warmup)
        plain = p["flips"] / p["wall_s"] / 1e9
    avg_ms = k["sweep_kernel_ms"] / max(k["launches"], 1)
    hbm_gbs = TRAFFIC_L16 / (avg_ms * 1e6) if avg_ms > 0 else 0.0
    flips_per_launch = k["sweep_kernel_flips"] / max(k["launches"], 1)
    achieved = W_ALG * flips_per_launch / (avg_ms * 1e-3) / 1e12
    line = {
        "metric": "spin flips/ns at L=16, batched vs plain",
        "value": round(value, 3),
        "unit": "spin flips/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(b["wall_s"] * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded Philox couplings and hot-start spins)",
        "config": {
            "workload": f"3D EA +-J PT, L={L}, {SAMPLES_BATCHED} samples (uint64 words), "
                        f"{T} temperatures geometric in T=[{T_MIN},{T_MAX}], {R} replicas, "
                        f"PT every {PT_EVERY} sweeps, measurement every {MEAS_EVERY}; "
                        f"step = {PT_EVERY} sweeps",
            "form": "batched",
            "plain_flips_per_ns": None if plain is None else round(plain, 3),
            "plain_workload": f"same, {SAMPLES_PLAIN} samples (one group), per-group launches",
            "batched_over_plain": None if plain is None else round(value / plain, 3),
            "shape": {"batched": SHAPE_NAMES[shape_b[0]], "plain": SHAPE_NAMES[shape_p[0]],
                      "tuner_ms_batched": {SHAPE_NAMES[m]: round(v, 4)
                                           for m, v in shape_b[1].items() if v > 0},
                      "tuner_ms_plain": {SHAPE_NAMES[m]: round(v, 4)
                                         for m, v in shape_p[1].items() if v > 0}},
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "valu",
            "kernel": "k_block<16,false>",
            "achieved": round(achieved, 3),
            "peak": VALU_PEAK_TOPS,
            "unit": "TOP/s (int32 VALU, W_alg = 1.22 ops per flip attempt)",
            "frac": round(achieved / VALU_PEAK_TOPS, 4),
            "avg_launch_ms": round(avg_ms, 4),
            "flips_per_launch": flips_per_launch,
            "timing": f"HIP events around each of the {k['launches']} k_block launches on the "
                      f"handle's stream, in a second timed pass over the same {args.steps} "
                      f"steps (direct launches); value and ms_per_step come from the first "
                      f"pass, which replays the captured intervals without events",
            "traffic": TRAFFIC_L16,
            # the same launch against HBM: counter-measured bytes per launch
            # over the event-timed launch duration
            "hbm": {"achieved": round(hbm_gbs, 1), "peak": 8000.0, "unit": "GB/s",
                    "frac": round(hbm_gbs / 8000.0, 4)},
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
