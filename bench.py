"""Spin-flip throughput and serving latency of the parallel-tempering engine
on one MI355X.

Default workload (--workload l16; per step = one PT interval of 16 sweeps
over the whole batch):
  3D Edwards-Anderson +-J, L=16, 24 temperatures geometric in T = [0.2, 2.0],
  2 replicas per sample, PT attempt every 16 sweeps (alternating even/odd
  pairs), a measurement every 1000 sweeps, hot start from seeded Philox.
  batched form: 2048 samples (32 groups of 64, one uint64 word per site), one
  launch per interval, the intervals replayed from captured hipGraphs;
  plain form: the same 2048 samples with one launch per 64-sample group and
  interval, and the drop-in shape itself (64 samples, one group).
value = batched spin-flip attempts per ns (whole job, inputs resident in HBM).

The same line carries the serving loop (jobs of mixed L arriving as a Poisson
stream at 200 jobs/s and at a rate one GPU sustains, p50/p99 arrival-to-result
latency),
the VALU issue-rate probe behind the roofline's peak, and the CPU baseline
(the C oracle on the host's cores: the reference's CPU case, L=8 with 32
samples, and the L=16 workload, one thread and all threads).

Other workloads (one JSON line each, same keys):
  mixed  8 groups of 64 samples, L per group drawn from {8,12,16,20,24}
         (seed 0x5EED0003), 32 temperatures in [0.2, 2.0], 2 replicas, 1e5
         sweeps each, all in one batch engine vs the same jobs launched one by one;
  large  L = 32 and 48, 64 samples, 32 temperatures in [0.3, 2.0], 2 replicas;
  serve  the serving loop alone (--rate, --jobs);
  field  the default shape at H != 0 (--field-h).
--valu-probe prints the probe's JSON alone (for rocprofv3 counter runs).

Prints one JSON line.  --gpus N with N > 1 exits: the engine runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "spin-glass_amd"))
# The batch engine gives every (L, field) class its own stream; with HIP's
# default of four hardware queues per process the five classes of the
# serving workload share queues, and a class's launches wait behind
# another's.  At least eight queues (set before the runtime starts) keep them
# apart.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

import numpy as np  # noqa: E402

import sg  # noqa: E402

L, T, R = 16, 24, 2
T_MIN, T_MAX = 0.2, 2.0
PT_EVERY, MEAS_EVERY = 16, 1000
SAMPLES_BATCHED, SAMPLES_PLAIN = 2048, 64
W_ALG = 1.22          # int32 VALU ops per spin-flip attempt (algorithmic)
# int32 VALU peak of one MI355X: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz.  The
# probe (sg_probe_valu) measures what one launch of independent v_bitop3_b32
# chains sustains on the box and reports it beside this figure;
# profiles/valu_probe_bitop3_mad64.txt holds its counter run.
VALU_PEAK_TOPS = 39.3
HBM_PEAK_GBS = 8000.0
# HBM bytes per k_block<16,false> launch of the l16 workload (768 two-lattice tasks,
# 16 sweeps): rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of that kernel, each
# scaled by what the counter reads for a device copy of known size (the
# --copy-probe run); profiles/k_block_L16_batched_2048_traffic.txt.
TRAFFIC_L16 = 109.3e6
# VALU wave-instructions per k_block<16,false> launch of the l16 workload
# (SQ_INSTS_VALU, profiles/k_block_L16_batched_2048_counters.txt): x 64 lanes
# over the launch duration gives the issued lane-ops/s.
VALU_WAVE_INSTS_L16 = 77.47e6


def make(samples, form):
    s = sg.Sim(L=L, n_samples=samples, n_replicas=R, T_range=(T_MIN, T_MAX, T),
               coupling_seed=0x5EED0001, spin_seed=0x5EED0002, pt_every=PT_EVERY,
               meas_every=MEAS_EVERY, form=form)
    s.keep_records = False
    s.init_lattice()
    return s


SHAPE_NAMES = {sg.SG_KERNEL_RESIDENT: "lattice pairs", sg.SG_KERNEL_RESIDENT_SINGLE: "single lattices",
               sg.SG_KERNEL_STREAMED: "streamed chunks", sg.SG_KERNEL_SPLIT: "split lattices"}


CLOCK_WARM_STEPS = 160  # about 27 ms of the timed workload's own launches


def time_form(samples, form, steps, warmup, profile=False):
    """Times `steps` PT intervals after `warmup` untimed ones.

    Before them the handle runs CLOCK_WARM_STEPS intervals that are then
    discarded by re-initialising the lattice (same seeds: the warm-up and
    timed steps start from exactly the state they would have without it).
    The GPU's engine clock drops to about 2.2 GHz whenever the device idles
    for a few ms and climbs to about 2.45 GHz over roughly 7 ms of sustained
    sweep load (GRBM_GUI_ACTIVE per dispatch); without this, a short timed
    window after a short warm-up measures the climb.  Every captured launch
    sequence the warm-up and timed steps replay is instantiated before either
    (sg_prepare), and the graph count is checked unchanged after timing.
    profile=False: the whole job as it runs (batched: captured intervals
    replayed).  profile=True: every sweep-kernel launch bracketed by HIP
    events on the handle's stream (direct launches), for the kernel's own
    duration."""
    s = make(samples, form)
    s.run(CLOCK_WARM_STEPS * PT_EVERY)
    s.init_lattice()
    s.set_profiling(profile)
    s.prepare(warmup * PT_EVERY)
    s.prepare(steps * PT_EVERY, from_sweep=warmup * PT_EVERY)
    graphs = s.graph_count
    s.run(warmup * PT_EVERY)
    n0, ms0, fl0 = s.kernel_stats(0)
    t0 = time.perf_counter()
    s.run(steps * PT_EVERY)          # sg_run synchronises the stream on return
    t1 = time.perf_counter()
    if s.graph_count != graphs:
        raise RuntimeError("a graph was instantiated inside the timed region")
    n, ms, fl = s.kernel_stats(0)
    n, ms, fl = n - n0, ms - ms0, fl - fl0     # the timed steps' launches only
    flips = float(steps * PT_EVERY) * L ** 3 * 64 * s.G * R * T
    out = dict(wall_s=t1 - t0, flips=flips, launches=n, sweep_kernel_ms=ms,
               sweep_kernel_flips=fl, graphs=graphs)
    s.close()
    return out


# ------------------------------------------------------------------ CPU baseline
def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_rate(oracle, Lc, samples, n_temps, t_lo, t_hi, pt_every, meas_every, sweeps,
                 warm, threads, repeats):
    """Median over `repeats` runs of the oracle's flips/ns (samples counted,
    not lanes), timing the sweep loop only (after `warm` untimed sweeps)."""
    oracle.set_threads(threads)
    betas = 1.0 / (t_lo * (t_hi / t_lo) ** (np.arange(n_temps) / (n_temps - 1)))
    rates = []
    for _ in range(repeats):
        o = oracle.OracleSim(Lc, (samples + 63) // 64, R, betas, coupling_seed=0x5EED0001,
                             spin_seed=0x5EED0002, pt_every=pt_every, meas_every=meas_every)
        o.init()
        o.run(warm)
        t0 = time.perf_counter()
        o.run(sweeps)
        dt = time.perf_counter() - t0
        rates.append(float(sweeps) * Lc ** 3 * samples * R * n_temps / dt / 1e9)
        del o
    return float(np.median(rates))


def cpu_baseline():
    """The C oracle (the same bit-sliced algorithm and Philox numbers as the
    GPU, OpenMP over lattices with one PT synchronisation per step) on this
    host's cores: the reference's CPU case (L=8, 32 samples, 16 T in [0.5, 2], PT every sweep,
    a measurement every 100) and the L=16 twin of the default line (64
    samples, 24 T, PT every 16), each with one thread and with all the
    threads this process may use (OMP_NUM_THREADS when set, else the CPU
    affinity mask), the median of repeated runs."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    host_cpus = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", host_cpus) or host_cpus)
    threads = max(1, min(threads, host_cpus))
    c1 = dict(Lc=8, samples=32, n_temps=16, t_lo=0.5, t_hi=2.0, pt_every=1, meas_every=100)
    c2 = dict(Lc=16, samples=64, n_temps=T, t_lo=T_MIN, t_hi=T_MAX, pt_every=PT_EVERY,
              meas_every=MEAS_EVERY)
    t0 = time.perf_counter()
    r = {
        "l8_1t": _oracle_rate(oracle, **c1, sweeps=2000, warm=100, threads=1, repeats=3),
        "l8_all": _oracle_rate(oracle, **c1, sweeps=10000, warm=100, threads=threads,
                                    repeats=3),
        "l16_1t": _oracle_rate(oracle, **c2, sweeps=256, warm=32, threads=1, repeats=3),
        "l16_all": _oracle_rate(oracle, **c2, sweeps=1024, warm=96, threads=threads, repeats=5),
    }
    spent = time.perf_counter() - t0
    return dict(
        value=round(r["l16_all"], 3), unit="spin flips/ns", cores=threads, kind="port",
        cpu_model=_cpu_model(), host_cpus=host_cpus,
        single_thread=round(r["l16_1t"], 3),
        reference_cpu_case_L8={"single_thread": round(r["l8_1t"], 3), "all_threads": round(r["l8_all"], 3),
                 "workload": "L=8, 32 samples (half a 64-lane word; samples counted), 2 replicas, "
                             "16 T geometric in [0.5,2.0], PT every sweep, measurement every 100; "
                             "1e4 sweeps with all threads, 2e3 with one"},
        sample=f"oracle/sg_oracle.c bit-sliced path (gcc -O3 -march=x86-64-v3, OpenMP over "
               f"lattices): value = L=16, 64 samples x {T} T x {R} R, PT every {PT_EVERY}, "
               f"1024 sweeps after 96 untimed, {threads} threads, median of 5; single_thread = "
               f"the same, 256 sweeps, median of 3; reference_cpu_case_L8 as stated; {spent:.1f} s in all")


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


def valu_probe():
    """The VALU probe: v_bitop3_b32 and v_mad_u64_u32 chains over every CU,
    each one timed launch of about 8 ms after an untimed one; T lane-ops/s
    each (a 0.5 ms launch measured 10 % below an 8 ms one)."""
    out = {}
    for kind, name in ((0, "v_bitop3_b32"), (1, "v_mad_u64_u32")):
        ops, ms, rate = sg.probe_valu(kind=kind, iters=32768)
        out[name] = {"lane_ops": ops, "ms": round(ms, 4), "tops": round(rate / 1e12, 2)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=64,
                    help="untimed PT intervals first; every captured interval shape the timed "
                         "steps replay is instantiated before the timed region whatever this is")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--workload", default="l16", choices=["l16", "mixed", "large", "serve", "field"])
    ap.add_argument("--sweeps", type=int, default=0, help="sweeps per job (mixed, large, field)")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: arrivals per second")
    ap.add_argument("--jobs", type=int, default=400, help="serve: number of jobs")
    ap.add_argument("--field-h", type=float, default=0.0625, help="field: H")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB")
    ap.add_argument("--valu-probe", action="store_true", help="print the VALU probe alone")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the engine runs on one GPU only (samples shard across GPUs with no "
              "exchange); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.valu_probe:
        print(json.dumps({"valu_probe": valu_probe(), "peak_tops": VALU_PEAK_TOPS}), flush=True)
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
    if args.workload != "l16":
        import bench_workloads
        line = bench_workloads.run(args)
        print(json.dumps(line), flush=True)
        return

    probe = valu_probe()             # also brings the GPU to its clocks before timing
    b = time_form(SAMPLES_BATCHED, sg.SG_FORM_BATCHED, args.steps, args.warmup)
    value = b["flips"] / b["wall_s"] / 1e9
    # the work-group shape and tile the autotuner picked for each form (cached per lattice
    # count), with the device ms of every (shape, tile) it timed
    shape_b = sg.tune_tiles(L, (SAMPLES_BATCHED // 64) * R * T)
    shape_p = sg.tune_tiles(L, (SAMPLES_PLAIN // 64) * R * T)
    k = time_form(SAMPLES_BATCHED, sg.SG_FORM_BATCHED, args.steps, args.warmup, profile=True)
    plain = plain_same = None
    if not args.no_plain:
        p = time_form(SAMPLES_PLAIN, sg.SG_FORM_PLAIN, args.steps, args.warmup)
        plain = p["flips"] / p["wall_s"] / 1e9
        q = time_form(SAMPLES_BATCHED, sg.SG_FORM_PLAIN, args.steps, args.warmup)
        plain_same = q["flips"] / q["wall_s"] / 1e9
    avg_ms = k["sweep_kernel_ms"] / max(k["launches"], 1)
    hbm_gbs = TRAFFIC_L16 / (avg_ms * 1e6) if avg_ms > 0 else 0.0
    flips_per_launch = k["sweep_kernel_flips"] / max(k["launches"], 1)
    achieved = W_ALG * flips_per_launch / (avg_ms * 1e-3) / 1e12
    issued = VALU_WAVE_INSTS_L16 * 64 / (avg_ms * 1e-3) / 1e12
    serve = None
    if not args.no_serve:
        import bench_workloads
        bench_workloads.pretune()
        serve = {"rate_200": bench_workloads.serve_run(400, 200.0),
                 "rate_50": bench_workloads.serve_run(200, 50.0)}
    line = {
        "metric": "spin flips/ns at L=16, batched vs plain; p50/p99 job latency in serve loop",
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
            "graphs_prepared": b["graphs"],
            "clock_warmup": f"{CLOCK_WARM_STEPS} intervals on the timed handle before its warm-up, "
                            f"discarded by re-initialising the lattice (the engine clock climbs "
                            f"from about 2.2 to 2.45 GHz over ~7 ms of sustained load)",
            "plain_same_samples_flips_per_ns": None if plain_same is None else round(plain_same, 3),
            "batched_over_plain_same_samples": None if plain_same is None
            else round(value / plain_same, 3),
            "plain_flips_per_ns": None if plain is None else round(plain, 3),
            "plain_workload": f"same, {SAMPLES_PLAIN} samples (one group: the drop-in shape, "
                              f"where batched and plain are the same launch)",
            "batched_over_plain": None if plain is None else round(value / plain, 3),
            "shape": {"batched": f"{SHAPE_NAMES[shape_b[0]]}, tile {shape_b[1]}",
                      "plain": f"{SHAPE_NAMES[shape_p[0]]}, tile {shape_p[1]}",
                      "tuner_ms_batched": {f"{SHAPE_NAMES[m]}/{t}": round(v, 4)
                                           for (m, t), v in shape_b[2].items()},
                      "tuner_ms_plain": {f"{SHAPE_NAMES[m]}/{t}": round(v, 4)
                                         for (m, t), v in shape_p[2].items()}},
            "serve": serve,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "valu",
            "kernel": "k_block<16,false>",
            "achieved": round(achieved, 3),
            "peak": VALU_PEAK_TOPS,
            "unit": "TOP/s (int32 VALU, W_alg = 1.22 ops per flip attempt)",
            "frac": round(achieved / VALU_PEAK_TOPS, 4),
            "issued": round(issued, 3),
            "issued_frac": round(issued / VALU_PEAK_TOPS, 4),
            "issued_basis": f"{VALU_WAVE_INSTS_L16:.4g} VALU wave-instructions per launch "
                            f"(SQ_INSTS_VALU) x 64 lanes / the event-timed launch duration",
            "valu_probe": probe,
            "avg_launch_ms": round(avg_ms, 4),
            "flips_per_launch": flips_per_launch,
            "timing": f"HIP events around each of the {k['launches']} k_block launches on the "
                      f"handle's stream, in a second timed pass over the same {args.steps} "
                      f"steps (direct launches); value and ms_per_step come from the first "
                      f"pass, which replays the captured intervals without events",
            "traffic": TRAFFIC_L16,
            # the same launch against HBM: counter-measured bytes per launch
            # over the event-timed launch duration
            "hbm": {"achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(hbm_gbs / HBM_PEAK_GBS, 4)},
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
