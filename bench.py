#!/usr/bin/env python3
"""MEMD throughput / latency benchmark on one MI355X.

Metric: MEMD epochs/s with p50/p99 latency on 32 ch x 4096 samples, K = 64
directions, the reference's schedule (9 output slots, 10 sifts per IMF, fix
mode, reference extrema window), FP64.  A step is one batched launch sequence
that decomposes `--batch` independent synthetic epochs already resident in HBM;
`value` is epochs per second over the timed steps.  The plain per-epoch path
(one memd_run_f64 call per epoch) and the FP64 CPU baseline (the host oracle on
this machine's cores, a bounded sample) are reported beside it.

    python bench.py [--steps K] [--warmup W] [--batch B]
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "memd-gpu_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector (the matrix cores add no FP64 rate)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64, help="epochs per step")
    ap.add_argument("--plain-epochs", type=int, default=12)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--T", type=int, default=4096)
    ap.add_argument("--K", type=int, default=64)
    ap.add_argument("--n-imf", type=int, default=9)
    ap.add_argument("--sifts", type=int, default=10)
    ap.add_argument("--cpu-reps", type=int, default=5,
                    help="eeg32x30504 / wide128x16384: whole CPU epochs timed (the median is reported)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="uneven512: skip the plain per-epoch pass (profiling runs)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in tile shapes (default: memd_autotune before the untimed warmup)")
    ap.add_argument("--tuning", type=str, default=None,
                    help="BLOCK_GROUPS,EVAL_TILE_WORDS: install these shapes for --n instead of autotuning "
                         "(profiling runs repeat a bench run's choice without the tuner's own kernels)")
    ap.add_argument("--workload", choices=("batched32x4096", "eeg32x30504", "wide128x16384", "uneven512", "served"),
                    default="batched32x4096",
                    help="batched32x4096: 64 epochs of 32ch x 4096 K=64 per batched call (headline); eeg32x30504: "
                         "32ch x 30504 K=64 single epochs vs the CPU oracle; wide128x16384: 128ch x 16384 K=256 "
                         "single epochs vs the CPU oracle; uneven512: 512 uneven epochs batched vs plain; served: "
                         "continuous batching under Poisson arrivals")
    ap.add_argument("--rates", type=float, nargs="+", default=[10.0, 25.0, 40.0],
                    help="served: arrival rates (epochs/s) to sweep")
    ap.add_argument("--duration", type=float, default=20.0, help="served: seconds of arrivals per rate")
    ap.add_argument("--pool-gb", type=int, default=64, help="served: server HBM pool (GiB)")
    ap.add_argument("--max-active", type=int, default=128, help="served: cap on epochs advanced together")
    ap.add_argument("--mix", choices=("c4", "c3"), default="c4",
                    help="served: c4 = 90%% uneven epochs + 10%% 32ch x 60000; c3 = uneven epochs only")
    ap.add_argument("--saturate", type=int, default=0,
                    help="served: also submit this many epochs of the mix at once, drain them through the server "
                         "and time the same epochs as one batched call")
    ap.add_argument("--mode", choices=("fix", "stop"), default="fix",
                    help="batched32x4096: sifting policy (stop = memd.m criterion 0.075/0.75/0.075, at most --sifts)")
    ap.add_argument("--boundary", choices=("clamp", "mirror"), default="clamp",
                    help="batched32x4096: end treatment (mirror = memd.m boundary_conditions)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain device copy of this many bytes (counter calibration) and exit")
    return ap.parse_args()


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    ent = rec.get(workload, {}).get(kernel)
    return int(ent["bytes_per_launch"]) if ent else None


def quiesce_gc():
    """Collect once and move every object allocated so far (torch's modules, the
    inputs) out of the collector's reach: a full collection over them takes about
    40 ms of host time, and left enabled one lands inside the timed calls."""
    gc.collect()
    gc.freeze()


def pct(xs, q):
    return float(np.percentile(np.asarray(xs, float), q)) if len(xs) else None


def cpu_info():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"model": model, "nproc": os.cpu_count(), "affinity": aff}


def cpu_threads():
    """Threads for the CPU legs: the CPUs this process may use, at most 16 (one GPU's share)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, aff))


def cpu_whole_epochs(gen, n, K, n_imf, S, reps, threads, **opts):
    """Median seconds of `reps` whole epochs (distinct inputs from gen(i)), FP64
    oracle with OpenMP over channels, after a one-sift warm-up."""
    from oracle import memd_oracle
    V = memd_oracle.directions(n, K)
    memd_oracle.memd(gen(0), V, 2, 1, threads=threads, trace=False, **opts)
    times = []
    for r in range(reps):
        y = gen(1 + r)
        t0 = time.perf_counter()
        memd_oracle.memd(y, V, n_imf, S, threads=threads, trace=False, **opts)
        times.append(time.perf_counter() - t0)
        print(f"cpu: whole epoch {r + 1}/{reps} of {n} x {y.shape[1]} in {times[-1]:.1f} s", file=sys.stderr, flush=True)
    return float(np.median(times)), times


def cpu_epoch_parallel(ys_Vs, n_imf, S, threads, **opts):
    """Epochs per second of the oracle's epoch-parallel batch (one epoch per thread)."""
    from oracle import memd_oracle
    t0 = time.perf_counter()
    memd_oracle.memd_batch(ys_Vs, n_imf, S, threads=threads, **opts)
    return len(ys_Vs) / (time.perf_counter() - t0)


def cpu_baseline(n, T, K, n_imf, S, **opts):
    """The FP64 oracle on the same workload, whole epochs, three ways: OpenMP over
    channels (median of 5 epochs after a warm-up), epochs in parallel (one per
    thread, as a batched CPU run would), and one thread.  value = the faster of the
    two parallel throughputs."""
    from oracle import memd_oracle
    import synth
    threads = cpu_threads()
    gen = lambda i: synth.sines_epoch(n, T, 5000 + i)
    med, times = cpu_whole_epochs(gen, n, K, n_imf, S, 5, threads, **opts)
    chan = 1.0 / med
    V = memd_oracle.directions(n, K)
    par = cpu_epoch_parallel([(gen(100 + i), V) for i in range(threads)], n_imf, S, threads, **opts)
    t0 = time.perf_counter()
    memd_oracle.memd(gen(200), V, n_imf, S, threads=1, trace=False, **opts)
    single = 1.0 / (time.perf_counter() - t0)
    best = max(chan, par)
    return {"value": best, "unit": "epochs/s", "cores": threads, "kind": "port",
            "sample": f"whole epochs of the same workload, FP64 oracle: median of 5 with OpenMP over channels "
                      f"({chan:.3f}/s), {threads} epochs in parallel one per thread ({par:.3f}/s), "
                      f"one epoch on one thread ({single:.4f}/s); value = the faster parallel form",
            "channel_parallel_epochs_per_s": round(chan, 4), "channel_parallel_times_s": [round(t, 3) for t in times],
            "epoch_parallel_epochs_per_s": round(par, 4), "single_core_epochs_per_s": round(single, 5),
            "cpu": cpu_info()}


def bench_eeg(a):
    """32 ch x 30504 samples (EEG-like generator), K = 64, 9 slots x 10 sifts:
    plain per-epoch GPU path vs the FP64 CPU oracle on this host."""
    import torch
    import memd_amd
    import synth
    from oracle import memd_oracle
    n, T, K, N_imf, S = 32, 30504, 64, 9, 10
    ctx = memd_amd.MemdContext(0)
    V = ctx.directions(n, K)
    ys = [torch.from_numpy(synth.eeg_epoch(n, T, 3000 + i)).to("cuda:0") for i in range(max(a.steps, 1))]
    out = torch.empty((N_imf, n, T), dtype=torch.float64, device="cuda:0")
    quiesce_gc()
    for _ in range(a.warmup):
        ctx.run(ys[0], V, N_imf, S, out=out)
    torch.cuda.synchronize()
    lat = []
    for y in ys:
        s0 = time.perf_counter()
        ctx.run(y, V, N_imf, S, out=out)
        lat.append(time.perf_counter() - s0)
    gpu_eps = len(lat) / sum(lat)
    ctx.close()
    if a.no_cpu:
        print(json.dumps({"metric": "plain GPU only (--no-cpu)", "value": round(gpu_eps, 3), "unit": "epochs/s",
                          "ms_per_step": round(sum(lat) / len(lat) * 1e3, 3)}), flush=True)
        return
    # CPU: whole epochs of the same shape (median after a one-sift warm-up), and one
    # sift of IMF 1 on one thread for the single-core rate
    threads = cpu_threads()
    med, times = cpu_whole_epochs(lambda i: synth.eeg_epoch(n, T, 3000 + i), n, K, N_imf, S, a.cpu_reps, threads)
    cpu_eps = 1.0 / med
    Vh = memd_oracle.directions(n, K)
    t0 = time.perf_counter()
    memd_oracle.memd(synth.eeg_epoch(n, T, 3000), Vh, 2, 1, threads=1, trace=False)
    single_sift = time.perf_counter() - t0
    print(json.dumps({
        "metric": "MEMD epochs/s, 32ch x 30504, K=64 (plain GPU vs FP64 CPU)", "value": round(gpu_eps, 3),
        "unit": "epochs/s", "n_gpus": 1, "steps": len(lat), "warmup": a.warmup,
        "ms_per_step": round(sum(lat) / len(lat) * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "memd_32ch_x30504_K64_fix9x10_plain", "generator": "eeg pink+alpha, fs=512"},
        "latency_ms": {"p50": round(pct(lat, 50) * 1e3, 3), "p99": round(pct(lat, 99) * 1e3, 3)},
        "cpu_baseline": {"value": round(cpu_eps, 5), "unit": "epochs/s", "cores": threads, "kind": "port",
                         "sample": f"median of {len(times)} whole epochs of the same shape after a one-sift "
                                   f"warm-up, FP64 oracle, OpenMP over channels with {threads} threads",
                         "times_s": [round(t, 2) for t in times],
                         "single_core_one_sift_s": round(single_sift, 2), "cpu": cpu_info()},
        "gpu_over_cpu": round(gpu_eps / cpu_eps, 1)}), flush=True)


def bench_wide(a):
    """128 ch x 16384 samples (EEG-like generator + 40 Hz on a quarter of the
    channels), K = 256, 9 slots x 10 sifts: plain per-epoch GPU path vs the FP64
    CPU oracle on this host."""
    import torch
    import memd_amd
    import synth
    from oracle import memd_oracle
    n, T, K, N_imf, S = 128, 16384, 256, 9, 10
    ctx = memd_amd.MemdContext(0)
    tune_s, tuning = 0.0, None
    if not a.no_autotune:
        t0 = time.perf_counter()
        tuning = ctx.autotune(n, T, K)
        tune_s = time.perf_counter() - t0
    V = ctx.directions(n, K)
    ys = [torch.from_numpy(synth.eeg_epoch(n, T, 4000 + i, gamma40=True)).to("cuda:0")
          for i in range(max(a.steps, 1))]
    out = torch.empty((N_imf, n, T), dtype=torch.float64, device="cuda:0")
    quiesce_gc()
    for _ in range(a.warmup):
        ctx.run(ys[0], V, N_imf, S, out=out)
    torch.cuda.synchronize()
    lat = []
    for y in ys:
        s0 = time.perf_counter()
        ctx.run(y, V, N_imf, S, out=out)
        lat.append(time.perf_counter() - s0)
    gpu_eps = len(lat) / sum(lat)
    ctx.close()
    if a.no_cpu:
        print(json.dumps({"metric": "plain GPU only (--no-cpu)", "value": round(gpu_eps, 3), "unit": "epochs/s",
                          "ms_per_step": round(sum(lat) / len(lat) * 1e3, 3)}), flush=True)
        return
    # CPU: whole epochs of the same shape (median after a one-sift warm-up), and one
    # sift of IMF 1 on one thread for the single-core rate
    threads = cpu_threads()
    med, times = cpu_whole_epochs(lambda i: synth.eeg_epoch(n, T, 4000 + i, gamma40=True), n, K, N_imf, S,
                                  a.cpu_reps, threads)
    cpu_eps = 1.0 / med
    Vh = memd_oracle.directions(n, K)
    t0 = time.perf_counter()
    memd_oracle.memd(synth.eeg_epoch(n, T, 4000, gamma40=True), Vh, 2, 1, threads=1, trace=False)
    single_sift = time.perf_counter() - t0
    print(json.dumps({
        "metric": "MEMD epochs/s, 128ch x 16384, K=256 (plain GPU vs FP64 CPU)", "value": round(gpu_eps, 3),
        "unit": "epochs/s", "n_gpus": 1, "steps": len(lat), "warmup": a.warmup,
        "ms_per_step": round(sum(lat) / len(lat) * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "memd_128ch_x16384_K256_fix9x10_plain",
                   "generator": "eeg pink+alpha, fs=512, 40 Hz on a quarter of the channels"},
        "latency_ms": {"p50": round(pct(lat, 50) * 1e3, 3), "p99": round(pct(lat, 99) * 1e3, 3)},
        "tuning": tuning, "autotune_s": round(tune_s, 3),
        "cpu_baseline": {"value": round(cpu_eps, 6), "unit": "epochs/s", "cores": threads, "kind": "port",
                         "sample": f"median of {len(times)} whole epochs of the same shape after a one-sift "
                                   f"warm-up, FP64 oracle, OpenMP over channels with {threads} threads",
                         "times_s": [round(t, 2) for t in times],
                         "single_core_one_sift_s": round(single_sift, 2), "cpu": cpu_info()},
        "gpu_over_cpu": round(gpu_eps / cpu_eps, 1)}), flush=True)


def bench_uneven(a):
    """512 independent epochs, n in {8,16,32,64}, T in [1000,8000], K = 2n (EEG-like
    generator, per-epoch seeds), 9 slots x 10 sifts: one batched call vs one call
    per epoch (plain path timed on the first 32 epochs)."""
    import torch
    import memd_amd
    import synth
    count = 512
    ctx = memd_amd.MemdContext(0)
    eps, Vs = [], {}
    shapes = synth.batch_shapes(count, 77)
    tunings, tune_s = {}, 0.0
    if not a.no_autotune:
        # one tuning per channel-count class, at the class's median length
        t0 = time.perf_counter()
        for n in sorted({s[0] for s in shapes}):
            Tm = int(np.median([T for (nn, T, _) in shapes if nn == n]))
            tunings[n] = ctx.autotune(n, Tm, 2 * n)
        tune_s = time.perf_counter() - t0
    for i, (n, T, K) in enumerate(shapes):
        if (n, K) not in Vs:
            Vs[(n, K)] = ctx.directions(n, K)
        y = torch.from_numpy(synth.eeg_epoch(n, T, 78 + i)).to("cuda:0")
        eps.append({"y": y, "V": Vs[(n, K)], "n_imf": 9, "S": 10,
                    "out": torch.empty((9, n, T), dtype=torch.float64, device="cuda:0")})
    quiesce_gc()
    for _ in range(a.warmup):
        ctx.run_batch(eps)
    torch.cuda.synchronize()
    lat = []
    for _ in range(max(a.steps, 1)):
        s0 = time.perf_counter()
        ctx.run_batch(eps)
        lat.append(time.perf_counter() - s0)
    batched = count * len(lat) / sum(lat)
    # every 16th epoch's batched result, kept to compare with its plain call
    sample = list(range(0, count, 16))
    kept = {i: eps[i]["out"].cpu().numpy().copy() for i in sample}
    plain_n = 0 if a.no_plain else count   # every epoch once, one call each
    plat, plain = [], None
    same = None
    if plain_n:
        ctx.run(eps[0]["y"], eps[0]["V"], 9, 10, out=eps[0]["out"])
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        for ep in eps[:plain_n]:
            s1 = time.perf_counter()
            ctx.run(ep["y"], ep["V"], 9, 10, out=ep["out"])
            plat.append(time.perf_counter() - s1)
        plain = plain_n / (time.perf_counter() - s0)
        same = all(np.array_equal(kept[i], eps[i]["out"].cpu().numpy()) for i in sample)
    # CPU: the oracle's epoch-parallel batch (one epoch per thread) on the first 32 epochs
    cpu = None
    if not a.no_cpu:
        threads = cpu_threads()
        sub = [(eps[i]["y"].cpu().numpy(), eps[i]["V"].cpu().numpy()) for i in range(32)]
        cpu_eps = cpu_epoch_parallel(sub, 9, 10, threads)
        cpu = {"value": round(cpu_eps, 4), "unit": "epochs/s", "cores": threads, "kind": "port",
               "sample": f"the first 32 epochs of the batch, FP64 oracle, {threads} epochs in parallel one per thread",
               "cpu": cpu_info()}
    ctx.close()
    print(json.dumps({
        "metric": "MEMD epochs/s, 512 uneven epochs (batched vs plain GPU)", "value": round(batched, 2),
        "unit": "epochs/s", "n_gpus": 1, "steps": len(lat), "warmup": a.warmup,
        "ms_per_step": round(sum(lat) / len(lat) * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "memd_512_uneven_fix9x10", "n": [8, 16, 32, 64], "T": [1000, 8000], "K": "2n"},
        "plain_gpu_epochs_per_s": round(plain, 2) if plain else None,
        "batched_over_plain": round(batched / plain, 2) if plain else None,
        "batched_equals_plain": {"epochs_checked": len(sample), "bit_identical": same} if plain_n else None,
        "cpu_baseline": cpu, "batched_over_cpu": round(batched / cpu["value"], 1) if cpu else None,
        "tuning": {str(k): v for k, v in tunings.items()}, "autotune_s": round(tune_s, 3),
        "latency_ms": {"batch_p50": round(pct(lat, 50) * 1e3, 1),
                       "plain_epoch_p50": round(pct(plat, 50) * 1e3, 3) if plat else None,
                       "plain_epoch_p99": round(pct(plat, 99) * 1e3, 3) if plat else None}}), flush=True)


def bench_served(a):
    """Served continuous batching: Poisson arrivals at each rate in --rates for
    --duration seconds.  Mix c4: 90% of requests are uneven epochs (n in
    {8,16,32,64}, T in [1000,8000], K = 2n), 10% are 32 ch x 60000, K = 64; mix
    c3: uneven epochs only.  9 slots x 10 sifts, fix mode.  Latency = arrival ->
    result ready on the host.  --saturate N: N epochs of the mix submitted at
    once and drained (served rate at saturation), then the same N epochs as one
    batched call (the batched rate of the same work)."""
    import torch
    import memd_amd
    import synth
    rng = np.random.default_rng(4)
    ctx = memd_amd.MemdContext(0)
    # distinct inputs resident in HBM, reused cyclically (inputs are read-only)
    short = [(int(n), int(T), 2 * int(n)) for n, T, _ in synth.batch_shapes(48, 91)]
    # built-in tile shapes: the autotuner times full-device batches, while a server at
    # moderate load advances a few epochs per step
    pool_in = {"short": [], "long": []}
    Vs = {}
    longs = [(32, 60000, 64)] * 4 if a.mix == "c4" else []
    for i, (n, T, K) in enumerate(short + longs):
        if (n, K) not in Vs:
            Vs[(n, K)] = ctx.directions(n, K)
        y = torch.from_numpy(synth.sines_epoch(n, T, 9000 + i)).to("cuda:0")
        pool_in["long" if T == 60000 else "short"].append((y, Vs[(n, K)]))
    p_long = 0.1 if a.mix == "c4" else 0.0

    def draw():
        kind = "long" if rng.random() < p_long else "short"
        choices = pool_in[kind]
        return kind, choices[int(rng.integers(len(choices)))]

    srv = memd_amd.MemdServer(ctx, pool_bytes=a.pool_gb << 30, max_active=a.max_active)
    results = []
    quiesce_gc()
    for rate in a.rates:
        # arrival schedule
        arr, t = [], 0.0
        while True:
            t += rng.exponential(1.0 / rate)
            if t > a.duration:
                break
            arr.append(t)
        pending = {}
        lat, kinds = [], []
        nxt = 0
        info0 = srv.info()
        t0 = time.perf_counter()
        last_print = t0
        while nxt < len(arr) or pending:
            now = time.perf_counter() - t0
            while nxt < len(arr) and arr[nxt] <= now:
                kind, (y, V) = draw()
                rid = srv.submit(y, V, 9, 10)
                pending[rid] = (arr[nxt], kind)
                nxt += 1
            if not pending:
                time.sleep(min(0.002, max(0.0, arr[nxt] - now)))
                continue
            active, queued = srv.step(1)
            done_t = time.perf_counter() - t0
            for rid in srv.completed():
                ta, kind = pending.pop(rid)
                lat.append(done_t - ta)
                kinds.append(kind)
                srv.release(rid)
            if time.perf_counter() - last_print > 10:
                last_print = time.perf_counter()
                print(f"served: rate {rate}/s t={done_t:.1f}s done={len(lat)} active={active} queued={queued}",
                      file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        info1 = srv.info()
        lat_s = [x for x, k in zip(lat, kinds) if k == "short"]
        lat_l = [x for x, k in zip(lat, kinds) if k == "long"]
        results.append({"rate": rate, "offered": len(arr), "completed": len(lat),
                        "epochs_per_s": round(len(lat) / wall, 2), "wall_s": round(wall, 2),
                        "p50_ms": round(pct(lat, 50) * 1e3, 1), "p99_ms": round(pct(lat, 99) * 1e3, 1),
                        "short_p50_ms": round(pct(lat_s, 50) * 1e3, 1) if lat_s else None,
                        "long_p50_ms": round(pct(lat_l, 50) * 1e3, 1) if lat_l else None,
                        "sift_steps": info1["steps"] - info0["steps"],
                        "plan_rebuilds": info1["rebuilds"] - info0["rebuilds"],
                        "rebuild_host_s": round(info1["rebuild_s"] - info0["rebuild_s"], 3)})
        print(f"served: rate {rate}/s {results[-1]}", file=sys.stderr, flush=True)
    saturated = None
    if a.saturate:
        items = [draw() for _ in range(a.saturate)]
        eps = [{"y": y, "V": V, "n_imf": 9, "S": 10} for _, (y, V) in items]
        work = sum(int(e["y"].shape[0]) * int(e["y"].shape[1]) * int(e["V"].shape[0]) for e in eps)
        ctx.run_batch(eps[:8])   # warm the batched path
        torch.cuda.synchronize()
        info0 = srv.info()
        t0 = time.perf_counter()
        for e in eps:
            srv.submit(e["y"], e["V"], 9, 10)
        while True:
            active, queued = srv.step(1)
            for rid in srv.completed():
                srv.release(rid)
            if active == 0 and queued == 0:
                break
        served_s = time.perf_counter() - t0
        info1 = srv.info()
        t0 = time.perf_counter()
        ctx.run_batch(eps)
        torch.cuda.synchronize()
        batched_s = time.perf_counter() - t0
        saturated = {"epochs": len(eps), "n_T_K_total": work,
                     "served_epochs_per_s": round(len(eps) / served_s, 2),
                     "batched_epochs_per_s": round(len(eps) / batched_s, 2),
                     "served_over_batched": round(batched_s / served_s, 3),
                     "sift_steps": info1["steps"] - info0["steps"],
                     "plan_rebuilds": info1["rebuilds"] - info0["rebuilds"]}
        print(f"served: saturated {saturated}", file=sys.stderr, flush=True)
    srv.close()
    ctx.close()
    best = max(results, key=lambda r: r["epochs_per_s"]) if results else None
    wl = "memd_served_90pct_uneven_10pct_32x60000_fix9x10" if a.mix == "c4" else "memd_served_uneven_fix9x10"
    print(json.dumps({
        "metric": "MEMD served epochs/s with p50/p99 latency (Poisson arrivals)",
        "value": best["epochs_per_s"] if best else (saturated or {}).get("served_epochs_per_s"),
        "unit": "epochs/s", "n_gpus": 1, "steps": len(results), "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl, "duration_s": a.duration, "pool_gb": a.pool_gb, "max_active": a.max_active},
        "latency_ms": {"p50": best["p50_ms"], "p99": best["p99_ms"]} if best else None, "sweep": results,
        "saturated": saturated}), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported: epochs are "
              "independent and shard across GPUs with no exchange (one process per GPU, each "
              "running this benchmark on its share).", file=sys.stderr)
        sys.exit(2)
    import torch
    import memd_amd
    import synth

    if a.copy_bytes:
        # known-size copy: reads and writes copy_bytes each (calibrates FETCH_SIZE / WRITE_SIZE)
        src = torch.ones(a.copy_bytes // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.copy_bytes}), flush=True)
        return

    if a.workload == "eeg32x30504":
        return bench_eeg(a)
    if a.workload == "served":
        return bench_served(a)
    if a.workload == "uneven512":
        return bench_uneven(a)
    if a.workload == "wide128x16384":
        return bench_wide(a)

    n, T, K, N_imf, S, B = a.n, a.T, a.K, a.n_imf, a.sifts, a.batch
    mode = memd_amd.MODE_STOP if a.mode == "stop" else memd_amd.MODE_FIX
    bnd = memd_amd.BOUNDARY_MIRROR if a.boundary == "mirror" else memd_amd.BOUNDARY_CLAMP
    dev = "cuda:0"
    ctx = memd_amd.MemdContext(0)
    tuning, tune_s = None, 0.0
    if a.tuning:
        bg, tw = (int(v) for v in a.tuning.split(","))
        ctx.set_tuning(n, bg, tw)
        tuning = ctx.get_tuning(n)
    elif not a.no_autotune:
        t0 = time.perf_counter()
        tuning = ctx.autotune(n, T, K)   # once per device and shape class; not timed
        tune_s = time.perf_counter() - t0
    V = ctx.directions(n, K)
    ys = [torch.from_numpy(synth.sines_epoch(n, T, 1000 + i)).to(dev) for i in range(B)]
    outs = [torch.empty((N_imf, n, T), dtype=torch.float64, device=dev) for _ in range(B)]
    eps = [{"y": ys[i], "V": V, "n_imf": N_imf, "S": S, "out": outs[i]} for i in range(B)]

    quiesce_gc()
    for _ in range(a.warmup):
        ctx.run_batch(eps, mode=mode, boundary=bnd)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        s0 = time.perf_counter()
        ctx.run_batch(eps, mode=mode, boundary=bnd)
        lat.append(time.perf_counter() - s0)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    value = B * a.steps / total
    print("batched call latencies (ms): " + " ".join(f"{1e3 * v:.1f}" for v in lat), file=sys.stderr, flush=True)

    # plain per-epoch path (one call per epoch), latency per epoch
    plain_lat = []
    if a.plain_epochs > 0:
        ctx.run(ys[0], V, N_imf, S, out=outs[0], mode=mode, boundary=bnd)
        torch.cuda.synchronize()
        for i in range(a.plain_epochs):
            s0 = time.perf_counter()
            ctx.run(ys[i % B], V, N_imf, S, out=outs[i % B], mode=mode, boundary=bnd)
            plain_lat.append(time.perf_counter() - s0)
    plain_eps = len(plain_lat) / sum(plain_lat) if plain_lat else None

    workload = f"memd_{n}ch_x{T}_K{K}_{a.mode}{N_imf}x{S}_b{B}" + ("_mirror" if a.boundary == "mirror" else "")
    fused = a.boundary == "clamp"   # the fused evaluation kernel runs (mirrored ends: the record path)
    roofline, kernels = None, None
    if not a.no_profile:
        # two profiled passes, the second reported: a kernel's first direct launch in a
        # process carries one-off costs that would skew its average
        ctx.run_batch(eps, profile=True, mode=mode, boundary=bnd)
        _, info = ctx.run_batch(eps, profile=True, mode=mode, boundary=bnd)
        pr = info["profile"]
        kernels = {k: round(v * 1e3, 2) for k, v in pr["ms_per_launch"].items()}   # us per launch
        steps = max(pr["steps"], 1)
        W = (T + 63) // 64
        # fused record+evaluation kernel, per launch: cur read+write (16 B/sample) and one read
        # of the sample-major copy that supplies the knot values (8 B/sample); the block
        # solution per knot-channel (8 B); per knot the position, the two spike entries and
        # the two reciprocals (36 B); flag word + prefix per (epoch, direction, envelope,
        # 64 samples) (12 B).  Flops: two 3-fma Horners + mean per evaluation point (15),
        # record build per knot-channel (10), divide + subtract per sample (2).
        if fused and a.mode == "fix":
            bytes_eval = (24 * pr["samples"] + 8 * pr["knot_channels"] + 36 * pr["knots"]) / steps \
                + 12 * 2 * K * W * B
            flops_eval = (15 * pr["eval_points"] + 10 * pr["knot_channels"] + 2 * pr["samples"]) / steps
        elif fused:
            # stop mode: the mean written instead of the signal updated (16 B/sample with the
            # knot-value reads), and one block sum per (valid direction, 16 channels, sample);
            # the difference, its square and the group adds per evaluation point (+3)
            G16 = (n + 15) // 16
            bytes_eval = (16 * pr["samples"] + 8 * pr["knot_channels"] + 36 * pr["knots"]
                          + 8 * pr["eval_points"] * G16 / n) / steps + 12 * 2 * K * W * B
            flops_eval = (18 * pr["eval_points"] + 10 * pr["knot_channels"] + 1 * pr["samples"]) / steps
        else:
            # record-path evaluation: one 32-byte record per knot-channel read, the mean
            # written and |mean|^2, amplitude per sample (8 + 16 B); two Horners, the mean
            # and the squared difference per evaluation point (18)
            bytes_eval = (32 * pr["knot_channels"] + 8 * pr["samples"]) / steps + 16 * T * B
            flops_eval = 18 * pr["eval_points"] / steps
        t_eval = pr["ms_per_launch"]["k_eval"] * 1e-3   # slot of the fused record+evaluation kernel
        gbs = bytes_eval / t_eval / 1e9
        tfs = flops_eval / t_eval / 1e12
        if gbs / HBM_PEAK_GBS >= tfs / FP64_PEAK_TFLOPS:
            roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None}
        else:
            # no matrix instruction runs: the compute bound is the FP64 vector rate
            roofline = {"bound": "fp64", "achieved": round(tfs, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tfs / FP64_PEAK_TFLOPS, 4), "traffic": None}
        roofline["traffic"] = measured_traffic(workload, "k_eval_fused")
        roofline.update({"kernel": ("k_eval_fused" if a.mode == "fix" else "k_eval_fused<stop>") if fused
                         else "k_eval_stop", "launch_us": round(t_eval * 1e6, 2),
                         "alg_bytes_per_launch": int(bytes_eval), "alg_flops_per_launch": int(flops_eval),
                         "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "fp64_frac": round(tfs / FP64_PEAK_TFLOPS, 4)})

    cpu = None if a.no_cpu else cpu_baseline(n, T, K, N_imf, S, mode=mode, boundary=bnd)
    pct_or_none = lambda xs, q: round(pct(xs, q) * 1e3, 3) if xs else None
    ctx.close()
    line = {
        "metric": "MEMD epochs/s (32ch x 4096, K=64); batched vs plain GPU vs CPU",
        "value": round(value, 2), "unit": "epochs/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(total / a.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "n": n, "T": T, "K": K,
                   "N_imf": N_imf, "S": S, "mode": a.mode, "boundary": a.boundary, "window": "refwarp32",
                   "batch_epochs": B,
                   "generator": "sines 5/20/60 Hz, N(0,1)/sqrt(n) mixing, N(0,0.1^2) noise"},
        "latency_ms": {"batched_p50": round(pct(lat, 50) * 1e3, 3), "batched_p99": round(pct(lat, 99) * 1e3, 3),
                       "plain_p50": pct_or_none(plain_lat, 50), "plain_p99": pct_or_none(plain_lat, 99)},
        "plain_gpu_epochs_per_s": round(plain_eps, 2) if plain_eps else None,
        "batched_over_plain": round(value / plain_eps, 2) if plain_eps else None,
        "roofline": roofline,
        "kernel_us_per_launch": kernels,
        "tuning": tuning, "autotune_s": round(tune_s, 3),
        "cpu_baseline": cpu,
        "gpu_over_cpu": round(value / cpu["value"], 1) if cpu else None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
