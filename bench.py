#!/usr/bin/env python3
"""PairHMM benchmark on one MI355X.

Workload (default): 4096 synthetic active regions with uneven shapes --
haplotypes U[2,64], reads log-uniform[10,500], haplotype length U[100,600],
read length one of {100,150,250} per region -- packed once into HBM and run
in one launch sequence per step (FP32 pass over length-binned tiles, then the
FP64 rescue). A step is one pass over all regions; inputs are resident in
HBM before the timed region starts.

Prints one JSON line: GCUPS (cell updates per second / 1e9), the roofline of
the dominant FP32 kernel against MI355X's 157.3 TFLOP/s FP32 peak at 12 flops
per cell, the plain one-region-per-launch GPU form on a sample, accuracy
against the FP64 oracle and the reference routing on a sample, the same
regions served one by one (Poisson arrivals from 32 producer threads, GCUPS
with p50/p99 region latency per arrival rate), the CPU baseline (Rovaca's
region scheme on the host cores) on a bounded sample, and the next kernel on
the same regions: every read realigned to its best haplotype by the device
aligner, as one batch and as one served call per region (`realign`).

Other modes: --workload stress (precision stress, 4096 regions), single (one
8 x 32 region, latency), served (the serving loop alone), deep (served with
1 % deep regions of 2000 reads x 128 haplotypes of 1000 bp), calls (the
region-level call_pairhmm from worker threads, reads preprocessed on the
device or on the calling threads, PCR model off and CONSERVATIVE), sw (the
realignment aligner on one batch of 1024 regions' reads, --assemble for the
assembly scores) and sw_calls (the aligner called once per region from
worker threads, served and batched).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_CELL = 12          # M: 4 mul + 2 add, I: 2 mul + 1 add, D: 2 mul + 1 add
FP32_PEAK_TFLOPS = 157.3     # MI355X published FP32 vector peak
FP64_PEAK_TFLOPS = 78.6      # MI355X published FP64 vector peak
METRIC = "PairHMM GCUPS + p99 region latency, 1 MI355X vs host AVX-512 & unbatched GPU"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(name, n_regions):
    from rovaca_amd import workload
    if name == "batched":
        return workload.batched_regions(n_regions, seed=1), f"batched_uneven_{n_regions}_regions"
    if name == "stress":
        return workload.stress_regions(n_regions, seed=2), f"precision_stress_{n_regions}_regions"
    if name == "single":
        return [workload.single_region(0)], "single_region_8x32"
    raise SystemExit(f"unknown workload {name}")


def accuracy(regions, batch, out, sample):
    """max/mean |delta log10 L| vs the FP64 oracle and the reference routing, on a sample."""
    import oracle
    d64, dref = [], []
    for x in sample:
        reg = regions[x]
        got = batch.region_result(out, x)
        want, route, ok = oracle.region_reference(reg)
        m = ~np.isnan(want)
        dref.append(np.abs(got[m] - want[m]))
        f64 = np.full_like(want, np.nan)
        H = reg["n_haps"]
        for r in range(reg["n_reads"]):
            o0, o1 = reg["read_off"][r], reg["read_off"][r + 1]
            if o1 == o0:
                continue
            args = [reg[k][o0:o1] for k in "qidc"]
            for h in range(H):
                hb = reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]]
                if len(hb):
                    f64[r, h] = oracle.pair_log10_f64(hb, reg["read_bases"][o0:o1], *args)
        m = ~np.isnan(f64)
        d64.append(np.abs(got[m] - f64[m]))
    d64 = np.concatenate(d64)
    dref = np.concatenate(dref)
    return {"sample_regions": len(sample), "sample_pairs": int(len(d64)),
            "max_abs_vs_fp64_oracle": float(d64.max()), "mean_abs_vs_fp64_oracle": float(d64.mean()),
            "max_abs_vs_reference_routing": float(dref.max()),
            "mean_abs_vs_reference_routing": float(dref.mean())}


def pair_accuracy(reg, got, pairs=None, routing=True):
    """|delta log10 L| of one region's result matrix (reads x haplotypes)
    against the FP64 oracle and Rovaca's own intra kernel (built from its
    sources into oracle/_ref; FP32 first, FP64 below 1e-28) on every pair or
    the (read, haplotype) pairs given, and against the reference routing on
    the whole region."""
    import oracle
    d64, dri = [], []
    if pairs is None:
        pairs = [(r, h) for r in range(reg["n_reads"]) for h in range(reg["n_haps"])]
    have_ref = oracle.ref() is not None
    for r, h in pairs:
        o0, o1 = reg["read_off"][r], reg["read_off"][r + 1]
        hb = reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]]
        if o1 == o0 or len(hb) == 0:
            continue
        rd = reg["read_bases"][o0:o1]
        args = [reg[k][o0:o1] for k in "qidc"]
        d64.append(abs(got[r, h] - oracle.pair_log10_f64(hb, rd, *args)))
        if have_ref:
            dri.append(abs(got[r, h] - oracle.ref_intra_log10(hb, rd, *args)))
    out = {"sample_pairs": len(d64), "max_abs_vs_fp64_oracle": float(np.max(d64)),
           "mean_abs_vs_fp64_oracle": float(np.mean(d64))}
    if dri:
        out["max_abs_vs_reference_intra"] = float(np.max(dri))
        out["mean_abs_vs_reference_intra"] = float(np.mean(dri))
    if routing:
        want, route, ok = oracle.region_reference(reg)
        m = ~np.isnan(want)
        d = np.abs(got[m] - want[m])
        out["max_abs_vs_reference_routing"] = float(d.max())
        out["mean_abs_vs_reference_routing"] = float(d.mean())
    return out


def cpu_baseline(regions, target_seconds):
    """Rovaca's region scheme (three passes, 16-lane inter kernel, Rovaca's own
    intra kernel for pass 3) on the host cores, over a bounded sample: the
    median of 3 timed runs after one warm-up run."""
    import oracle
    from rovaca_amd import workload
    L = oracle.baseline()
    threads = max(1, min(16, os.cpu_count() or 1))
    # calibrate on one core with a few regions, then size the sample so that
    # the three timed runs take about target_seconds together
    cal = regions[:4]
    oracle.baseline_run(cal, 1)
    secs1, _ = oracle.baseline_run(cal, 1)
    rate1 = workload.cells(cal) / max(secs1, 1e-6)
    want_cells = rate1 * threads * target_seconds / 3
    sample, tot = [], 0
    for r in regions[4:]:
        sample.append(r)
        tot += workload.cells([r])
        if tot >= want_cells:
            break
    oracle.baseline_run(sample[:max(1, len(sample) // 8)], threads)
    runs = sorted(oracle.baseline_run(sample, threads)[0] for _ in range(3))
    secs = runs[1]
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": tot / secs / 1e9, "unit": "GCUPS", "cores": threads,
            "kind": "port",
            "sample": f"{len(sample)} regions of the same workload ({tot / 1e9:.1f} Gcells), median of 3 runs "
                      f"({', '.join(f'{x:.2f}' for x in runs)} s wall)",
            "isa": "avx512" if L.orc_baseline_isa() == 512 else "avx2",
            "pass3_intra": "rovaca_compiled" if L.uses_reference_intra else "restatement",
            "single_core_gcups": rate1 / 1e9, "cpu_model": model}


def plain_gpu(regions, n):
    """One region per launch (rvc_phmm_region_sync), sequentially, on a sample."""
    import rovaca_amd as rv
    from rovaca_amd import workload
    sample = regions[:n]
    rv.region_sync(sample[0])
    t0 = time.perf_counter()
    for r in sample:
        rv.region_sync(r)
    secs = time.perf_counter() - t0
    return {"value": workload.cells(sample) / secs / 1e9, "unit": "GCUPS", "regions": len(sample),
            "mean_region_ms": secs / len(sample) * 1e3}


def run_single(args):
    """One 8 x 32 region through the plain form, repeated: latency and GCUPS."""
    rec = single_region_config(args.no_cpu, reps=max(200, args.steps))
    return {"metric": "PairHMM single-region latency (8 haplotypes x 32 reads, 1 MI355X)",
            "value": rec["median_ms"], "unit": "ms", "higher_is_better": False, "n_gpus": 1,
            "steps": rec["calls"], "warmup": 20, "dtype": "f32", "rescue_dtype": "f64", "data": "synthetic",
            "config": {"workload": "single_region_8x32", "cells": rec["cells"]}, **rec}


def served_stream(pool, deep, rates, duration, producers, warmup_s=0.3, check=None):
    """Poisson arrivals of pool regions (and 1 % deep ones if given) from
    'producers' host threads at each rate, through the serving loop. Before
    each timed window a short untimed stream at the same rate (and with the
    same mix) sizes the slot arenas, staging and result buffers, so that the
    timed window reallocates nothing (buffer_grows reports it). check = n:
    after the timed windows, n arrivals at 8k/s with their outputs kept, each
    compared bit for bit with the region computed alone (untimed); the
    count goes into the returned stats as 'served_check'."""
    import rovaca_amd as rv
    regions = pool + deep
    rv.init(0, server=True, hbm_budget_bytes=64 << 30, max_batch_regions=4096)
    rng = np.random.default_rng(9)

    def arrivals(rate, n):
        arr = np.cumsum(rng.exponential(1.0 / rate, n))
        order = rng.integers(0, len(pool), n)
        if deep:
            is_deep = rng.random(n) < 0.01
            order[is_deep] = len(pool) + rng.integers(0, len(deep), int(is_deep.sum()))
        return arr, order

    out = []
    for rate in rates:
        wa, wo = arrivals(rate, max(64, int(rate * warmup_s)))
        if deep:
            wo[:len(deep)] = len(pool) + np.arange(len(deep))  # every deep region once in the warm-up
        rv.loadgen_run(regions, wa, wo, producers)
        n = max(64, int(rate * duration))
        arr, order = arrivals(rate, n)
        st0 = rv.server_stats()
        res, lat = rv.loadgen_run(regions, arr, order, producers)
        st1 = rv.server_stats()
        launches = st1["launches"] - st0["launches"]
        rec = {"rate_per_s": rate, "regions": int(res["completed"]),
               "gcups": round(res["cells"] / res["wall_s"] / 1e9, 1),
               "offered_gcups": round(res["cells"] / arr[-1] / 1e9, 1),
               "p50_ms": round(res["lat_p50_ms"], 3), "p99_ms": round(res["lat_p99_ms"], 3),
               "max_ms": round(res["lat_max_ms"], 3), "queue_p50_ms": round(res["queue_p50_ms"], 3),
               "queue_p99_ms": round(res["queue_p99_ms"], 3), "submit_late_p99_ms": round(res["late_p99_ms"], 3),
               "wall_s": round(res["wall_s"], 3), "launches": int(launches),
               "mean_regions_per_launch": round(int(res["completed"]) / max(1, launches), 1),
               "host_stage_ms": round(st1["stage_ms"] - st0["stage_ms"], 1),
               "buffer_grows": int(st1["buffer_grows"] - st0["buffer_grows"])}
        if deep:
            dl = lat[order >= len(pool)]
            rec["deep_regions"] = int((order >= len(pool)).sum())
            rec["deep_p50_ms"] = round(float(np.median(dl)), 3) if len(dl) else None
            rec["deep_p99_ms"] = round(float(np.percentile(dl, 99)), 3) if len(dl) else None
            rec["cells"] = int(res["cells"])
        out.append(rec)
        log(f"rate {rate:.0f}/s: {rec['gcups']} GCUPS, p50 {rec['p50_ms']} ms, p99 {rec['p99_ms']} ms, "
            f"{rec['launches']} launches, {rec['buffer_grows']} buffer grows")
    st = rv.server_stats()
    if check:
        arr, order = arrivals(8000.0, check)
        order = order % len(pool)
        _, _, outs = rv.loadgen_run(regions, arr, order, producers, outputs=True)
        alone = {int(x): rv.region_sync(pool[int(x)])[0] for x in set(order.tolist())}
        same = sum(1 for i in range(len(order))
                   if np.array_equal(np.isnan(alone[int(order[i])]), np.isnan(outs[i]))
                   and np.array_equal(alone[int(order[i])][~np.isnan(outs[i])], outs[i][~np.isnan(outs[i])]))
        st["served_check"] = {"arrivals": int(len(order)), "rate_per_s": 8000.0,
                              "bit_identical_to_region_alone": int(same)}
    rv.shutdown()
    return out, st


def highest_rate_under(rates, p99_ms=10.0):
    ok = [r["rate_per_s"] for r in rates if r["p99_ms"] < p99_ms]
    return max(ok) if ok else None


def single_region_config(no_cpu, reps=400):
    """The plain single active region (8 haplotypes x 32 reads): per-call
    latency of the one-region path as a C++ caller sees it (the calls made
    from native code, rvc_loadgen_region_latency) and through Python, the
    entry point that allocates nothing, the region's launch sequence with
    resident inputs eager and as a hipGraph replay, and Rovaca's region
    scheme on one host thread for the same region."""
    import rovaca_amd as rv
    from rovaca_amd import workload
    reg = workload.single_region(0)
    cells = workload.cells([reg])
    rv.init(0)
    rv.region_latency(reg, 50)
    native = rv.region_latency(reg, reps)
    native_na = rv.region_latency(reg, reps, noalloc=True)
    for _ in range(20):
        rv.region_sync(reg)
    got, _ = rv.region_sync(reg)
    acc = pair_accuracy(reg, got)
    lat = []
    for _ in range(reps):
        t0 = time.perf_counter()
        rv.region_sync(reg)
        lat.append((time.perf_counter() - t0) * 1e3)
    # the region's launch sequence with its inputs resident, queued eagerly
    # and replayed as a hipGraph (rvc_phmm_batch with use_graph)
    seq = {}
    for mode in ("eager", "graph"):
        b = rv.Batch([reg], use_graph=(mode == "graph"), untimed=True)
        for _ in range(20):
            b.run()
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            b.run()
            t.append((time.perf_counter() - t0) * 1e3)
        seq[mode] = round(float(np.median(t)), 4)
        b.close()
    rv.shutdown()
    lat = np.array(lat)
    out = {"workload": "single_region_8x32", "cells": cells, "calls": reps,
           "median_ms": round(float(np.median(native)), 4), "p99_ms": round(float(np.percentile(native, 99)), 4),
           "gcups_at_median": round(cells / (np.median(native) * 1e-3) / 1e9, 2),
           "caller": "C++ (rvc_loadgen_region_latency)",
           "noalloc_median_ms": round(float(np.median(native_na)), 4),
           "python_median_ms": round(float(np.median(lat)), 4), "python_p99_ms": round(float(np.percentile(lat, 99)), 4),
           "resident_launch_sequence_ms": {"eager": seq["eager"], "hipgraph_replay": seq["graph"]},
           "accuracy": acc}
    if not no_cpu:
        import oracle
        oracle.baseline_run([reg], 1)
        secs = min(oracle.baseline_run([reg], 1)[0] for _ in range(5))
        out["cpu_baseline"] = {"value": round(secs * 1e3, 4), "unit": "ms", "cores": 1, "kind": "port",
                               "sample": "the same region, one worker thread, best of 5"}
        out["cpu_over_gpu_latency"] = round(secs * 1e3 / out["median_ms"], 1)
    return out


def fixture_accuracy():
    """The device against Rovaca's own intra kernel on the committed fixtures
    (inputs and reference outputs stored together): max/mean |delta log10 L|."""
    import rovaca_amd as rv
    d, n = [], 0
    gold = os.path.join(ROOT, "tests", "golden")
    for name in sorted(os.listdir(gold)):
        if not (name.endswith(".npz") and name.startswith(("single_", "batched_", "stress_"))):
            continue
        g = np.load(os.path.join(gold, name))
        reg = {k: g[k] for k in ("hap_off", "hap_bases", "read_off", "read_bases", "q", "i", "d", "c")}
        reg["n_haps"], reg["n_reads"] = len(reg["hap_off"]) - 1, len(reg["read_off"]) - 1
        got, _ = rv.region_sync(reg)
        m = ~np.isnan(g["ref_intra"])
        d.append(np.abs(got[m] - g["ref_intra"][m]))
        n += 1
    d = np.concatenate(d)
    return {"fixtures": n, "pairs": int(len(d)), "max_abs_vs_reference_intra": float(d.max()),
            "mean_abs_vs_reference_intra": float(d.mean())}


def stress_config(args):
    """Precision stress (2048 regions, 30 % of reads 250 bp with 40 %
    substitutions and Q 2-10): batched GCUPS, FP32/FP64 split, rescued
    fraction, accuracy against the FP64 oracle and Rovaca's intra kernel
    (committed fixtures), CPU baseline, and the plain and served forms."""
    import rovaca_amd as rv
    from rovaca_amd import workload
    regions = workload.stress_regions(2048, seed=2)
    cells = workload.cells(regions)
    rv.init(0)
    b = rv.Batch(regions, use_graph=True)
    b.run()
    t0 = time.perf_counter()
    steps = 4
    for _ in range(steps):
        b.run()
    wall = time.perf_counter() - t0
    b.close()
    e = rv.Batch(regions)
    e.run()
    split = []
    for _ in range(2):
        e.run()
        split.append(e.info())
    out_v, flags = e.results()
    rescued_cells = 0
    for x, reg in enumerate(regions):
        fl = e.region_result(flags, x)
        if (fl & 2).any():
            hl, rl = np.diff(reg["hap_off"]), np.diff(reg["read_off"])
            rr, hh = np.nonzero(fl & 2)
            rescued_cells += int((rl[rr].astype(np.int64) * hl[hh]).sum())
    fp32_ms = float(np.mean([x["fp32_ms"] for x in split]))
    fp64_ms = float(np.mean([x["fp64_ms"] for x in split]))
    n_pairs = int(split[-1]["n_pairs"])
    n_resc = int(split[-1]["n_rescued"])
    value = cells * steps / wall / 1e9
    rec = {"workload": "precision_stress_2048_regions", "regions": len(regions), "pairs": n_pairs, "cells": cells,
           "gcups": round(value, 1), "ms_per_step": round(wall / steps * 1e3, 3),
           "device_ms": {"fp32_pass": round(fp32_ms, 3), "fp64_rescue": round(fp64_ms, 3)},
           "rescue": {"pairs": n_resc, "fraction": round(n_resc / max(1, n_pairs), 4), "cells": rescued_cells,
                      "fp64_tflops": round(rescued_cells * FLOPS_PER_CELL / (fp64_ms * 1e-3) / 1e12, 2),
                      "fp64_frac": round(rescued_cells * FLOPS_PER_CELL / (fp64_ms * 1e-3) / 1e12
                                         / FP64_PEAK_TFLOPS, 4)}}
    log(f"stress: {rec['gcups']} GCUPS, FP32 {fp32_ms:.1f} ms + FP64 {fp64_ms:.1f} ms, {n_resc} rescued")
    small = [x for x, r in enumerate(regions) if r["n_haps"] * r["n_reads"] <= 1000][:8]
    if small:
        rec["accuracy"] = accuracy(regions, e, out_v, small)
    rec["accuracy_vs_reference_fixtures"] = fixture_accuracy()
    e.close()
    rec["plain_gpu"] = plain_gpu(regions, 128)
    rec["batched_over_plain"] = round(value / rec["plain_gpu"]["value"], 2)
    rv.shutdown()
    srv, _ = served_stream(regions, [], [16000.0], 0.5, args.producers)
    rec["served"] = srv[0]
    rec["served_over_plain"] = round(srv[0]["gcups"] / rec["plain_gpu"]["value"], 2)
    if not args.no_cpu:
        try:
            rec["cpu_baseline"] = cpu_baseline(regions, 10.0)
            rec["gpu_over_cpu"] = round(value / rec["cpu_baseline"]["value"], 1)
            rec["served_over_cpu"] = round(srv[0]["gcups"] / rec["cpu_baseline"]["value"], 1)
        except Exception as ex:  # noqa: BLE001
            rec["cpu_baseline"] = {"error": str(ex)}
    return rec


def deep_config(pool, args):
    """Deep-coverage tail under serving: 1 % of the served stream replaced by
    regions of 2000 reads (250 bp) x 128 haplotypes (1000 bp), at each rate,
    with the plain per-region form and Rovaca's region scheme on the host
    cores on a sample of the same mix (99 pool regions + 1 deep region)."""
    from rovaca_amd import workload
    deep = [workload.deep_region(seed=40 + x) for x in range(4)]
    rates = [float(x) for x in args.rates.split(",")]
    srv, st = served_stream(pool, deep, rates, args.duration, args.producers)
    rec = {"workload": "served_with_1pct_deep_regions", "deep_region": "2000 reads x 250 bp, 128 haplotypes x 1000 bp",
           "deep_cells": workload.cells(deep[:1]), "pool_regions": len(pool), "producers": args.producers,
           "rates": srv, "highest_rate_p99_under_10ms": highest_rate_under(srv),
           "peak_resident_gib": round(st["peak_resident_bytes"] / 2**30, 3)}
    import rovaca_amd as rv
    mix = pool[:99] + deep[:1]
    rv.init(0)
    # accuracy on 256 seeded pairs of a deep region (the whole region is
    # 6.4e10 cells, beyond the oracle's reach in a bench run)
    got, _ = rv.region_sync(deep[0])
    rng = np.random.default_rng(77)
    sample = list(zip(rng.integers(0, deep[0]["n_reads"], 256).tolist(), rng.integers(0, deep[0]["n_haps"], 256).tolist()))
    rec["accuracy"] = pair_accuracy(deep[0], got, sample, routing=False)
    rv.region_sync(mix[0])
    t0 = time.perf_counter()
    for r in mix:
        rv.region_sync(r)
    secs = time.perf_counter() - t0
    rv.shutdown()
    mix_cells = workload.cells(mix)
    rec["plain_gpu"] = {"value": round(mix_cells / secs / 1e9, 1), "unit": "GCUPS",
                        "sample": "99 pool regions + 1 deep region, one region per launch"}
    top = srv[-1]["gcups"]
    rec["served_over_plain"] = round(top / rec["plain_gpu"]["value"], 2)
    if not args.no_cpu:
        import oracle
        threads = max(1, min(16, os.cpu_count() or 1))
        oracle.baseline_run(mix[:8], threads)
        runs = sorted(oracle.baseline_run(mix, threads)[0] for _ in range(3))
        rec["cpu_baseline"] = {"value": round(mix_cells / runs[1] / 1e9, 2), "unit": "GCUPS", "cores": threads,
                               "kind": "port",
                               "sample": f"99 pool regions + 1 deep region ({mix_cells / 1e9:.1f} Gcells), "
                                         f"median of 3 runs ({', '.join(f'{x:.2f}' for x in runs)} s wall)"}
        rec["served_over_cpu"] = round(top / rec["cpu_baseline"]["value"], 1)
    return rec


def run_served(args):
    """Served stream: Poisson arrivals from 32 producer threads at each rate;
    'deep' mixes 1 % regions of 2000 reads x 128 haplotypes (1000 bp)."""
    from rovaca_amd import workload
    t = time.time()
    pool = workload.batched_regions(min(args.regions, 2048), seed=3)
    log(f"pool of {len(pool)} regions generated in {time.time() - t:.1f} s")
    if args.workload == "deep":
        rec = deep_config(pool, args)
        top = rec["rates"][-1]
        return {"metric": "PairHMM served_with_1pct_deep_regions GCUPS + p99 region latency (1 MI355X)",
                "value": top["gcups"], "unit": "GCUPS", "higher_is_better": True, "n_gpus": 1, "dtype": "f32",
                "rescue_dtype": "f64", "data": "synthetic", "config": {"workload": rec["workload"]}, **rec}
    rates = [float(x) for x in args.rates.split(",")]
    out, st = served_stream(pool, [], rates, args.duration, args.producers)
    top = out[-1]
    return {"metric": "PairHMM served GCUPS + p99 region latency (1 MI355X)", "value": top["gcups"],
            "unit": "GCUPS", "higher_is_better": True, "n_gpus": 1, "dtype": "f32", "rescue_dtype": "f64",
            "data": "synthetic", "config": {"workload": "served", "pool_regions": len(pool),
                                            "producers": args.producers, "hbm_budget_gib": 64,
                                            "arrivals": "poisson", "duration_s_per_rate": args.duration},
            "rates": out, "highest_rate_p99_under_10ms": highest_rate_under(out),
            "server": {"launches": st["launches"], "peak_resident_gib": round(st["peak_resident_bytes"] / 2**30, 3),
                       "rescued_pairs": st["rescued_pairs"]}}


def run_calls(args):
    """Rovaca's worker pattern: args.producers threads call call_pairhmm (raw
    reads, region-level) one region at a time through the serving loop, with
    the reads preprocessed on the device or on the calling threads."""
    import rovaca_amd as rv
    from rovaca_amd import workload
    pool = workload.batched_regions(min(args.regions, 2048), seed=3)
    rng = np.random.default_rng(17)
    raws = [rv.raw_region(r, rng.integers(20, 61, int(r["n_reads"]))) for r in pool]
    n_calls = max(256, int(args.calls))
    order = rng.integers(0, len(pool), n_calls)
    out = []
    for pcr_name, pcr in (("NONE", rv.PCR_NONE), ("CONSERVATIVE", rv.PCR_CONSERVATIVE)):
        for where in ("device", "host"):
            rv.init(0, server=True, hbm_budget_bytes=64 << 30, host_prep=1 if where == "host" else 2)
            rv.loadgen_call_pairhmm(raws, 128, order[:128], 10, pcr, args.producers)  # warm-up
            res, lat = rv.loadgen_call_pairhmm(raws, n_calls, order, 10, pcr, args.producers)
            rv.shutdown()
            rec = {"pcr": pcr_name, "preprocessing": where, "calls": int(res["completed"]),
                   "gcups": round(res["cells"] / res["wall_s"] / 1e9, 1), "wall_s": round(res["wall_s"], 3),
                   "p50_ms": round(res["lat_p50_ms"], 3), "p99_ms": round(res["lat_p99_ms"], 3)}
            out.append(rec)
            log(f"pcr {pcr_name}, preprocessing on {where}: {rec['gcups']} GCUPS, p50 {rec['p50_ms']} ms, "
                f"p99 {rec['p99_ms']} ms")
    top = max(r["gcups"] for r in out if r["preprocessing"] == "device")
    return {"metric": "PairHMM region calls (call_pairhmm from worker threads, served, 1 MI355X)", "value": top,
            "unit": "GCUPS", "higher_is_better": True, "n_gpus": 1, "dtype": "f32", "rescue_dtype": "f64",
            "data": "synthetic", "config": {"workload": "region_calls", "pool_regions": len(pool),
                                            "workers": args.producers, "calls": n_calls, "min_quality": 10,
                                            "mq": "U[20,60]"}, "runs": out}


def realign_inputs(args):
    """The first regions of the batched workload and, per region, its
    haplotypes, its reads and each read's best haplotype (the PairHMM matrix's
    largest entry in the read's row; reads without one are left out)."""
    import rovaca_amd as rv
    from rovaca_amd import workload
    regions = workload.batched_regions(min(args.regions, 1024), seed=1)
    rv.init(0)
    b = rv.Batch(regions)
    b.run()
    lk_all, _ = b.results()
    per_region = []
    for x, reg in enumerate(regions):
        lk = b.region_result(lk_all, x)  # [n_reads][n_haps]
        haps = [reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]].tobytes() for h in range(reg["n_haps"])]
        reads, best = [], []
        for r in range(reg["n_reads"]):
            row = lk[r]
            if np.all(np.isnan(row)):
                continue
            reads.append(reg["read_bases"][reg["read_off"][r]:reg["read_off"][r + 1]].tobytes())
            best.append(int(np.nanargmax(row)))
        per_region.append((haps, reads, best))
    b.close()
    return regions, per_region


def run_sw_calls(args):
    """Realignment the way Rovaca's workers would call it: one call per region
    (its reads against their best haplotypes) from 'producers' native worker
    threads (rvc_loadgen_sw_align), each region's batch packed beforehand,
    through the served call (concurrent calls merged into shared device
    passes) and, for comparison, the batched call made per region. GCUPS over
    the whole run and per-call latency; the best of 'steps' runs after an
    untimed one."""
    import rovaca_amd as rv
    regions, per_region = realign_inputs(args)
    batches = [rv.SwBatch(h, r, b) for h, r, b in per_region if r]
    n_thr = max(1, args.producers)
    log(f"sw_calls: {len(batches)} region calls from {n_thr} threads")

    def stream(served):
        runs = [rv.loadgen_sw_align(batches, served, n_thr) for _ in range(max(1, args.steps) + 1)][1:]
        res, lat = min(runs, key=lambda r: r[0]["wall_s"])
        return {"gcups": round(res["cells"] / res["wall_s"] / 1e9, 1), "wall_ms": round(res["wall_s"] * 1e3, 2),
                "p50_ms": round(res["lat_p50_ms"], 3), "p99_ms": round(res["lat_p99_ms"], 3)}, res["cells"]

    served, cells = stream(True)
    direct, _ = stream(False)
    rv.shutdown()
    return {"metric": f"Smith-Waterman realignment, one served call per region from {n_thr} threads (1 MI355X)",
            "value": served["gcups"], "unit": "GCUPS", "higher_is_better": True, "n_gpus": 1,
            "dtype": "int16", "data": "synthetic",
            "config": {"workload": f"sw_realign_calls_{len(regions)}_regions", "calls": len(batches), "cells": cells,
                       "producers": n_thr},
            "served": served, "batched_call_per_region": direct,
            "mean_reads_per_call": round(sum(len(r) for _, r, _ in per_region) / max(1, len(batches)), 1)}


def run_sw(args):
    """Realignment of reads to their best haplotype (the next kernel after
    PairHMM on the same region): every read of the first regions of the
    batched workload against the haplotype the PairHMM matrix ranks first,
    through the batched device aligner. The CPU baseline is Rovaca's own
    aligner built from its sources (oracle/_ref) on the host's cores."""
    import rovaca_amd as rv
    regions, per_region = realign_inputs(args)
    haps, reads, read_hap = [], [], []
    for rh, rr, rb in per_region:
        h0 = len(haps)
        haps += rh
        reads += rr
        read_hap += [h0 + x for x in rb]
    cells = sum(len(haps[h]) * len(r) for h, r in zip(read_hap, reads))
    log(f"sw: {len(reads)} reads of {len(regions)} regions against their best haplotype, {cells / 1e9:.1f} Gcells")
    # the batch packed once and run repeatedly with the same result buffers
    # (the first run also touches the fresh output pages); timed runs after it
    assemble = bool(getattr(args, "assemble", False))
    sb = rv.SwBatch(haps, reads, read_hap, assemble)
    sb.run()
    runs = [sb.run() for _ in range(max(1, args.steps))]
    res = sb.results()
    dev_ms = min(s["fill_ms"] + s["trace_ms"] for s in runs)
    call_ms = min(s["total_ms"] for s in runs)
    best = min(runs, key=lambda s: s["fill_ms"] + s["trace_ms"])
    rv.shutdown()
    # parity sample against the oracle
    import oracle
    sample = range(0, len(reads), max(1, len(reads) // 3000))
    mism = sum(1 for i in sample if res[i] != oracle.sw_align(haps[read_hap[i]], reads[i], assemble))
    out = {"metric": "Smith-Waterman realignment GCUPS (reads vs best haplotype, 1 MI355X)",
           "value": round(cells / (dev_ms * 1e-3) / 1e9, 1), "unit": "GCUPS", "higher_is_better": True, "n_gpus": 1,
           # genotype scores on reads of 16+ bases: the packed int16 sweep (two
           # pairs per lane); shorter reads, and assembly scores on pairs of
           # 61+ bases, take the int32 sweep
           "dtype": "int32" if assemble else "int16", "data": "synthetic",
           "config": {"workload": f"sw_realign_{len(regions)}_regions", "pairs": len(reads), "cells": cells,
                      "scores": "assembly 200/-150/-260/-11" if assemble else "genotype 10/-15/-30/-5"},
           "device_ms": {"sweep": round(best["fill_ms"], 3), "traceback": round(best["trace_ms"], 3)},
           "call_ms": round(call_ms, 3), "gcups_with_copies": round(cells / (call_ms * 1e-3) / 1e9, 1),
           "launches": best["launches"], "chunks": best["chunks"],
           "parity": {"sample_pairs": len(sample), "mismatches_vs_oracle": mism}}
    if not args.no_cpu and oracle.ref_sw() is not None:
        cores = min(16, os.cpu_count() or 1)
        n_cpu = len(reads)
        cpu_cells = cells
        # all pairs, repeated for about 5 s; the median pass
        passes = []
        while len(passes) < 3 or (sum(passes) < 5.0 and len(passes) < 40):
            passes.append(oracle.ref_sw_align_many(haps, reads, read_hap, cores, assemble))
        secs = float(np.median(passes))
        out["cpu_baseline"] = {"value": round(cpu_cells / secs / 1e9, 2), "unit": "GCUPS", "cores": cores,
                               "kind": "reference",
                               "sample": f"all {n_cpu} pairs ({cpu_cells / 1e9:.1f} Gcells), Rovaca's aligner "
                                         f"(oracle/_ref/libref_sw.so, its AVX-512 cores: 16-bit, 32-bit where the scores could "
                                         f"overflow 16 bits), {cores} threads, "
                                         f"median of {len(passes)} passes, {secs:.3f} s each"}
        out["gpu_over_cpu"] = round(out["value"] / out["cpu_baseline"]["value"], 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "stress", "single", "served", "deep", "calls", "sw", "sw_calls"])
    ap.add_argument("--calls", type=int, default=8192, help="calls: region calls per run")
    ap.add_argument("--rates", default="2000,8000,16000,24000,32000", help="served: arrival rates, regions/s")
    ap.add_argument("--duration", type=float, default=1.0, help="served: seconds of arrivals per rate")
    ap.add_argument("--producers", type=int, default=32)
    ap.add_argument("--assemble", action="store_true", help="sw workload: the assembly scores (the int32 sweep)")
    ap.add_argument("--regions", type=int, default=4096)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip plain-GPU and accuracy samples")
    ap.add_argument("--no-served", action="store_true", help="skip the served stream of the batched workload")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other configs in the default line (single region, precision stress, deep tail)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (known bytes) to scale HBM byte counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 28  # 2^28 float32 = 1 GiB read + 1 GiB written per copy
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (regions are independent: "
              "more GPUs would each take a share of them from a queue of their own)", file=sys.stderr)
        sys.exit(2)

    import rovaca_amd as rv
    from rovaca_amd import workload

    if args.workload in ("served", "deep"):
        print(json.dumps(run_served(args)), flush=True)
        return
    if args.workload == "single":
        print(json.dumps(run_single(args)), flush=True)
        return
    if args.workload == "calls":
        print(json.dumps(run_calls(args)), flush=True)
        return
    if args.workload == "sw":
        print(json.dumps(run_sw(args)), flush=True)
        return
    if args.workload == "sw_calls":
        print(json.dumps(run_sw_calls(args)), flush=True)
        return

    t = time.time()
    regions, wname = make_workload(args.workload, args.regions)
    cells = workload.cells(regions)
    log(f"workload {wname}: {len(regions)} regions, {cells / 1e9:.1f} Gcells, generated in {time.time() - t:.1f} s")
    rv.init(0)
    t = time.time()
    batch = rv.Batch(regions, use_graph=bool(args.graph))
    info0 = batch.info()
    log(f"packed + uploaded in {time.time() - t:.1f} s, {info0['device_bytes'] / 2**30:.2f} GiB on device, "
        f"{info0['n_pairs']} pairs")
    for w in range(args.warmup):
        batch.run()
        log(f"warmup {w} done")
    # timed region: each run() returns after the device finished (stream synchronised)
    dev_total, dev_fp32 = [], []
    t0 = time.perf_counter()
    for s in range(args.steps):
        batch.run()
        inf = batch.info()
        dev_total.append(inf["total_ms"])
    t1 = time.perf_counter()
    wall = t1 - t0
    ms_per_step = wall / args.steps * 1e3
    value = cells * args.steps / wall / 1e9
    log(f"timed {args.steps} steps: {ms_per_step:.1f} ms/step, {value:.1f} GCUPS")
    # FP32 / FP64 split from an eager (non-graph) run with HIP events on the batch stream
    ebatch = batch
    if args.graph:
        batch.close()
        ebatch = rv.Batch(regions, use_graph=False)
    ebatch.run()
    split = []
    for s in range(max(1, min(3, args.steps))):
        ebatch.run()
        split.append(ebatch.info())
    fp32_ms = float(np.mean([x["fp32_ms"] for x in split]))
    fp64_ms = float(np.mean([x["fp64_ms"] for x in split]))
    n_rescued = int(split[-1]["n_rescued"])
    kernels = ebatch.kernels()  # per-launch HIP events of the last eager run, on the batch's stream
    out, flags = ebatch.results()
    rescued_cells = 0
    # cells of the pairs that went to FP64
    for x, reg in enumerate(regions):
        fl = ebatch.region_result(flags, x)
        if (fl & 2).any():
            hl = np.diff(reg["hap_off"])
            rl = np.diff(reg["read_off"])
            rr, hh = np.nonzero(fl & 2)
            rescued_cells += int((rl[rr].astype(np.int64) * hl[hh]).sum())
    # roofline of the dominant kernel: the FP32 launch with the most device time
    dom = max((k for k in kernels if k["precision"] == 32), key=lambda k: k["ms"])
    dom_tflops = dom["cells"] * FLOPS_PER_CELL / (dom["ms"] * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"k_fp32_{wname.replace('_regions', '')}_hbm_traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))["kernels"].get(dom["name"])
        if t:
            traffic = round(t["read_bytes"] + t["write_bytes"])
    fp32_tflops = cells * FLOPS_PER_CELL / (fp32_ms * 1e-3) / 1e12
    roof = {"bound": "valu", "achieved": round(dom_tflops, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(dom_tflops / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
            "kernel": dom["name"], "kernel_ms": round(dom["ms"], 3), "kernel_cells": dom["cells"],
            "flops_per_cell": FLOPS_PER_CELL,
            "fp32_pass": {"achieved": round(fp32_tflops, 2), "frac": round(fp32_tflops / FP32_PEAK_TFLOPS, 4),
                          "ms": round(fp32_ms, 3), "cells": cells},
            "traffic_source": os.path.relpath(tpath, ROOT) if traffic else None}
    result = {
        "metric": METRIC,
        "value": round(value, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "rescue_dtype": "f64", "data": "synthetic",
        "config": {"workload": wname, "regions": len(regions), "pairs": int(info0["n_pairs"]),
                   "cells": int(cells), "graph": bool(args.graph), "seed": 1},
        "roofline": roof,
        "device_ms": {"fp32_pass": round(fp32_ms, 3), "fp64_rescue": round(fp64_ms, 3),
                      "launch_sequence_mean": round(float(np.mean(dev_total)), 3)},
        "rescue": {"pairs": n_rescued, "fraction": n_rescued / max(1, int(info0["n_pairs"])),
                   "cells": rescued_cells,
                   # FP64 pass: the rescued cells' algorithmic flops against the 78.6 TFLOP/s FP64 peak
                   "fp64_tflops": round(rescued_cells * FLOPS_PER_CELL / (fp64_ms * 1e-3) / 1e12, 2) if fp64_ms > 0 else None,
                   "fp64_frac": round(rescued_cells * FLOPS_PER_CELL / (fp64_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)
                   if fp64_ms > 0 else None},
        "kernels": [{"name": k["name"], "ms": round(k["ms"], 3), "cells": k["cells"], "tiles": k["tiles"],
                     "gcups": round(k["cells"] / (k["ms"] * 1e-3) / 1e9, 1) if k["ms"] > 0 and k["cells"] else None}
                    for k in kernels],
    }
    if not args.no_extras:
        log("plain one-region-per-launch GPU form")
        result["plain_gpu"] = plain_gpu(regions, min(256, len(regions)))
        result["plain_gpu"]["batched_over_plain"] = round(value / result["plain_gpu"]["value"], 2)
        log("accuracy sample")
        small = [x for x, r in enumerate(regions) if r["n_haps"] * r["n_reads"] <= 1200][:12]
        if small:
            result["accuracy"] = accuracy(regions, ebatch, out, small)
    rv.shutdown()
    if not args.no_served:
        # the same regions arriving one by one through the serving loop
        log("served stream")
        rates = [float(x) for x in args.rates.split(",")]
        srv, sst = served_stream(regions, [], rates, args.duration, args.producers, check=400)
        result["served"] = {"pool": wname, "producers": args.producers, "arrivals": "poisson",
                            "duration_s_per_rate": args.duration, "hbm_budget_gib": 64,
                            "warmup": "0.3 s of arrivals at each rate before its window", "rates": srv,
                            "highest_rate_p99_under_10ms": highest_rate_under(srv),
                            "accuracy": sst.get("served_check")}
        if "plain_gpu" in result:
            result["served"]["top_rate_over_plain"] = round(srv[-1]["gcups"] / result["plain_gpu"]["value"], 2)
    if not args.no_cpu:
        log("CPU baseline")
        try:
            result["cpu_baseline"] = cpu_baseline(regions, args.cpu_seconds)
            result["cpu_baseline"]["gpu_over_cpu"] = round(value / result["cpu_baseline"]["value"], 1)
            if "served" in result:
                result["served"]["top_rate_over_cpu"] = round(
                    result["served"]["rates"][-1]["gcups"] / result["cpu_baseline"]["value"], 1)
        except Exception as e:  # noqa: BLE001
            result["cpu_baseline"] = {"error": str(e)}
    if not args.no_configs and args.workload == "batched":
        # the other configs of the benchmark, each with its own baselines
        log("config: single region 8 x 32")
        result["single_region"] = single_region_config(args.no_cpu)
        log("config: precision stress, 2048 regions")
        result["precision_stress"] = stress_config(args)
        log("config: deep-coverage tail regions under serving")
        pool = [r for r in regions[:2048]]
        result["deep_tail"] = deep_config(pool, argparse.Namespace(rates="2000,8000,32000", duration=args.duration,
                                                                    producers=args.producers, no_cpu=args.no_cpu))
    if not args.no_extras and args.workload == "batched":
        # the next kernel on the same regions: every read realigned to its
        # best haplotype, as one batch and as one served call per region
        log("realignment aligner")
        try:
            sw = run_sw(argparse.Namespace(regions=1024, steps=3, no_cpu=args.no_cpu))
            calls = run_sw_calls(argparse.Namespace(regions=1024, steps=2, producers=32))
            result["realign"] = {"workload": sw["config"]["workload"], "pairs": sw["config"]["pairs"],
                                 "cells": sw["config"]["cells"], "dtype": sw["dtype"],
                                 "gcups_device": sw["value"], "device_ms": sw["device_ms"],
                                 "gcups_with_copies": sw["gcups_with_copies"], "parity": sw["parity"],
                                 "cpu_baseline": sw.get("cpu_baseline"),
                                 "served_call_per_region_32_workers": calls["served"],
                                 "batched_call_per_region_32_workers": calls["batched_call_per_region"]}
        except Exception as e:  # noqa: BLE001
            result["realign"] = {"error": str(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
