#!/usr/bin/env python3
"""Benchmark of the k-mer hot path on one MI355X.

Workload (one step = one plain GPU job on one sample): 20M 2x150 paired-end
reads from a 50k-transcript synthetic transcriptome (1% substitutions, 0.05%
N), resident in HBM as ASCII before the timed region.  The job is what
insilico_read_normalization.pl runs per sample: pack reads, size the table,
count canonical 25-mers of both.fa (jellyfish count -C), per-read coverage
stats of every read against the k-mers with count >= 2 (missing -> 1,
fastaToKmerCoverageStats --DS), and the max_cov=50 pair selection
(nbkc_merge_left_right_stats.pl + nbkc_normalize.pl).

Prints one JSON line.  value = k-mer windows counted and scored per second of
whole-job time.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "k-mers/s count+score vs plain GPU & host CPU; p50/p99 job latency, bit-exact"
HBM_PEAK_GBS = 8000.0
K = 25
READ_LEN = 150
# Algorithmic bytes (per k-mer window / per read) of the two dominant kernels.
# count: 0.375 B packed base + 64 B bucket line read + 64 B write-back; per read 12 B (len, offset)
# stats: 0.375 B packed base + 64 B bucket line read; per read 12 B in + 32 B stats out
COUNT_B_PER_WIN, COUNT_B_PER_READ = 128.375, 12.0
STATS_B_PER_WIN, STATS_B_PER_READ = 64.375, 44.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(tx, n_pairs_sample: int, max_cov: int) -> dict:
    """Oracle (single-threaded C restatement) on the first n pairs of the same stream."""
    from oracle import oracle as O
    from trinityrnaseq_amd import synth
    buf = np.empty(2 * n_pairs_sample * READ_LEN, np.uint8)
    synth.pe_reads(tx, n_pairs_sample, READ_LEN, seed=1002, out=buf)
    seq = buf.tobytes()
    off = np.arange(2 * n_pairs_sample + 1, dtype=np.uint64) * np.uint64(READ_LEN)
    t0 = time.perf_counter()
    keys, counts = O.count(seq, off, K, True)
    t1 = time.perf_counter()
    st = O.read_stats(seq, off, keys, counts, K, True, 2, 1, 0, 0)
    t2 = time.perf_counter()
    wins = 2 * n_pairs_sample * (READ_LEN - K + 1)
    return {"value": wins / (t2 - t0), "unit": "k-mers/s", "cores": 1, "kind": "port",
            "sample": f"first {n_pairs_sample} pairs ({2 * n_pairs_sample} reads) of the same "
                      f"stream; oracle count {t1 - t0:.2f}s + stats {t2 - t1:.2f}s",
            "count_s": t1 - t0, "stats_s": t2 - t1}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=20_000_000)
    ap.add_argument("--max-cov", type=int, default=50)
    ap.add_argument("--cpu-pairs", type=int, default=100_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-ref-gb", type=float, default=0.0,
                    help="also run one device-to-device copy of this many GB (HBM counter reference)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    from trinityrnaseq_amd import api, synth

    n_pairs = args.pairs
    n_reads = 2 * n_pairs
    t = time.perf_counter()
    tx = synth.large_transcriptome(1002)
    host = np.empty(n_reads * READ_LEN, np.uint8)
    synth.pe_reads(tx, n_pairs, READ_LEN, seed=1002, out=host)
    off = np.arange(n_reads + 1, dtype=np.uint64) * np.uint64(READ_LEN)
    log(f"generated {n_reads} reads in {time.perf_counter() - t:.1f}s")

    ctx = api.Context(0)
    d_seq = api.DeviceBuffer(ctx, host.nbytes + 64)
    d_off = api.DeviceBuffer(ctx, off.nbytes)
    d_seq.upload(host)
    d_off.upload(off)
    del host
    d_stats = api.DeviceBuffer(ctx, n_reads * api.READ_STAT_DTYPE.itemsize)
    d_sel = api.DeviceBuffer(ctx, n_pairs + 64)
    wins_per_step = n_reads * (READ_LEN - K + 1)
    log("inputs resident in HBM")
    if args.copy_ref_gb > 0:
        nb = int(args.copy_ref_gb * 1e9)
        a_, b_ = api.DeviceBuffer(ctx, nb), api.DeviceBuffer(ctx, nb)
        b_.copy_from(a_, nb)
        log(f"copy reference: {nb} bytes device-to-device")
        a_.free()
        b_.free()

    def job():
        reads = api.Reads.from_device(ctx, d_seq.ptr, d_off.ptr, n_reads)
        est = api.estimate_distinct(ctx, reads, K, True)
        table = api.Table(ctx, K, True, int(est * 1.05) + 4096, 0.5)
        table.count(reads)
        api.read_cov_stats_d(ctx, table, reads, d_stats.ptr, min_count=2)
        nsel = api.select_pairs_d(ctx, d_stats.ptr, d_stats.ptr + n_pairs * 32, n_pairs,
                                  args.max_cov, 1, 10000, READ_LEN - K + 1, d_sel.ptr)
        ctx.sync()
        info = table.info()
        table.free()
        reads.free()
        return nsel, est, info

    for i in range(args.warmup):
        t = time.perf_counter()
        nsel, est, info = job()
        log(f"warmup {i}: {time.perf_counter() - t:.3f}s selected={nsel} est_distinct={est:.3e} "
            f"capacity={info['capacity']}")
    ctx.timing(True)
    ctx.timing_reset()
    lat = []
    ctx.sync()
    for i in range(args.steps):
        t = time.perf_counter()
        nsel, est, info = job()
        lat.append(time.perf_counter() - t)
        log(f"step {i}: {lat[-1]:.3f}s selected={nsel}")
    ctx.sync()
    total = sum(lat)
    ms = 1000 * total / args.steps

    kern = {}
    for name in ["pack", "read_words", "hll", "count", "part_hist", "part_scatter", "part_count",
                 "add_keys", "compact", "dump_count", "ovf_fold", "stats", "select", "select_flags",
                 "scan_u64", "scan_u32", "windows", "rehash"]:
        tot, n = ctx.kernel_time(name)
        if n:
            kern[name] = {"total_ms": round(tot, 3), "launches": n, "avg_ms": round(tot / n, 4)}
    count_bytes = wins_per_step * COUNT_B_PER_WIN + n_reads * COUNT_B_PER_READ
    stats_bytes = wins_per_step * STATS_B_PER_WIN + n_reads * STATS_B_PER_READ
    roof = {}
    for name, by in (("count", count_bytes), ("part_count", count_bytes), ("stats", stats_bytes)):
        if name in kern:
            sec = kern[name]["total_ms"] / 1000.0 / args.steps
            roof[name] = {"achieved_gbs": by / sec / 1e9, "bytes_per_step": by,
                          "kernel_ms_per_step": sec * 1000}
    dom = max(roof, key=lambda n: roof[n]["kernel_ms_per_step"]) if roof else None
    roofline = None
    if dom:
        a = roof[dom]["achieved_gbs"]
        roofline = {"bound": "hbm", "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(a / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                    "other": {n: round(v["achieved_gbs"], 1) for n, v in roof.items()}}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(tx, args.cpu_pairs, args.max_cov)
        log(f"cpu baseline {cpu['value']:.3e} k-mers/s")

    lat_sorted = sorted(lat)
    p = lambda q: lat_sorted[min(len(lat_sorted) - 1, int(round(q * (len(lat_sorted) - 1))))]
    out = {
        "metric": METRIC, "value": wins_per_step / (ms / 1000.0), "unit": "k-mers/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded transcriptome + reads generated in-process)",
        "config": {"workload": f"large_pe: {n_pairs} 2x150 PE reads, 50k transcripts, k=25 canonical, "
                               f"count + coverage stats (dump -L 2 semantics) + max_cov={args.max_cov} selection",
                   "pairs": n_pairs, "k": K, "read_len": READ_LEN, "form": "plain per-job"},
        "reads_per_s": n_reads / (ms / 1000.0),
        "latency_s": {"p50": p(0.5), "p99": p(0.99), "min": lat_sorted[0], "max": lat_sorted[-1]},
        "kernels": kern, "roofline": roofline, "cpu_baseline": cpu,
        "selected_pairs": nsel, "distinct_estimate": est,
    }
    print(json.dumps(out), flush=True)
    for b in (d_seq, d_off, d_stats, d_sel):
        b.free()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
