#!/usr/bin/env python3
"""Benchmark of the k-mer hot path on one MI355X.

Default workload (--workload stress; one step = one plain GPU job on one
sample): 150M 2x150 paired-end reads from a 150k-transcript synthetic
transcriptome with heavy expression skew (lognormal sigma 2.5) and 1.5%
substitutions, resident in HBM as ASCII before the timed region: pack,
size, count canonical 25-mers (jellyfish count -C), coverage stats of every
read against the k-mers with count >= 2 (fastaToKmerCoverageStats --DS),
Inchworm's prune (min count 2, entropy >= 1.5) and ReadsToTranscripts voting
against 400k contig-like sequences.

--workload large_pe: the 20M-pair single sample (50k transcripts, 1%
substitutions, 0.05% N) through count, coverage stats and the max_cov=50
pair selection (nbkc_merge_left_right_stats.pl + nbkc_normalize.pl), as
insilico_read_normalization.pl runs it per sample.  --workload components /
serving: the batched and served forms.

Prints one JSON line.  value = k-mer windows processed per second of
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
# Algorithmic bytes of the table kernels.
# count with table atomics: 0.375 B packed base + 64 B bucket line read + 64 B write-back per
#   window; per read 12 B (len, offset)
# region-resident count, per merged word w (8 B) of the stream:
#   count_merge   0.375 B per packed base read + 8 B per word written
#   count_hist    8 B per word read, two passes (coarse bins, then regions)
#   count_scatter 8 B read + 8 B written per word, two passes
#   count_regions 8 B per word read + every table line written once (fresh table)
# stats: 0.375 B packed base + 64 B bucket line read; per read 12 B in + 32 B stats out
COUNT_B_PER_WIN, COUNT_B_PER_READ = 128.375, 12.0
STATS_B_PER_WIN, STATS_B_PER_READ = 64.375, 44.0
# vote: 0.375 B packed base + 64 B component-table line read; per read 12 B in + 12 B vote out
VOTE_B_PER_WIN, VOTE_B_PER_READ = 64.375, 24.0
# Memory-side bytes per window actually moved (FETCH_SIZE + WRITE_SIZE of one step over its
# 5.04 G windows, counters calibrated against kernels of the same access shape), from
# profiles/large_pe_20M_pairs_count_stats.md.  The counters cannot be read inside this
# process, so the profiled figures are carried here for roofline.traffic.
TRAFFIC_B_PER_WIN = {"count": (121.5 + 83.4) / 5.04, "stats": (143.49 + 1.61) / 5.04}
# The same for the stressed sample's kernels (37.8 G windows per step), from
# profiles/stress_150M_pairs_trace.md: FETCH_SIZE + WRITE_SIZE of one job at
# face value (their random 64-B line reads read as 64 B each, atomics as 32 B
# of WRITE_SIZE).
STRESS_TRAFFIC_B_PER_WIN = {"count": (880.30 + 574.82) / 37.8, "stats": (861.49 + 12.01) / 37.8,
                            "vote": (790.17 + 14.40) / 37.8, "count_merge": (87.79 + 115.15) / 37.8,
                            "count_scatter": (130.83 + 307.42) / 37.8, "count_hist": (116.33 + 14.80) / 37.8,
                            "count_regions": (103.86 + 17.54 + 129.94 + 64.96) / 37.8}


# every kernel name the library times (TK_LAUNCH / timing names)
KERNEL_NAMES = ["pack", "read_words", "hll", "sort_pairs", "minimizer", "count", "count_merge", "count_hist",
                "count_scatter", "count_bounds", "count_regions", "count_retry", "ovf_fold", "ovf_rehash", "rehash",
                "stats", "select", "select_flags", "scan_u32", "scan_u64", "windows", "gather_stats", "prune",
                "ctable_build", "vote", "dump", "dump_bins", "dump_count", "histo", "tickets"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(tx, n_pairs_sample: int, max_cov: int) -> dict:
    """Count + coverage stats of the first n pairs of the same stream on the host
    cores with the OpenMP CPU restatement (oracle/tk_cpuref.cpp: lock-free
    table, per-read nth_element median; checked against the oracle in tests).
    The reference binaries are not in the tree, so this stands in for them."""
    from oracle import oracle as O
    from trinityrnaseq_amd import synth
    buf = np.empty(2 * n_pairs_sample * READ_LEN, np.uint8)
    synth.pe_reads(tx, n_pairs_sample, READ_LEN, seed=1002, out=buf)
    off = np.arange(2 * n_pairs_sample + 1, dtype=np.uint64) * np.uint64(READ_LEN)
    threads = max(1, min(16, os.cpu_count() or 1))
    wins = 2 * n_pairs_sample * (READ_LEN - K + 1)
    r = O.cpuref_count_stats(buf, off, K, True, 2, threads, expected_distinct=wins // 3)
    sec = r["count_s"] + r["stats_s"]
    return {"value": wins / sec, "unit": "k-mers/s", "cores": threads, "kind": "port",
            "sample": f"first {n_pairs_sample} pairs ({2 * n_pairs_sample} reads, {wins} windows) of the same "
                      f"stream; CPU restatement (OpenMP, not the reference binaries): count {r['count_s']:.2f}s "
                      f"+ stats {r['stats_s']:.2f}s, {r['distinct']} distinct",
            "count_s": r["count_s"], "stats_s": r["stats_s"]}


def random_access_ceiling(ctx, table_bytes: int, wins_per_step: int, kern: dict, steps: int,
                          n_ops: int | None = None) -> dict:
    """Rates of random 64-byte line reads and random 8-byte atomics (add, CAS,
    workgroup-scope add) over a buffer the size of the step's count table and
    over smaller ones (tk_micro.hip), the XCD each workgroup runs on, and the
    count and scoring kernels' per-window rates."""
    import ctypes
    from trinityrnaseq_amd import api
    lib = ctypes.CDLL(str(ROOT / "trinityrnaseq_amd" / "lib" / "libtk_micro.so"))
    lib.tkm_run.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                            ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint64)]
    lib.tkm_run.restype = ctypes.c_int
    lib.tkm_xcc_map.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    nbytes = max(1 << 30, table_bytes) // 4096 * 4096
    buf = api.DeviceBuffer(ctx, nbytes)  # contents are irrelevant to the rates
    rates, ops = {}, {}
    for kind, name, nb in ((1, "line64_read", nbytes), (2, "line128_read", nbytes), (3, "atomic_add", nbytes),
                           (4, "atomic_add_ret", nbytes), (1, "line64_read_128MB", 128 << 20),
                           (1, "line64_read_2MB", 2 << 20), (4, "atomic_add_ret_128MB", 128 << 20),
                           (1, "line64_read_1GB", 1 << 30), (1, "line64_read_4GB", min(nbytes, 4 << 30)),
                           (5, "cas", nbytes), (6, "atomic_add_wg", nbytes),
                           (6, "atomic_add_wg_2MB", 2 << 20)):
        best = 0.0
        for _ in range(3):
            ms, done = ctypes.c_float(), ctypes.c_uint64()
            rc = lib.tkm_run(kind, ctypes.c_void_p(buf.ptr), nb, n_ops or wins_per_step, ctypes.byref(ms),
                             ctypes.byref(done))
            if rc:
                raise RuntimeError(f"tkm_run({name}) failed: {rc}")
            best = max(best, done.value / (ms.value / 1e3))
        rates[name] = best
        ops[name] = done.value
    # which XCD each workgroup of a launch runs on (b % 8 expected)
    nblk = 4096
    if lib.tkm_xcc_map(ctypes.c_void_p(buf.ptr), nblk):
        raise RuntimeError("tkm_xcc_map failed")
    xcc = np.empty(nblk, np.uint32)
    buf.download(xcc)
    buf.free()
    out = {"buffer_bytes": nbytes, "ops_per_launch": ops,
           "xcc_of_block_mod8_match": float(np.mean(xcc == (np.arange(nblk) % 8))),
           "xcc_ids": sorted(set(int(v) for v in xcc)),
           "line64_reads_per_s_128MB": rates["line64_read_128MB"],
           "line64_reads_per_s_2MB": rates["line64_read_2MB"],
           "returning_atomics_per_s_128MB": rates["atomic_add_ret_128MB"],
           "line64_reads_per_s_1GB": rates["line64_read_1GB"],
           "line64_reads_per_s_4GB": rates["line64_read_4GB"],
           "cas_per_s": rates["cas"],
           "workgroup_scope_atomics_per_s": rates["atomic_add_wg"],
           "workgroup_scope_atomics_per_s_2MB": rates["atomic_add_wg_2MB"],
           "line64_reads_per_s": rates["line64_read"],
           "line128_reads_per_s": rates["line128_read"],
           "atomics_per_s": rates["atomic_add"],
           "returning_atomics_per_s": rates["atomic_add_ret"]}
    # windows per second of the two table passes (the profile relates them to the ceilings
    # through the counters: atomics per window for count, L2 misses per window for stats)
    for name in ("count", "stats"):
        if name in kern:
            out[f"{name}_windows_per_s"] = wins_per_step / (kern[name]["total_ms"] / 1e3 / steps)
    return out


def make_components(n_comp: int, seed: int = 1003):
    """Chrysalis-style components: reads per component ~ power law (alpha 1.8)
    on [10, 200k]; 1-20 transcripts of 300-5000 nt; 150-nt reads at 1%
    error; contigs = 1-3 pieces of each transcript (at most 50)."""
    from trinityrnaseq_amd import synth
    rng = np.random.default_rng(seed)
    a = 1.8
    lo, hi = 10.0, 200_000.0
    u = rng.random(n_comp)
    sizes = (lo ** (1 - a) + u * (hi ** (1 - a) - lo ** (1 - a))) ** (1 / (1 - a))
    sizes = np.maximum(10, sizes.astype(np.int64))
    rbuf, cparts, rcount, ccount = [], [], [], []
    for c in range(n_comp):
        ntx = int(rng.integers(1, 21))
        tx = synth.Transcriptome(seed * 100_003 + c, ntx, 0, 300, 5000)
        s, _ = synth.se_reads(tx, int(sizes[c]), READ_LEN, seed=seed * 7 + c, sub_rate=0.01)
        rbuf.append(s)
        rcount.append(int(sizes[c]))
        contigs = []
        for j in range(ntx):
            t = tx.seq(j)
            cuts = sorted(rng.integers(50, max(51, len(t) - 50), size=int(rng.integers(0, 3))))
            b0 = 0
            for x in list(cuts) + [len(t)]:
                if x - b0 >= 30:
                    contigs.append(t[b0:x])
                b0 = x
        contigs = contigs[:50] or [tx.seq(0)]
        cparts.append(contigs)
        ccount.append(len(contigs))
    reads = np.concatenate(rbuf)
    roff = np.arange(int(sum(rcount)) + 1, dtype=np.uint64) * np.uint64(READ_LEN)
    cseq = "".join(x for cs in cparts for x in cs).encode()
    clen = [len(x) for cs in cparts for x in cs]
    coff = np.zeros(len(clen) + 1, np.uint64)
    coff[1:] = np.cumsum(clen)
    rco = np.zeros(n_comp + 1, np.uint64)
    rco[1:] = np.cumsum(rcount)
    cco = np.zeros(n_comp + 1, np.uint64)
    cco[1:] = np.cumsum(ccount)
    return reads, roff, rco, cseq, coff, cco


def bench_components(args) -> dict:
    from trinityrnaseq_amd import api
    t = time.perf_counter()
    reads, roff, rco, cseq, coff, cco = make_components(args.components)
    n_reads = len(roff) - 1
    log(f"generated {args.components} components, {n_reads} reads, {len(coff) - 1} contigs "
        f"in {time.perf_counter() - t:.1f}s")
    ctx = api.Context(0)
    r = api.Reads(ctx, reads.tobytes(), roff)
    cg = api.Reads(ctx, cseq, coff)
    wins = r.windows(K)
    b = api.Batch(ctx, r, rco, cg, cco, use_graph=True)
    d_st = api.DeviceBuffer(ctx, n_reads * 32)
    d_vo = api.DeviceBuffer(ctx, n_reads * 12)
    tuned_ms = b.autotune(d_st.ptr, d_vo.ptr, min_count=1)
    info = b.info()
    for _ in range(args.warmup):
        b.run_d(d_st.ptr, d_vo.ptr)
    ctx.sync()
    lat = []
    for _ in range(args.steps):
        t = time.perf_counter()
        b.run_d(d_st.ptr, d_vo.ptr)
        ctx.sync()
        lat.append(time.perf_counter() - t)
    batched_s = sum(lat) / len(lat)
    # plain form: one job per component (the ParaFly fan-out), on a prefix
    n_plain = min(args.components, args.plain_components)
    t = time.perf_counter()
    w_plain = 0
    for c in range(n_plain):
        a, z = int(rco[c]), int(rco[c + 1])
        rr = api.Reads(ctx, reads[a * READ_LEN:z * READ_LEN].tobytes(), roff[a:z + 1] - roff[a])
        w_plain += rr.windows(K)
        tb = api.Table(ctx, K, True, int(api.estimate_distinct(ctx, rr, K, True) * 1.05) + 64)
        tb.count(rr)
        api.read_cov_stats(ctx, tb, rr)
        ca, cz = int(cco[c]), int(cco[c + 1])
        cr = api.Reads(ctx, cseq[int(coff[ca]):int(coff[cz])], coff[ca:cz + 1] - coff[ca])
        ct = api.CTable(ctx, cr, np.arange(cz - ca, dtype=np.uint32), K, True)
        ct.vote(rr)
        tb.free(); ct.free(); rr.free(); cr.free()
    plain_s = time.perf_counter() - t
    plain_rate = w_plain / plain_s
    out = {
        "metric": METRIC + " [batched components]", "value": wins / batched_s, "unit": "k-mers/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(batched_s * 1000, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded components generated in-process)",
        "config": {"workload": f"components: {args.components} Chrysalis-style components (power law "
                               f"alpha=1.8 on [10,200k] reads), count + coverage stats + R2T vote per "
                               f"component, one graph replay", "reads": n_reads, "windows": wins,
                   "form": "batched (hipGraph)", "autotuned": info},
        "reads_per_s": n_reads / batched_s,
        "plain": {"components": n_plain, "k-mers/s": plain_rate, "seconds": plain_s,
                  "note": "one job per component through the per-sample entry points"},
        "batched_over_plain": (wins / batched_s) / plain_rate,
        "autotune_best_ms": tuned_ms,
    }
    for x in (d_st, d_vo):
        x.free()
    b.free()
    return out


def bench_stress(args) -> dict:
    """The default workload: the large stressed sample, the largest
    single-GPU configuration.  150M 2x150 PE reads (300M reads, 37.8 G
    windows) from 150k transcripts with heavy expression skew (lognormal sigma
    2.5), 1.5% substitutions, resident in HBM as ASCII before the timed
    region.  One step = one plain job: pack, size (HyperLogLog), count
    canonical 25-mers, coverage stats of every read (k-mers with count >= 2,
    missing -> 1), Inchworm's prune (min count 2, entropy >= 1.5) of the same
    table, and ReadsToTranscripts voting of every read against 400k
    contig-like sequences.  After the timed steps the same job runs from the
    reads in pageable host memory (upload included), for the PCIe-inclusive
    job time; its results must equal the device job's."""
    from trinityrnaseq_amd import api, synth
    n_pairs = int(args.stress_pairs)
    n_reads = 2 * n_pairs
    t = time.perf_counter()
    tx = synth.Transcriptome(1005, 150_000, 1, 1200.0, 0.6, 200, 15000, share_frac=0.10,
                             expr_sigma=2.5)
    host = np.empty(n_reads * READ_LEN, np.uint8)
    synth.pe_reads(tx, n_pairs, READ_LEN, seed=1005, sub_rate=0.015, out=host)
    off = np.arange(n_reads + 1, dtype=np.uint64) * np.uint64(READ_LEN)
    # 400k contig-like pieces: three overlapping thirds of each transcript, the
    # first 400k of them; component = transcript index
    pieces, comps = [], []
    for i in range(tx.n):
        sq = tx.seq(i)
        L = len(sq)
        for j in range(3):
            a, b = j * L // 3, min(L, (j + 2) * L // 3)
            if b - a >= K:
                pieces.append(sq[a:b])
                comps.append(i)
        if len(pieces) >= 400_000:
            break
    pieces, comps = pieces[:400_000], np.array(comps[:400_000], np.uint32)
    cseq, coff = api.pack_sequences(pieces)
    log(f"generated {n_reads} reads and {len(pieces)} contigs in {time.perf_counter() - t:.1f}s")
    ctx = api.Context(0)
    d_seq = api.DeviceBuffer(ctx, host.nbytes + 64)
    d_off = api.DeviceBuffer(ctx, off.nbytes)
    d_seq.upload(host)
    d_off.upload(off)
    d_stats = api.DeviceBuffer(ctx, n_reads * api.READ_STAT_DTYPE.itemsize)
    d_votes = api.DeviceBuffer(ctx, n_reads * api.R2T_VOTE_DTYPE.itemsize)
    keep = api.entropy_keep_table(K, 1.5)
    contigs = api.Reads(ctx, cseq, coff)
    wins = n_reads * (READ_LEN - K + 1)
    log("inputs resident in HBM")

    def assigned_reads() -> int:
        """Reads the vote assigned to a component (outside the timed region)."""
        v = np.empty(n_reads, api.R2T_VOTE_DTYPE)
        d_votes.download(v)
        return int((v["comp"] != 0xFFFFFFFF).sum())

    def step(reads_from_host: bool):
        if reads_from_host:
            reads = api.Reads(ctx, host, off)
        else:
            reads = api.Reads.from_device(ctx, d_seq.ptr, d_off.ptr, n_reads)
        est = api.estimate_distinct(ctx, reads, K, True)
        table = api.Table(ctx, K, True, int(est * 1.05) + 4096, 0.5)
        table.count(reads)
        api.read_cov_stats_d(ctx, table, reads, d_stats.ptr, min_count=2)
        kept = table.prune(2, keep)  # Inchworm's dictionary, after the stats used the full counts
        ct = api.CTable(ctx, contigs, comps, K, True)
        ct.vote_d(reads, d_votes.ptr)  # results stay in HBM, like the stats
        ctx.sync()
        info = table.info()
        ct.free()
        table.free()
        reads.free()
        return est, kept, info

    # one job after another in one process: each job's table and read set take the blocks the
    # previous job freed (no hipFree + hipMalloc of a 65 GB table per job)
    api.thread_cache(200 << 30)
    for i in range(args.warmup):
        t = time.perf_counter()
        est, kept, info = step(False)
        log(f"warmup {i}: {time.perf_counter() - t:.3f}s est_distinct={est:.3e} kept={kept} "
            f"capacity={info['capacity']}")
    ctx.timing(True)
    ctx.timing_reset()
    lat = []
    ctx.sync()
    for i in range(args.steps):
        t = time.perf_counter()
        est, kept, info = step(False)
        lat.append(time.perf_counter() - t)
        log(f"step {i}: {lat[-1]:.3f}s")
    ctx.sync()
    assigned = assigned_reads()
    kern = {}
    for name in ["pack", "read_words", "hll", "sort_pairs", "count", "count_merge", "count_hist", "count_scatter",
                 "count_bounds", "count_regions", "count_retry", "stats", "prune", "ctable_build", "vote", "ovf_fold",
                 "scan_u64"]:
        tot, n = ctx.kernel_time(name)
        if n:
            kern[name] = {"total_ms": round(tot, 3), "launches": n, "avg_ms": round(tot / n, 4)}
    words = ctx.count_words()  # merged words of the last step's count (0: table atomics)
    ctx.timing(False)
    ms = 1000 * sum(lat) / len(lat)
    # dominant kernel by time per step, priced at its algorithmic bytes
    stage_bytes = {"count": wins * COUNT_B_PER_WIN + n_reads * COUNT_B_PER_READ,
                   "stats": wins * STATS_B_PER_WIN + n_reads * STATS_B_PER_READ,
                   "vote": wins * VOTE_B_PER_WIN + n_reads * VOTE_B_PER_READ}
    if words:
        # region-resident count in read chunks: each chunk's region pass reads and writes the
        # table's lines once (the first, on a cleared table, only writes them)
        chunks = kern.get("count_regions", {}).get("launches", args.steps) / args.steps
        table_bytes = info["capacity"] * 8
        stage_bytes.update({"count_merge": n_reads * READ_LEN * 0.375 + 8 * words,
                            "count_hist": 2 * 8 * words, "count_scatter": 2 * 16 * words,
                            "count_regions": 8 * words + table_bytes * (2 * chunks - 1)})
    roof = {}
    for name, by in stage_bytes.items():
        if name in kern:
            sec = kern[name]["total_ms"] / 1000.0 / args.steps
            roof[name] = {"gbs": by / sec / 1e9, "bytes_per_step": by, "ms": sec * 1000,
                          "launches_per_step": kern[name]["launches"] / args.steps}
    dom = max(roof, key=lambda n: roof[n]["ms"])
    per_launch = roof[dom]["bytes_per_step"] / roof[dom]["launches_per_step"]
    avg_s = kern[dom]["avg_ms"] / 1000.0
    achieved = per_launch / avg_s / 1e9
    traffic = STRESS_TRAFFIC_B_PER_WIN.get(dom)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(traffic * wins / roof[dom]["launches_per_step"]) if traffic else None,
                "traffic_source": "profiles/stress_150M_pairs_trace.md (FETCH_SIZE + WRITE_SIZE per dispatch)",
                "kernel": dom, "algorithmic_bytes_per_launch": round(per_launch),
                "avg_launch_ms": kern[dom]["avg_ms"],
                "other": {n: {"gbs": round(v["gbs"], 1), "ms_per_step": round(v["ms"], 3)} for n, v in roof.items()},
                "merged_words": words}
    # the same job from pageable host memory: upload + pack inside the job (not `value`)
    host_lat = []
    for i in range(args.host_steps):
        t = time.perf_counter()
        est_h, kept_h, info_h = step(True)
        host_lat.append(time.perf_counter() - t)
        log(f"host-input job {i}: {host_lat[-1]:.3f}s")
        if (kept_h, assigned_reads(), info_h["capacity"]) != (kept, assigned, info["capacity"]):
            raise RuntimeError("the host-input job's results differ from the device job's")
    dump = None
    if not args.no_dump_check:
        # `jellyfish count` then `dump -L 1` of the whole table (Trinity:2612-2625): more keys
        # than one device sort call takes, so the dump goes out in sorted key ranges
        t = time.perf_counter()
        reads = api.Reads.from_device(ctx, d_seq.ptr, d_off.ptr, n_reads)
        table = api.Table(ctx, K, True, int(est * 1.05) + 4096, 0.5)
        table.count(reads)
        n_dump = table.dump_size()
        hist = table.histo()
        keys, counts = table.dump()
        dump_s = time.perf_counter() - t
        ok = len(keys) == n_dump == int(hist.sum())
        step_ = 1 << 27
        for a in range(0, len(keys), step_):
            z = min(len(keys), a + step_ + 1)
            ok = ok and bool(np.all(keys[a + 1:z] > keys[a:z - 1]))
        dump = {"entries": int(n_dump), "sorted_and_complete": ok, "count_dump_histo_s": round(dump_s, 3),
                "counts_sum": int(counts.sum(dtype=np.uint64))}
        log(f"dump check: {dump}")
        del keys, counts
        table.free()
        reads.free()
        if not ok:
            raise RuntimeError("the full dump is not complete and sorted")
    for b in (d_seq, d_off, d_stats, d_votes):
        b.free()
    contigs.free()
    api.thread_cache(0)
    if args.ceiling:
        # random 64-B line reads / atomics over a buffer the size of this job's count table
        roofline["ceiling"] = random_access_ceiling(ctx, info["capacity"] * 8, wins, kern, args.steps,
                                                    n_ops=wins // 8)
    ctx.close()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_stress(tx, args.cpu_pairs)
        log(f"cpu baseline {cpu['value']:.3e} k-mers/s")
    return {
        "metric": METRIC, "value": wins / (ms / 1000.0), "unit": "k-mers/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded transcriptome + reads generated in-process)",
        "config": {"workload": f"stress: {n_pairs} 2x150 PE reads, 150k transcripts (expression lognormal "
                               f"sigma 2.5), 1.5% error, k=25 canonical: count + coverage stats (min count 2) "
                               f"+ Inchworm prune (min 2, entropy 1.5) + R2T vs {len(pieces)} contigs",
                   "pairs": n_pairs, "contigs": len(pieces), "k": K, "read_len": READ_LEN,
                   "form": "plain per-job, inputs resident in HBM",
                   "device_blocks": "reused from job to job (tk_thread_cache)"},
        "reads_per_s": n_reads / (ms / 1000.0),
        "latency_s": percentiles(lat),
        "host_input_job_s": dict(percentiles(host_lat), **{
            "k-mers/s_p50": wins / percentiles(host_lat)["p50"] if host_lat else None,
            "input": "pageable host memory (staged through pinned buffers): upload + pack inside the job"}),
        "kernels": kern, "roofline": roofline, "cpu_baseline": cpu,
        "distinct_estimate": est, "kept_after_prune": kept, "full_dump": dump,
        "table_capacity": info["capacity"], "reads_assigned": assigned,
    }


def percentiles(xs) -> dict:
    if not xs:
        return {"p50": None, "p99": None, "n": 0}
    v = sorted(xs)
    p = lambda q: v[min(len(v) - 1, int(round(q * (len(v) - 1))))]  # noqa: E731
    return {"p50": p(0.5), "p99": p(0.99), "min": v[0], "max": v[-1], "n": len(v)}


def cpu_baseline_stress(tx, n_pairs_sample: int) -> dict:
    """Count + coverage stats of the first n pairs of the stressed stream with
    the OpenMP CPU restatement (oracle/tk_cpuref.cpp) on the host cores this
    process is given (OMP_NUM_THREADS when set, else every core of its
    affinity mask); the median of three runs (warm page cache)."""
    from oracle import oracle as O
    from trinityrnaseq_amd import synth
    buf = np.empty(2 * n_pairs_sample * READ_LEN, np.uint8)
    synth.pe_reads(tx, n_pairs_sample, READ_LEN, seed=1005, sub_rate=0.015, out=buf)
    off = np.arange(2 * n_pairs_sample + 1, dtype=np.uint64) * np.uint64(READ_LEN)
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    wins = 2 * n_pairs_sample * (READ_LEN - K + 1)
    runs = []
    for _ in range(3):
        r = O.cpuref_count_stats(buf, off, K, True, 2, threads, expected_distinct=wins // 2)
        runs.append(r)
    runs.sort(key=lambda r: r["count_s"] + r["stats_s"])
    r = runs[1]
    sec = r["count_s"] + r["stats_s"]
    return {"value": wins / sec, "unit": "k-mers/s", "cores": threads, "kind": "port",
            "sample": f"first {n_pairs_sample} pairs ({2 * n_pairs_sample} reads, {wins} windows) of the same "
                      f"stream, count + coverage stats (min count 2); CPU restatement (OpenMP, not the reference "
                      f"binaries), median of 3 runs: count {r['count_s']:.2f}s + stats {r['stats_s']:.2f}s, "
                      f"{r['distinct']} distinct",
            "runs_s": [round(x["count_s"] + x["stats_s"], 3) for x in runs]}


def bench_inchworm(args) -> dict:
    """Inchworm's greedy contigs (tk_inchworm) over the dictionary of the
    small workload (500 transcripts, 200k 76-nt single-end reads, 0.5% errors)
    and of a 500k-pair slice of the 50k-transcript stream, after Trinity's
    phase-1 prune (entropy >= 1.0); seeds need count >= 2 and entropy >= 1.5.
    Each run is checked against the oracle's serial greedy (which is also the
    single-thread CPU figure)."""
    from oracle import oracle as O
    from trinityrnaseq_amd import api, synth
    ctx = api.Context(0)
    out = {}
    def small():
        return synth.se_reads(synth.small_transcriptome(1001), 200_000, 76, seed=1001)

    def large():
        n = 500_000
        buf = np.empty(2 * n * READ_LEN, np.uint8)
        synth.pe_reads(synth.large_transcriptome(1002), n, READ_LEN, seed=1002, out=buf)
        return buf, np.arange(2 * n + 1, dtype=np.uint64) * np.uint64(READ_LEN)

    cases = [("small_se_200k_reads", small), ("large_pe_500k_pairs", large)]
    keep1 = api.entropy_keep_table(K, 1.0)
    keep15 = api.entropy_keep_table(K, 1.5)
    for name, gen in cases:
        seq, off = gen()
        reads = api.Reads(ctx, seq, off)
        del seq
        est = api.estimate_distinct(ctx, reads, K, True)
        table = api.Table(ctx, K, True, int(est * 1.05) + 4096, 0.5)
        table.count(reads)
        table.prune(1, keep1)
        keys, counts = table.dump()
        api.inchworm(ctx, table, 2, keep15)  # warm-up
        lat = []
        for _ in range(args.steps):
            t = time.perf_counter()
            contigs, sums, seeds, info = api.inchworm(ctx, table, 2, keep15)
            lat.append(time.perf_counter() - t)
        log(f"inchworm {name}: GPU p50 {percentiles(lat)['p50']:.3f}s, {len(keys)} k-mers; serial oracle next")
        t = time.perf_counter()
        oc, osum, oseeds = O.inchworm(keys, counts, K, True, 2, keep15)
        cpu_s = time.perf_counter() - t
        if oc != contigs or not np.array_equal(osum, sums):
            raise RuntimeError(f"inchworm {name}: contigs differ from the oracle")
        g = percentiles(lat)["p50"]
        nk = int(sum(len(c) - K + 1 for c in contigs))
        out[name] = {"dictionary_kmers": int(len(keys)), "contigs": len(contigs),
                     "contigs_ge_200nt": int(sum(len(c) >= 200 for c in contigs)), "contig_kmers": nk,
                     "gpu_s_p50": round(g, 4), "gpu_kmers_per_s": nk / g, "rounds": info["rounds"],
                     "walks": info["walks"], "cpu_oracle_s": round(cpu_s, 3),
                     "cpu_kmers_per_s": nk / cpu_s, "gpu_over_cpu": cpu_s / g}
        log(f"inchworm {name}: {out[name]}")
        table.free()
        reads.free()
    if args.iworm_pairs:
        # Trinity's phase-1 Inchworm on the dictionary of a whole large sample (every k-mer of
        # count >= 1 with entropy >= 1.0: Trinity:117,1050,2625), too large for the serial
        # oracle: checked through invariants on the device instead (no k-mer in two places,
        # every contig k-mer in the dictionary)
        n = args.iworm_pairs
        buf = np.empty(2 * n * READ_LEN, np.uint8)
        synth.pe_reads(synth.large_transcriptome(1002), n, READ_LEN, seed=1002, out=buf)
        reads = api.Reads(ctx, buf, np.arange(2 * n + 1, dtype=np.uint64) * np.uint64(READ_LEN))
        del buf
        est = api.estimate_distinct(ctx, reads, K, True)
        table = api.Table(ctx, K, True, int(est * 1.05) + 4096, 0.5)
        table.count(reads)
        reads.free()
        kept = table.prune(1, keep1)
        log(f"inchworm full: {kept} k-mers in the dictionary; contigs next")
        t = time.perf_counter()
        contigs, sums, seeds, info = api.inchworm(ctx, table, 2, keep15)
        g = time.perf_counter() - t
        nk = int(sum(len(c) - K + 1 for c in contigs))
        log(f"inchworm full: {len(contigs)} contigs, {nk} contig k-mers in {g:.2f}s ({info['rounds']} rounds)")
        cr = api.Reads.from_list(ctx, contigs)
        ct = api.Table(ctx, K, True, nk + 4096, 0.5)
        ct.count(cr)
        ck, cc = ct.dump()
        once = len(ck) == nk and int(cc.max(initial=0)) <= 1
        inside = bool(np.all(table.lookup(ck) >= 1))
        if not (once and inside):
            raise RuntimeError("inchworm full: a k-mer sits in two places or outside the dictionary")
        out[f"large_pe_{n}_pairs"] = {"dictionary_kmers": int(kept), "contigs": len(contigs),
                                       "contigs_ge_200nt": int(sum(len(c) >= 200 for c in contigs)),
                                       "contig_kmers": nk, "gpu_s": round(g, 3), "gpu_kmers_per_s": nk / g,
                                       "rounds": info["rounds"], "walks": info["walks"],
                                       "checked": "every contig k-mer once, and in the dictionary"}
        for o in (ct, cr, table):
            o.free()
    ctx.close()
    first = out["large_pe_500k_pairs"]
    return {"metric": METRIC + " [inchworm contigs]", "value": first["gpu_kmers_per_s"], "unit": "contig k-mers/s",
            "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": round(1000 * first["gpu_s_p50"], 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
            "data": "synthetic (seeded transcriptome + reads generated in-process)",
            "config": {"workload": "inchworm: greedy contigs over the pruned 25-mer dictionary (entropy 1.0; "
                                   "seeds count >= 2, entropy >= 1.5)"},
            "cases": out,
            "cpu_baseline": {"value": first["cpu_kmers_per_s"], "unit": "contig k-mers/s", "cores": 1,
                             "kind": "port", "sample": "the same dictionary, oracle/tk_oracle_iworm.c (serial)"}}


def _kmer_keys(codes: np.ndarray, kk: int, canonical: bool = True) -> tuple[np.ndarray, np.ndarray]:
    """(lo, hi) words of the kk-mers in the rows of `codes` (n x kk, values 0-3),
    as tk_kset reads them: the 2kk-bit value's low 64 bits and the bits above;
    canonical: the smaller of the k-mer and its reverse complement."""
    def words(c):
        lo = np.zeros(len(c), np.uint64)
        hi = np.zeros(len(c), np.uint64)
        for i in range(kk):
            sh = 2 * (kk - 1 - i)
            v = c[:, i].astype(np.uint64)
            if sh >= 64:
                hi |= v << np.uint64(sh - 64)
            else:
                lo |= v << np.uint64(sh)
                if sh > 62:
                    hi |= v >> np.uint64(64 - sh)
        return lo, hi
    lo, hi = words(codes)
    if canonical:
        rlo, rhi = words(3 - codes[:, ::-1])
        use_rc = (rhi < hi) | ((rhi == hi) & (rlo < lo))
        lo, hi = np.where(use_rc, rlo, lo), np.where(use_rc, rhi, hi)
    return lo, hi


def bench_kset(args) -> dict:
    """Counts of a fixed k-mer set (tk_kset), the GPU part of GraphFromFasta
    (Trinity:2180-2196) and QuantifyGraph (Trinity:2348-2356):
    welds: 2 M distinct canonical 48-mers drawn from the 50k-transcript set
    (candidate welds, each spanning two contigs) counted over every 48-mer
    window of the first --kset-pairs pairs of the 20M-pair stream, reads
    resident in HBM; nodes: QuantifyGraph --batch over the first 2,000
    components of the components workload, every distinct 24-mer of each
    component's contigs tagged by the component, counted over the
    component's reads.  Each checked against the oracle on a read sample,
    whose single-thread time is the CPU baseline."""
    from trinityrnaseq_amd import api, synth
    from oracle import oracle as O  # the checker and the CPU baseline only
    ctx = api.Context(0)
    rng = np.random.default_rng(4801)
    t = time.perf_counter()
    tx = synth.large_transcriptome(1002)
    n_pairs = args.kset_pairs
    host = np.empty(2 * n_pairs * READ_LEN, np.uint8)
    synth.pe_reads(tx, n_pairs, READ_LEN, seed=1002, out=host)
    off = np.arange(2 * n_pairs + 1, dtype=np.uint64) * np.uint64(READ_LEN)
    lut = np.full(256, 0, np.uint8)
    for ch, v in zip(b"ACGT", range(4)):
        lut[ch] = v
    picks = []
    while sum(len(p) for p in picks) < 2_200_000:
        j = int(rng.integers(0, tx.n))
        sq = tx.seq(j).encode()
        if len(sq) < 48:
            continue
        starts = rng.integers(0, len(sq) - 47, size=64)
        arr = np.frombuffer(sq, np.uint8)
        picks.append(lut[arr[starts[:, None] + np.arange(48)]])
    codes = np.concatenate(picks)
    lo, hi = _kmer_keys(codes, 48)
    pairs = np.unique(np.stack([hi, lo], axis=1), axis=0)[:2_000_000]
    keys = np.ascontiguousarray(np.stack([pairs[:, 1], pairs[:, 0]], axis=1))
    log(f"kset: {len(keys)} weld 48-mers, {2 * n_pairs} reads in {time.perf_counter() - t:.1f}s")
    reads = api.Reads(ctx, host, off)
    ks = api.KSet(ctx, keys)
    windows = 2 * n_pairs * (READ_LEN - 48 + 1)
    ks.count(reads, 48, True)  # warm-up
    ctx.timing(True)
    ctx.timing_reset()
    lat = []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        counts = ks.count(reads, 48, True)
        lat.append(time.perf_counter() - t1)
    kms, kn = ctx.kernel_time("kset_count")
    ctx.timing(False)
    step_s = sum(lat) / len(lat)
    log(f"kset welds: {step_s * 1000:.1f} ms per count, kernel {kms / kn:.1f} ms, {int((counts > 0).sum())} supported")
    # parity on a sample and the CPU baseline (single-thread oracle)
    n_s = 200_000
    s_seq = host[:n_s * READ_LEN].tobytes()
    s_off = off[:n_s + 1]
    s_reads = api.Reads(ctx, host[:n_s * READ_LEN], s_off)
    got = ks.count(s_reads, 48, True)
    t1 = time.perf_counter()
    want = O.kset_count(keys, s_seq, s_off, 48, True)
    cpu_s = time.perf_counter() - t1
    if not np.array_equal(got, want):
        raise RuntimeError("weld counts differ from the oracle's on the sample")
    s_reads.free()
    ks.free()
    reads.free()
    algo = windows * 64.375 + 2 * n_pairs * 12.0  # bucket line per window; length + offset per read
    kernel_s = kms / kn / 1000.0
    # QuantifyGraph --batch: tagged (component, 24-mer) node keys of 2,000 components
    r_, ro_, rco, cseq, coff, cco = make_components(10_000)
    nc = 2000
    nr = int(rco[nc])
    q_reads = api.Reads(ctx, r_[:int(ro_[nr])], ro_[:nr + 1])
    tags = np.repeat(np.arange(nc, dtype=np.uint32), np.diff(rco[:nc + 1]).astype(np.int64))
    ca = lut[np.frombuffer(cseq, np.uint8)[:int(coff[int(cco[nc])])]]
    n_ct = int(cco[nc])
    clen = np.diff(coff[:n_ct + 1]).astype(np.int64)
    nwin = np.maximum(clen - 23, 0)
    # window starts of every contig of the first nc components, and their component
    first = np.repeat(coff[:n_ct].astype(np.int64) - np.cumsum(nwin) + nwin, nwin)
    starts = first + np.arange(int(nwin.sum()), dtype=np.int64)
    comp_of = np.repeat(np.repeat(np.arange(nc, dtype=np.uint64), np.diff(cco[:nc + 1]).astype(np.int64)), nwin)
    lo24, _ = _kmer_keys(np.lib.stride_tricks.sliding_window_view(ca, 24)[starts], 24)
    nk = np.unique(np.stack([lo24, comp_of], axis=1), axis=0)
    qs = api.KSet(ctx, nk)
    qs.count(q_reads, 24, True, tags)
    ctx.timing(True)
    ctx.timing_reset()
    t1 = time.perf_counter()
    qs.count(q_reads, 24, True, tags)
    q_s = time.perf_counter() - t1
    q_kms, _ = ctx.kernel_time("kset_count")
    ctx.timing(False)
    q_windows = nr * (READ_LEN - 24 + 1)
    # the oracle on the first 200 components
    n_q = int(rco[200])
    qwant = O.kset_count(nk, r_[:int(ro_[n_q])].tobytes(), ro_[:n_q + 1], 24, True, tags[:n_q])
    qs2 = api.KSet(ctx, nk)
    q_reads2 = api.Reads(ctx, r_[:int(ro_[n_q])], ro_[:n_q + 1])
    if not np.array_equal(qs2.count(q_reads2, 24, True, tags[:n_q]), qwant):
        raise RuntimeError("node weights differ from the oracle's on the first 200 components")
    for o in (qs, qs2, q_reads, q_reads2):
        o.free()
    log(f"kset nodes: {len(nk)} tagged 24-mers, {q_windows} windows in {q_s * 1000:.1f} ms "
        f"(count kernel {q_kms:.1f} ms)")
    ctx.close()
    return {
        "metric": METRIC + " [fixed k-mer set: weld 48-mers]", "value": windows / step_s, "unit": "k-mers/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": round(step_s * 1000, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u128",
        "data": "synthetic (seeded transcriptome + reads generated in-process)",
        "config": {"workload": f"GraphFromFasta weld support: {len(keys)} canonical 48-mers over {2 * n_pairs} "
                               f"2x150 reads ({windows} windows), reads resident in HBM",
                   "welds": int(len(keys)), "reads": 2 * n_pairs},
        "latency_s": percentiles(lat),
        "roofline": {"bound": "hbm", "kernel": "kset_count", "achieved": round(algo / kernel_s / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(algo / kernel_s / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "avg_launch_ms": round(kms / kn, 3),
                     "algorithmic_bytes_per_launch": round(algo)},
        "supported_welds": int((counts > 0).sum()),
        "cpu_baseline": {"value": n_s * (READ_LEN - 48 + 1) / cpu_s, "unit": "k-mers/s", "cores": 1,
                         "kind": "port", "sample": f"first {n_s} reads, the same 2 M keys (or_kset_count, "
                                                   f"single-threaded C)"},
        "nodes": {"workload": f"QuantifyGraph --batch: {nc} components, {len(nk)} (component, 24-mer) node keys, "
                              f"{nr} reads", "ms": round(q_s * 1000, 3), "kernel_ms": round(q_kms, 3),
                  "k-mers/s_kernel": q_windows / (q_kms / 1000.0),
                  "note": "ms includes gathering and downloading one u64 count per node key",
                  "checked": "first 200 components equal to the oracle's"},
    }


def plain_job(ctx, seq: np.ndarray, off: np.ndarray, n: int, max_cov: int):
    """One sample through the plain per-job API: upload + pack, size, count,
    stats, select, selected flags back to the host, free everything."""
    from trinityrnaseq_amd import api
    reads = api.Reads(ctx, seq, off)
    est = api.estimate_distinct(ctx, reads, K, True)
    table = api.Table(ctx, K, True, int(est * 1.05) + 4096, 0.5)
    table.count(reads)
    d_stats = api.DeviceBuffer(ctx, 2 * n * api.READ_STAT_DTYPE.itemsize)
    d_sel = api.DeviceBuffer(ctx, n + 64)
    api.read_cov_stats_d(ctx, table, reads, d_stats.ptr, min_count=2)
    max_nwin = max(0, reads.info()[2] - K + 1)
    nsel = api.select_pairs_d(ctx, d_stats.ptr, d_stats.ptr + n * api.READ_STAT_DTYPE.itemsize, n, max_cov,
                              1, 10000, max_nwin, d_sel.ptr)
    sel = np.empty(n, np.uint8)
    d_sel.download(sel)
    for o in (d_sel, d_stats, table, reads):
        o.free()
    return nsel, sel


def bench_serving(args) -> dict:
    from trinityrnaseq_amd import api, synth
    # both forms count with table atomics (the server's workers always do), so
    # the comparison isolates the serving machinery: overlapped uploads, slots,
    # admission, block reuse
    os.environ["TK_RESIDENT"] = "0"
    rng = np.random.default_rng(2000)
    lo, hi = 0.5e6 * args.serve_scale, 30e6 * args.serve_scale
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), size=args.samples)).astype(np.int64)
    t = time.perf_counter()
    samples = []
    for j, n in enumerate(sizes):
        tx = synth.large_transcriptome(2000 + j)  # 50k transcripts, as the single large sample
        # ordinary (pageable) host memory, as a client holds parsed reads; --serve-pinned puts
        # the samples in page-locked memory instead
        pinned = api.HostBuffer(2 * int(n) * READ_LEN) if args.serve_pinned else None
        buf = pinned.array if pinned is not None else np.empty(2 * int(n) * READ_LEN, np.uint8)
        synth.pe_reads(tx, int(n), READ_LEN, seed=2000 + j, out=buf)
        off = np.arange(2 * int(n) + 1, dtype=np.uint64) * np.uint64(READ_LEN)
        samples.append((buf, off, int(n), pinned))
        if (j + 1) % 16 == 0:
            log(f"generated {j + 1} of {args.samples} samples")
    total_pairs = int(sizes.sum())
    log(f"generated {args.samples} samples, {total_pairs} pairs in {time.perf_counter() - t:.1f}s")
    # plain per-job GPU form, one sample at a time; like the server's workers it keeps the
    # device blocks a job frees for the next job (tk_thread_cache), so the comparison is of the
    # serving machinery (overlap, slots, admission), not of block reuse
    ctx = api.Context(0)
    api.thread_cache(int(args.hbm_budget_gb * 2**30))
    plain_job(ctx, samples[0][0], samples[0][1], samples[0][2], args.max_cov)  # warm-up
    ctx.timing(True)
    ctx.timing_reset()
    t = time.perf_counter()
    plain_lat = []
    plain_sel = []
    for b, o, n, _ in samples:
        t1 = time.perf_counter()
        plain_sel.append(plain_job(ctx, b, o, n, args.max_cov))
        plain_lat.append(time.perf_counter() - t1)
    plain_s = time.perf_counter() - t
    # kernel time of the whole stream's jobs (HIP events on the context's stream): what no
    # overlap of uploads and host work can hide
    gpu_s = sum(ctx.kernel_time(name)[0] for name in KERNEL_NAMES) / 1000.0
    ctx.timing(False)
    log(f"plain serial: kernels {gpu_s:.3f}s of {plain_s:.3f}s")
    api.thread_cache(0)
    ctx.close()
    log(f"plain serial: {plain_s:.3f}s")
    # the server with one slot (same code path as served, no overlap)
    srv1 = api.Server(slots=1, max_cov=args.max_cov, hbm_budget_bytes=int(args.hbm_budget_gb * 2**30))
    srv1.wait(srv1.submit(samples[0][0], samples[0][1], samples[0][2]))
    t = time.perf_counter()
    for b, o, n, _ in samples:
        srv1.wait(srv1.submit(b, o, n))
    slot1_s = time.perf_counter() - t
    srv1.close()
    log(f"server 1 slot: {slot1_s:.3f}s")
    # served: burst for capacity, then Poisson arrivals at 0.8 x capacity
    srv = api.Server(slots=args.slots, max_cov=args.max_cov, hbm_budget_bytes=int(args.hbm_budget_gb * 2**30))
    srv.wait(srv.submit(samples[0][0], samples[0][1], samples[0][2]))
    t = time.perf_counter()
    ids = [srv.submit(b, o, n) for b, o, n, _ in samples]
    served_sel = []
    for jid in ids:
        res, flags = srv.wait(jid)
        served_sel.append((res.n_selected, flags))
    burst_s = time.perf_counter() - t
    log(f"served burst: {burst_s:.3f}s")
    for (ns_p, f_p), (ns_s, f_s) in zip(plain_sel, served_sel):
        if ns_p != ns_s or not np.array_equal(f_p, f_s):
            raise RuntimeError("a served selection differs from the plain per-job one")
    lam = 0.8 * args.samples / burst_s
    gaps = rng.exponential(1.0 / lam, size=args.samples)
    t0 = time.perf_counter()
    ids = []
    for (b, o, n, _), g in zip(samples, gaps):
        time.sleep(g)
        ids.append(srv.submit(b, o, n))
    lat = []
    for jid in ids:
        res, _ = srv.wait(jid)
        lat.append(res.latency_s)
    poisson_s = time.perf_counter() - t0
    srv.close()
    for s_ in samples:
        if s_[3] is not None:
            s_[3].free()
    lat = sorted(lat)
    q = lambda v, x: v[min(len(v) - 1, int(round(x * (len(v) - 1))))]
    return {
        "metric": METRIC + " [served stream]", "value": 2 * total_pairs / burst_s, "unit": "reads/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(burst_s * 1000, 3),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded samples generated in-process, in "
                + ("page-locked" if args.serve_pinned else "pageable") + " host memory)",
        "config": {"workload": f"serving: {args.samples} samples, read pairs log-uniform "
                               f"[{lo:.0f}, {hi:.0f}] (scale {args.serve_scale} of 0.5M-30M), 2x150 PE, "
                               f"count + coverage stats + max_cov={args.max_cov} selection per sample",
                   "slots": args.slots, "hbm_budget_gb": args.hbm_budget_gb, "total_pairs": total_pairs},
        "serial_plain_reads_per_s": 2 * total_pairs / plain_s,
        "stream_kernel_s": round(gpu_s, 3),
        "served_over_kernel_bound": round(gpu_s / burst_s, 3),
        "served_over_serial": plain_s / burst_s,
        "server_1slot_reads_per_s": 2 * total_pairs / slot1_s,
        "served_over_1slot": slot1_s / burst_s,
        "selections_match_plain": "every sample's selection flags equal",
        "poisson": {"lambda_per_s": lam, "p50_s": q(lat, 0.5), "p99_s": q(lat, 0.99),
                    "reads_per_s": 2 * total_pairs / poisson_s},
        "serial_latency_p50_s": q(sorted(plain_lat), 0.5), "serial_latency_p99_s": q(sorted(plain_lat), 0.99),
    }


def _write_fastq(path: str, reads: np.ndarray, tag: bytes, mate: int) -> None:
    """Reads (n x READ_LEN ASCII) as FASTQ records "@<tag><i:09d>/<mate>", quality 'I'."""
    n = reads.shape[0]
    head = np.frombuffer(b"@" + tag, np.uint8)
    h = len(head)
    width = h + 9 + 3 + READ_LEN + 3 + READ_LEN + 1
    rec = np.empty((n, width), np.uint8)
    rec[:, :h] = head
    x = np.arange(n, dtype=np.int64)
    for d in range(8, -1, -1):
        rec[:, h + d] = 48 + (x % 10)
        x //= 10
    p = h + 9
    rec[:, p:p + 3] = np.frombuffer(b"/%d\n" % mate, np.uint8)
    p += 3
    rec[:, p:p + READ_LEN] = reads
    p += READ_LEN
    rec[:, p:p + 3] = np.frombuffer(b"\n+\n", np.uint8)
    p += 3
    rec[:, p:p + READ_LEN] = ord("I")
    rec[:, -1] = ord("\n")
    with open(path, "wb") as f:
        f.write(rec.tobytes())


def bench_cli(args) -> dict:
    """The per-sample normalization as the Perl driver runs it, from FASTQ
    files to normalized FASTQ files (util/insilico_read_normalization.pl
    --pairs_together): (a) tool by tool, one process per step, the script's
    own commands with this build's binaries; (b) bin/insilico_read_normalization,
    the whole job in one process; (c) a stream of samples through the resident
    service with concurrent clients against (b) run one sample after another.
    Every output pair is compared with (a)'s / (b)'s."""
    import shutil
    import subprocess
    import tempfile
    from concurrent.futures import ThreadPoolExecutor
    from trinityrnaseq_amd import synth
    bin_ = os.path.join(os.path.dirname(os.path.abspath(__file__)), "trinityrnaseq_amd", "bin")
    tool = lambda name: os.path.join(bin_, name)  # noqa: E731
    work = tempfile.mkdtemp(prefix="tk_cli_")
    env = dict(os.environ, LC_ALL="C")

    def sh(cmd: str, cwd: str) -> None:
        r = subprocess.run(["bash", "-c", cmd], cwd=cwd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
        if r.returncode:
            raise RuntimeError(f"{cmd} -> {r.returncode}: {r.stderr.decode()[-2000:]}")

    def sample(j: int, n_pairs: int) -> tuple[str, str]:
        tx = synth.large_transcriptome(3000 + j)
        buf = np.empty(2 * n_pairs * READ_LEN, np.uint8)
        synth.pe_reads(tx, n_pairs, READ_LEN, seed=3000 + j, out=buf)
        r = buf.reshape(2 * n_pairs, READ_LEN)
        lp, rp = os.path.join(work, f"s{j}_1.fq"), os.path.join(work, f"s{j}_2.fq")
        _write_fastq(lp, r[:n_pairs], b"s%d_" % j, 1)
        _write_fastq(rp, r[n_pairs:], b"s%d_" % j, 2)
        return lp, rp

    irn_opts = "--seqType fq --JM 20G --max_cov {c} --min_cov 1 --CPU 16 --max_CV 10000 --pairs_together " \
               "--PARALLEL_STATS".format(c=args.max_cov)
    try:
        t = time.perf_counter()
        L, R = sample(0, args.cli_pairs)
        log(f"cli: wrote {args.cli_pairs} pairs of FASTQ in {time.perf_counter() - t:.1f}s")
        # (a) tool by tool (the script's commands, :274-452)
        da = os.path.join(work, "chain")
        os.makedirs(da)
        steps = [
            ("seqtk-trinity", f"{tool('seqtk-trinity')} seq -A -R 1 {L} >> left.fa & "
                              f"{tool('seqtk-trinity')} seq -A -R 2 {R} >> right.fa & wait"),
            ("cat", "cat left.fa right.fa > both.fa"),
            ("jellyfish count", f"{tool('jellyfish')} count -t 16 -m 25 -s 100000000 --canonical both.fa"),
            ("jellyfish histo", f"{tool('jellyfish')} histo -t 16 -o jellyfish.K25.min2.kmers.fa.histo mer_counts.jf"),
            ("jellyfish dump", f"{tool('jellyfish')} dump -L 2 mer_counts.jf > jellyfish.K25.min2.kmers.fa"),
            ("fastaToKmerCoverageStats", " & ".join(
                f"{tool('fastaToKmerCoverageStats')} --reads {x}.fa --kmers jellyfish.K25.min2.kmers.fa "
                f"--kmer_size 25 --num_threads 8 --DS > {x}.fa.K25.stats" for x in ("left", "right")) + " & wait"),
            ("sort", " & ".join(
                f"(head -n1 {x}.fa.K25.stats > {x}.fa.K25.stats.sort && tail -n +2 {x}.fa.K25.stats | "
                f"sort --parallel=8 -k1,1 -T . -S 8G >> {x}.fa.K25.stats.sort)" for x in ("left", "right"))
             + " & wait"),
            ("nbkc_normalize", f"{tool('nbkc_normalize')} --left left.fa.K25.stats.sort --right "
                               f"right.fa.K25.stats.sort --sorted --max_cov {args.max_cov} --min_cov 1 "
                               f"--max_CV 10000 --merged_out pairs.K25.stats > pairs.accs"),
            ("extract", " & ".join(
                f"{tool('extract_normalized_reads')} --seqType fq --selected pairs.accs --output {x}.norm.fq {f}"
                for x, f in (("left", L), ("right", R))) + " & wait"),
        ]
        chain = {}
        t0 = time.perf_counter()
        for name, cmd in steps:
            t = time.perf_counter()
            sh(cmd, da)
            chain[name] = round(time.perf_counter() - t, 3)
            log(f"cli chain: {name} {chain[name]:.2f}s")
        chain_s = time.perf_counter() - t0
        # (b) one process
        db = os.path.join(work, "one")
        t = time.perf_counter()
        sh(f"{tool('insilico_read_normalization')} {irn_opts} --left {L} --right {R} --output {db}", work)
        one_s = time.perf_counter() - t
        log(f"cli one process: {one_s:.2f}s (chain {chain_s:.2f}s)")
        tag = f"normalized_K25_maxC{args.max_cov}_minC1_maxCV10000.fq"
        same = all(open(os.path.join(da, f"{x}.norm.fq"), "rb").read() ==
                   open(os.path.join(db, f"s0_{m}.fq.{tag}"), "rb").read() for x, m in (("left", 1), ("right", 2)))
        if not same:
            raise RuntimeError("the one-process outputs differ from the tool-by-tool chain's")
        n_sel = sum(1 for _ in open(os.path.join(da, "pairs.accs")))
        for f in (L, R):
            os.remove(f)
        shutil.rmtree(da)
        shutil.rmtree(db)
        # (c) a stream of samples: one process per sample, one after another, vs the service
        S, P = args.cli_samples, args.cli_stream_pairs
        files = []
        t = time.perf_counter()
        for j in range(S):
            files.append(sample(1 + j, P))
        log(f"cli: wrote {S} samples of {P} pairs in {time.perf_counter() - t:.1f}s")
        lat_seq = []
        t0 = time.perf_counter()
        for j, (l_, r_) in enumerate(files):
            t = time.perf_counter()
            sh(f"{tool('insilico_read_normalization')} {irn_opts} --left {l_} --right {r_} "
               f"--output {work}/seq{j}", work)
            lat_seq.append(time.perf_counter() - t)
        seq_s = time.perf_counter() - t0
        log(f"cli stream, one process per sample: {seq_s:.2f}s")
        sock = os.path.join(work, "norm.sock")
        srv = subprocess.Popen([tool("insilico_read_normalization"), "--serve", sock, "--slots", str(args.slots),
                                "--hbm_budget_gb", str(args.hbm_budget_gb)], stdout=subprocess.PIPE,
                               stderr=subprocess.PIPE)
        try:
            t = time.perf_counter()
            while not os.path.exists(sock):
                if srv.poll() is not None or time.perf_counter() - t > 120:
                    raise RuntimeError("the normalization service did not start")
                time.sleep(0.1)
            start_s = time.perf_counter() - t

            def client(j):
                t1 = time.perf_counter()
                l_, r_ = files[j]
                sh(f"{tool('insilico_read_normalization')} --server {sock} {irn_opts} --left {l_} --right {r_} "
                   f"--output {work}/srv{j}", work)
                return time.perf_counter() - t1

            t0 = time.perf_counter()
            with ThreadPoolExecutor(args.cli_clients) as ex:
                lat_srv = list(ex.map(client, range(S)))
            srv_s = time.perf_counter() - t0
            sh(f"{tool('insilico_read_normalization')} --server {sock} --shutdown", work)
            srv.wait(timeout=120)
        finally:
            if srv.poll() is None:
                srv.kill()
                srv.wait()
        log(f"cli stream through the service: {srv_s:.2f}s")
        for j in range(S):
            for m in (1, 2):
                a_ = open(os.path.join(work, f"seq{j}", f"s{1 + j}_{m}.fq.{tag}"), "rb").read()
                b_ = open(os.path.join(work, f"srv{j}", f"s{1 + j}_{m}.fq.{tag}"), "rb").read()
                if a_ != b_:
                    raise RuntimeError(f"sample {j}: the service's output differs from the one-process run's")
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return {
        "metric": METRIC + " [CLI: FASTQ files to normalized FASTQ]", "value": 2 * args.cli_pairs / one_s,
        "unit": "reads/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(one_s * 1000, 3),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic FASTQ files written to local disk before each timed run",
        "config": {"workload": f"insilico_read_normalization --pairs_together --max_cov {args.max_cov} on "
                               f"{args.cli_pairs} 2x150 pairs (one sample); stream: {S} samples of {P} pairs",
                   "slots": args.slots, "clients": args.cli_clients},
        "one_process_s": round(one_s, 3), "chain_s": round(chain_s, 3), "chain_steps_s": chain,
        "one_process_over_chain": chain_s / one_s, "selected_pairs": n_sel,
        "outputs_match": "one-process FASTQ == tool-by-tool FASTQ; service FASTQ == one-process FASTQ",
        "stream": {"samples": S, "pairs_each": P, "one_process_each_s": round(seq_s, 3),
                   "service_s": round(srv_s, 3), "service_over_one_process_each": seq_s / srv_s,
                   "service_start_s": round(start_s, 3),
                   "one_process_latency": percentiles(lat_seq), "service_latency": percentiles(lat_srv)},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=20_000_000)
    ap.add_argument("--max-cov", type=int, default=50)
    ap.add_argument("--cpu-pairs", type=int, default=2_000_000,
                    help="pairs of the CPU baseline's sample (median of 3 runs on all host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ceiling", action="store_true",
                    help="large_pe: skip the random-access ceiling microbenchmarks after the timed steps")
    ap.add_argument("--no-dump-check", action="store_true",
                    help="stress: skip the full sorted dump (dump -L 1) after the timed steps")
    ap.add_argument("--ceiling", action="store_true",
                    help="stress: run the random-access ceiling microbenchmarks over a buffer the size of "
                         "the count table after the timed steps")
    ap.add_argument("--copy-ref-gb", type=float, default=0.0,
                    help="also run one device-to-device copy of this many GB (HBM counter reference)")
    ap.add_argument("--workload", choices=["stress", "large_pe", "components", "serving", "inchworm", "cli", "kset"],
                    default="stress")
    ap.add_argument("--host-steps", type=int, default=2,
                    help="stress: jobs run from pageable host memory after the timed steps")
    ap.add_argument("--stress-pairs", type=float, default=150e6)
    ap.add_argument("--components", type=int, default=10_000)
    ap.add_argument("--plain-components", type=int, default=1000)
    ap.add_argument("--samples", type=int, default=64)
    ap.add_argument("--serve-scale", type=float, default=1.0,
                    help="serving: sample sizes as a fraction of 0.5M-30M pairs (1.0 = the configured "
                         "stream, ~140 GB of host memory; 0.1 for a quick run)")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--serve-pinned", action="store_true",
                    help="serving: hold the samples in page-locked host memory")
    ap.add_argument("--hbm-budget-gb", type=float, default=96.0)
    ap.add_argument("--cli-pairs", type=int, default=2_000_000, help="cli: pairs of the one-sample comparison")
    ap.add_argument("--kset-pairs", type=int, default=10_000_000, help="kset: pairs whose 48-mer windows are counted")
    ap.add_argument("--iworm-pairs", type=int, default=0,
                    help="inchworm: also assemble the whole dictionary of this many pairs (0: skip)")
    ap.add_argument("--cli-samples", type=int, default=8, help="cli: samples of the stream")
    ap.add_argument("--cli-stream-pairs", type=int, default=1_000_000, help="cli: pairs per stream sample")
    ap.add_argument("--cli-clients", type=int, default=4, help="cli: concurrent clients of the service")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    if args.workload == "components":
        print(json.dumps(bench_components(args)), flush=True)
        return 0
    if args.workload == "serving":
        print(json.dumps(bench_serving(args)), flush=True)
        return 0
    if args.workload == "stress":
        print(json.dumps(bench_stress(args)), flush=True)
        return 0
    if args.workload == "inchworm":
        print(json.dumps(bench_inchworm(args)), flush=True)
        return 0
    if args.workload == "cli":
        print(json.dumps(bench_cli(args)), flush=True)
        return 0
    if args.workload == "kset":
        print(json.dumps(bench_kset(args)), flush=True)
        return 0

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

    exact = {"want": True}

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
        if exact["want"]:  # warm-up only, never inside the timed steps
            exact["distinct"] = table.dump_size()
        table.free()
        reads.free()
        return nsel, est, info

    # one sample after another in one process: the blocks of one job's table and read set go to
    # the next job instead of hipFree + hipMalloc (a large hipMalloc after large frees can stall
    # for seconds on a device that earlier processes used)
    api.thread_cache(64 << 30)
    for i in range(args.warmup):
        t = time.perf_counter()
        nsel, est, info = job()
        log(f"warmup {i}: {time.perf_counter() - t:.3f}s selected={nsel} est_distinct={est:.3e} "
            f"capacity={info['capacity']}")
    exact["want"] = False
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
    for name in ["pack", "read_words", "hll", "sort_pairs", "minimizer", "count", "count_merge", "count_hist",
                 "count_scatter", "count_bounds", "count_regions", "count_retry", "part_hist", "part_scatter", "part_count",
                 "add_keys", "compact", "dump_count", "ovf_fold", "stats", "select", "select_flags",
                 "scan_u64", "scan_u32", "windows", "rehash"]:
        tot, n = ctx.kernel_time(name)
        if n:
            kern[name] = {"total_ms": round(tot, 3), "launches": n, "avg_ms": round(tot / n, 4)}
    words = ctx.count_words()  # merged words of the last step's count (0: table atomics)
    count_bytes = wins_per_step * COUNT_B_PER_WIN + n_reads * COUNT_B_PER_READ
    stats_bytes = wins_per_step * STATS_B_PER_WIN + n_reads * STATS_B_PER_READ
    table_bytes = info["capacity"] * 8
    stage_bytes = [("count", count_bytes), ("part_count", count_bytes), ("stats", stats_bytes)]
    if words:
        stage_bytes += [("count_merge", n_reads * READ_LEN * 0.375 + 8 * words),
                        ("count_hist", 2 * 8 * words), ("count_scatter", 2 * 16 * words),
                        ("count_regions", 8 * words + table_bytes)]
    roof = {}
    for name, by in stage_bytes:
        if name in kern:
            sec = kern[name]["total_ms"] / 1000.0 / args.steps
            roof[name] = {"achieved_gbs": by / sec / 1e9, "bytes_per_step": by,
                          "kernel_ms_per_step": sec * 1000}
    dom = max(roof, key=lambda n: roof[n]["kernel_ms_per_step"]) if roof else None
    roofline = None
    if dom:
        a = roof[dom]["achieved_gbs"]
        launches = kern[dom]["launches"] / args.steps
        traffic = (TRAFFIC_B_PER_WIN[dom] * wins_per_step / launches
                   if dom in TRAFFIC_B_PER_WIN and n_pairs == 20_000_000 else None)
        roofline = {"bound": "hbm", "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(a / HBM_PEAK_GBS, 4),
                    "traffic": round(traffic) if traffic else None,
                    "traffic_source": "profiles/large_pe_20M_pairs_count_stats.md (FETCH_SIZE + WRITE_SIZE)",
                    "algorithmic_bytes_per_launch": round(roof[dom]["bytes_per_step"] / launches),
                    "kernel": dom,
                    "other": {n: {"gbs": round(v["achieved_gbs"], 1), "ms": round(v["kernel_ms_per_step"], 3)}
                              for n, v in roof.items()},
                    "merged_words": words}

    # the same job with the reads still in pageable host memory (upload included): not `value`
    t = time.perf_counter()
    reads = api.Reads(ctx, host, off)
    est_h = api.estimate_distinct(ctx, reads, K, True)
    table = api.Table(ctx, K, True, int(est_h * 1.05) + 4096, 0.5)
    table.count(reads)
    api.read_cov_stats_d(ctx, table, reads, d_stats.ptr, min_count=2)
    nsel_h = api.select_pairs_d(ctx, d_stats.ptr, d_stats.ptr + n_pairs * 32, n_pairs,
                                args.max_cov, 1, 10000, READ_LEN - K + 1, d_sel.ptr)
    ctx.sync()
    pcie_s = time.perf_counter() - t
    table.free()
    reads.free()
    del host
    if nsel_h != nsel:
        raise RuntimeError("the host-input job selected different pairs")
    pcie = {"ms": round(pcie_s * 1000, 3), "k-mers/s": wins_per_step / pcie_s,
            "input": "pageable host memory (staged through pinned buffers)"}

    if roofline is not None and not args.no_ceiling:
        roofline["ceiling"] = random_access_ceiling(ctx, info["capacity"] * 8, wins_per_step, kern, args.steps)

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
                   "pairs": n_pairs, "k": K, "read_len": READ_LEN, "form": "plain per-job",
                   "device_blocks": "reused from job to job (tk_thread_cache)"},
        "reads_per_s": n_reads / (ms / 1000.0),
        "latency_s": {"p50": p(0.5), "p99": p(0.99), "min": lat_sorted[0], "max": lat_sorted[-1]},
        "kernels": kern, "roofline": roofline, "cpu_baseline": cpu, "host_input_job": pcie,
        "selected_pairs": nsel, "distinct_estimate": est, "distinct_exact": exact.get("distinct"),
    }
    print(json.dumps(out), flush=True)
    for b in (d_seq, d_off, d_stats, d_sel):
        b.free()
    api.thread_cache(0)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
