"""The large workloads' production paths at their real shapes (GPU).

* The 20M-pair single-sample job (2x150 PE from 50k transcripts) with the
  default switches: region-resident count over 5.04 G windows, 64 KB table
  regions, locality order, chunked dump.  Checked against the OpenMP CPU
  restatement on the host cores: distinct keys, a digest of every
  (key, count), the histogram, Sigma counts = N-free windows, and every read's
  (median, nwin, sum).
* A slice of the stress sample (expression lognormal sigma 2.5, 1.5% errors)
  through count -> coverage stats -> Inchworm prune (min 2, entropy 1.5) ->
  ReadsToTranscripts, on one table, each step against the oracle.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np
import pytest

from data import concat
from oracle import oracle as O

pytestmark = pytest.mark.gpu

K = 25
READ_LEN = 150


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def api():
    from trinityrnaseq_amd import api as A
    return A


def _host_threads() -> int:
    return max(1, min(16, len(os.sched_getaffinity(0))))


def test_large_pe_20M_pairs_full_size(api, ctx):
    from trinityrnaseq_amd import synth
    n_pairs = int(os.environ.get("TK_TEST_LARGE_PAIRS", "20000000"))
    n_reads = 2 * n_pairs
    say("large_pe_full_size pairs", n_pairs)
    t0 = time.time()
    tx = synth.large_transcriptome(1002)
    host = np.empty(n_reads * READ_LEN, np.uint8)
    synth.pe_reads(tx, n_pairs, READ_LEN, seed=1002, out=host)
    off = np.arange(n_reads + 1, dtype=np.uint64) * np.uint64(READ_LEN)
    say(f"generated in {time.time() - t0:.1f}s")
    d_seq = api.DeviceBuffer(ctx, host.nbytes + 64)
    d_off = api.DeviceBuffer(ctx, off.nbytes)
    d_seq.upload(host)
    d_off.upload(off)
    reads = api.Reads.from_device(ctx, d_seq.ptr, d_off.ptr, n_reads)
    est = api.estimate_distinct(ctx, reads, K, True)
    table = api.Table(ctx, K, True, int(est * 1.05) + 4096, 0.5)
    table.count(reads)
    words = ctx.count_words()
    assert words > 0, "the 5.04 G-window job counts through the region-resident passes"
    st = api.read_cov_stats(ctx, table, reads, min_count=2)
    hist = table.histo()
    keys, counts = table.dump()
    say(f"GPU done in {time.time() - t0:.1f}s: {len(keys)} distinct, {words} merged words")
    assert np.all(keys[1:] > keys[:-1])
    dg = O.table_digest(keys, counts)
    csum = int(counts.sum(dtype=np.uint64))
    del keys, counts
    table.free()
    reads.free()
    d_seq.free()
    d_off.free()
    total, valid = O.window_totals(host, off, K)
    assert total == n_reads * (READ_LEN - K + 1)
    assert csum == valid
    cr = O.cpuref_count_stats(host, off, K, True, 2, _host_threads(), expected_distinct=int(est * 1.05),
                              summary=True)
    say(f"CPU restatement done in {time.time() - t0:.1f}s ({_host_threads()} threads)")
    assert cr["distinct"] == int(hist.sum())
    assert cr["digest"] == dg
    assert np.array_equal(cr["histo"], hist.astype(np.uint64))
    assert abs(est / cr["distinct"] - 1) < 0.02
    assert np.array_equal(st["median"], cr["median"])
    assert np.array_equal(st["nwin"], cr["nwin"])
    assert np.array_equal(st["sum"], cr["sum"])


def test_stress_slice_chain(api, ctx):
    """The stress sample's generator (150k-transcript model cut to 3,000
    transcripts, expression sigma 2.5, 1.5% errors), 150k pairs: one table
    through count, stats (min_count 2), prune (2, entropy 1.5) and the vote
    against contig pieces (three overlapping thirds of each transcript)."""
    from trinityrnaseq_amd import synth
    say("stress_slice_chain")
    tx = synth.Transcriptome(1005, 3000, 1, 1200.0, 0.6, 200, 15000, share_frac=0.10, expr_sigma=2.5)
    n_pairs = 150_000
    host = np.empty(2 * n_pairs * READ_LEN, np.uint8)
    synth.pe_reads(tx, n_pairs, READ_LEN, seed=1005, sub_rate=0.015, out=host)
    off = np.arange(2 * n_pairs + 1, dtype=np.uint64) * np.uint64(READ_LEN)
    seq = host.tobytes()
    reads = api.Reads(ctx, seq, off)
    est = api.estimate_distinct(ctx, reads, K, True)
    table = api.Table(ctx, K, True, int(est * 1.05) + 4096, 0.5)
    table.count(reads)
    ok, oc = O.count(seq, off, K, True)
    gk, gc = table.dump()
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
    # heavy skew: the top 1% of k-mers hold a large share of the windows
    top = np.sort(oc)[::-1][: len(oc) // 100]
    assert top.sum() > 0.2 * oc.sum()
    st = api.read_cov_stats(ctx, table, reads, min_count=2)
    want = O.read_stats(seq, off, ok, oc, K, True, 2, 1, 0, 0)
    assert np.array_equal(st, want.astype(st.dtype))
    kept = table.prune(2, api.entropy_keep_table(K, 1.5))
    pk, pc = O.prune(ok, oc, K, 2, 1.5)
    assert kept == len(pk)
    gk, gc = table.dump()
    assert np.array_equal(gk, pk) and np.array_equal(gc, pc)
    # stats against the pruned dictionary: pruned k-mers are missing (1)
    st2 = api.read_cov_stats(ctx, table, reads, min_count=1)
    want2 = O.read_stats(seq, off, pk, pc, K, True, 1, 1, 0, 0)
    assert np.array_equal(st2, want2.astype(st2.dtype))
    pieces, comps = [], []
    for i in range(tx.n):
        sq = tx.seq(i)
        L = len(sq)
        for j in range(3):
            a, b = j * L // 3, min(L, (j + 2) * L // 3)
            if b - a >= K:
                pieces.append(sq[a:b])
                comps.append(i)
    comps = np.array(comps, np.uint32)
    cseq, coff = concat(pieces)
    ct = api.CTable(ctx, api.Reads(ctx, cseq, coff), comps, K, True)
    got = ct.vote(reads)
    bk, bc = O.r2t_build(cseq, coff, comps, K, True)
    want = O.r2t_vote(seq, off, bk, bc, K, True)
    assert np.array_equal(got, want.astype(got.dtype))
    ct.free()
    table.free()
    reads.free()
