"""Parity of the HIP path (through the C ABI) with the oracle.  Bit-exact."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import brute
from data import concat, edge_reads, random_reads, repetitive_reads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def api():
    from trinityrnaseq_amd import api as A
    return A


@pytest.fixture(scope="module")
def small_se():
    from trinityrnaseq_amd import synth
    tx = synth.small_transcriptome(1001)
    seq, off = synth.se_reads(tx, 200_000, 76, seed=1001)
    return seq.tobytes(), off


def _count(api, ctx, seqs_or_pair, k=25, canonical=True, expected=None, max_load=0.5):
    if isinstance(seqs_or_pair, tuple):
        seq, off = seqs_or_pair
    else:
        seq, off = concat(seqs_or_pair)
    r = api.Reads(ctx, seq, off)
    t = api.Table(ctx, k, canonical, expected or max(1024, len(seq)), max_load)
    t.count(r)
    return r, t, seq, off


@pytest.mark.parametrize("canonical", [True, False])
def test_count_dump_small_se(api, ctx, small_se, canonical):
    say("count_dump_small_se canonical", canonical)
    r, t, seq, off = _count(api, ctx, small_se, 25, canonical)
    gk, gc = t.dump()
    ok, oc = O.count(seq, off, 25, canonical)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gc, oc)
    assert np.array_equal(t.histo(), O.histo(oc))
    # dump -L 2 (the normalization dump)
    k2, c2 = t.dump(2)
    m = oc >= 2
    assert np.array_equal(k2, ok[m]) and np.array_equal(c2, oc[m])


@pytest.mark.parametrize("k", [1, 5, 13, 25, 31])
def test_count_edge_cases_all_k(api, ctx, k):
    say("count_edge_cases k", k)
    seqs = random_reads(100 + k, 500, 0, 200, n_rate=0.03) + edge_reads() + repetitive_reads(k, 300)
    for canonical in (True, False):
        r, t, seq, off = _count(api, ctx, seqs, k, canonical, expected=64)
        gk, gc = t.dump()
        want = brute.count(seqs, k, canonical)
        assert {brute.kmer_of(int(a), k): int(b) for a, b in zip(gk, gc)} == dict(want)


def test_empty_and_short_reads(api, ctx):
    say("empty_and_short_reads")
    for seqs in ([""], ["A" * 24], ["N" * 100, ""], []):
        r, t, seq, off = _count(api, ctx, seqs, 25, True, expected=64)
        gk, gc = t.dump()
        assert len(gk) == 0
        st = api.read_cov_stats(ctx, t, r)
        assert len(st) == len(seqs) and (len(st) == 0 or int(st["nwin"].max()) <= 100)


def test_count_overflow_high_multiplicity(api, ctx):
    """Counts far past the 13-bit slot field (k=25) fold through the side table."""
    say("count_overflow")
    base = "ACGTTGCAAGGCTTACCGATAGCTAGCATCGACTAGCATCAG"
    seqs = [base] * 20000 + [brute.rc(base)] * 5000 + ["GATTACAGATTACAGATTACAGATTACA"] * 9000
    r, t, seq, off = _count(api, ctx, seqs, 25, True, expected=64)
    gk, gc = t.dump()
    ok, oc = O.count(seq, off, 25, True)
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
    assert int(gc.max()) == 25000
    # a second count call accumulates on top of finalised counts
    t.count(r)
    gk2, gc2 = t.dump()
    assert np.array_equal(gk2, ok) and np.array_equal(gc2, 2 * oc)
    st = api.read_cov_stats(ctx, t, r)
    assert int(st["median"][0]) == 50000


def test_table_growth(api, ctx):
    say("table_growth")
    seqs = random_reads(9, 20000, 100, 150, n_rate=0.0)
    r, t, seq, off = _count(api, ctx, seqs, 25, True, expected=64)
    cap = t.info()["capacity"]
    ok, oc = O.count(seq, off, 25, True)
    assert cap * 0.5 >= len(ok)
    gk, gc = t.dump()
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)


def test_hll_estimate(api, ctx, small_se):
    say("hll_estimate")
    seq, off = small_se
    r = api.Reads(ctx, seq, off)
    est = api.estimate_distinct(ctx, r, 25, True)
    ok, _ = O.count(seq, off, 25, True)
    assert abs(est - len(ok)) / len(ok) < 0.08


@pytest.mark.parametrize("median_policy", [0, 1, 2])
@pytest.mark.parametrize("nwin_policy", [0, 1])
def test_read_stats_parity(api, ctx, small_se, median_policy, nwin_policy):
    say("read_stats", median_policy, nwin_policy)
    seq0, off0 = small_se
    extra = edge_reads() + random_reads(3, 300, 0, 300, n_rate=0.05)
    eseq, eoff = concat(extra)
    seq = seq0 + eseq
    off = np.concatenate([off0, off0[-1] + eoff[1:]])
    r, t, _, _ = _count(api, ctx, (seq, off))
    ok, oc = O.count(seq, off, 25, True)
    for min_count in (1, 2):
        got = api.read_cov_stats(ctx, t, r, min_count, 1, nwin_policy, median_policy)
        want = O.read_stats(seq, off, ok, oc, 25, True, min_count, 1, nwin_policy, median_policy)
        assert np.array_equal(got, want.astype(got.dtype))


def test_read_stats_long_reads(api, ctx):
    say("read_stats_long_reads")
    seqs = random_reads(77, 40, 1000, 9000, n_rate=0.001) + repetitive_reads(78, 200)
    seq, off = concat(seqs)
    r, t, _, _ = _count(api, ctx, (seq, off))
    ok, oc = O.count(seq, off, 25, True)
    got = api.read_cov_stats(ctx, t, r)
    want = O.read_stats(seq, off, ok, oc, 25, True, 1, 1, 0, 0)
    assert np.array_equal(got, want.astype(got.dtype))


def test_load_prune_parity(api, ctx, small_se):
    say("load_prune")
    seq, off = small_se
    ok, oc = O.count(seq, off, 25, True)
    for mc, me in [(1, 0.0), (2, 1.0), (2, 1.5)]:
        t = api.Table(ctx, 25, True, len(ok))
        t.load(ok, oc)
        kept = t.prune(mc, O.entropy_keep_table(25, me))
        pk, pc = O.prune(ok, oc, 25, mc, me)
        assert kept == len(pk)
        gk, gc = t.dump()
        assert np.array_equal(gk, pk) and np.array_equal(gc, pc)
        # pruned keys read as absent
        got = t.lookup(ok[:5000])
        exp = np.zeros(5000, np.uint32)
        idx = np.searchsorted(pk, ok[:5000])
        hit = (idx < len(pk)) & (pk[np.minimum(idx, len(pk) - 1)] == ok[:5000])
        exp[hit] = pc[idx[hit]]
        assert np.array_equal(got, exp)


def test_load_big_counts(api, ctx):
    say("load_big_counts")
    rng = np.random.default_rng(1)
    keys = np.unique(rng.integers(0, 1 << 50, 100000, dtype=np.uint64))
    counts = rng.integers(1, 1 << 31, len(keys), dtype=np.uint64).astype(np.uint32)
    counts[:10] = [8191, 8192, 8193, 16384, 1, 2, 0xFFFFFFFF, 0x80000000, 12345678, 8190]
    t = api.Table(ctx, 25, True, 64)
    t.load(keys, counts)
    gk, gc = t.dump()
    assert np.array_equal(gk, keys) and np.array_equal(gc, counts)
    # loading the same pairs again doubles (saturating at 2^32-1 in the dump)
    t.load(keys, counts)
    gk, gc = t.dump()
    want = np.minimum(counts.astype(np.uint64) * 2, 0xFFFFFFFF).astype(np.uint32)
    assert np.array_equal(gc, want)


def test_select_pairs_parity(api, ctx):
    say("select_pairs")
    from trinityrnaseq_amd import synth
    tx = synth.large_transcriptome(1002, 2000)
    left, right, off = synth.pe_reads(tx, 60000, 150, seed=1002)
    lr = api.Reads(ctx, left.tobytes(), off)
    rr = api.Reads(ctx, right.tobytes(), off)
    t = api.Table(ctx, 25, True, 1 << 20)
    t.count(lr)
    t.count(rr)
    sl = api.read_cov_stats(ctx, t, lr, min_count=2)
    sr = api.read_cov_stats(ctx, t, rr, min_count=2)
    for max_cov in (5, 50, 200):
        sel, ns = api.select_pairs(ctx, sl, sr, max_cov, 1, 10000)
        row = lambda s: O.format_stats_row("x", s).split("\t")
        lf = [row(s) for s in sl]
        rf = [row(s) for s in sr]
        want = O.select_pairs([float(x[1]) for x in lf], [float(x[1]) for x in rf],
                              [float(x[2]) for x in lf], [float(x[2]) for x in rf],
                              [float(x[3]) for x in lf], [float(x[3]) for x in rf], max_cov, 1, 10000)
        assert np.array_equal(sel, want)
        assert ns == int(want.sum())


@pytest.mark.parametrize("canonical", [True, False])
def test_r2t_parity(api, ctx, canonical):
    say("r2t canonical", canonical)
    from trinityrnaseq_amd import synth
    tx = synth.large_transcriptome(77, 600)
    # bundles: components of 1-5 transcripts joined by X
    rng = np.random.default_rng(4)
    bundles, comps = [], []
    i = 0
    cid = 0
    while i < tx.n:
        m = int(rng.integers(1, 6))
        bundles.append("X".join(tx.seq(j) for j in range(i, min(i + m, tx.n))))
        comps.append(cid)
        cid += 1
        i += m
    bseq, boff = concat(bundles)
    seq, off = synth.se_reads(tx, 50000, 100, seed=5, sub_rate=0.01)
    seq = seq.tobytes()
    b = api.Reads(ctx, bseq, boff)
    ct = api.CTable(ctx, b, np.array(comps, np.uint32), 25, canonical)
    r = api.Reads(ctx, seq, off)
    got = ct.vote(r)
    keys, cv = O.r2t_build(bseq, boff, comps, 25, canonical)
    assert ct.info()[0] == len(keys)
    want = O.r2t_vote(seq, off, keys, cv, 25, canonical)
    assert np.array_equal(got, want.astype(got.dtype))
    # edge reads
    eseq, eoff = concat(edge_reads() + random_reads(8, 100, 0, 400))
    got = ct.vote(api.Reads(ctx, eseq, eoff))
    want = O.r2t_vote(eseq, eoff, keys, cv, 25, canonical)
    assert np.array_equal(got, want.astype(got.dtype))
