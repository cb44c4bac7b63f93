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


@pytest.fixture(params=["direct", "regions", "ordered", "resident"])
def layout(request, monkeypatch):
    """Count through the direct kernel, through the region-partitioned
    passes (1 MB regions so that small tables span many regions), through
    the block-merging kernel with table atomics taking reads in the locality
    order (forced on small inputs), or through the region-resident passes
    (4 KB regions, so that small tables span many regions and both sort
    passes have many bins)."""
    if request.param == "direct":
        monkeypatch.setenv("TK_PARTITION", "0")
    elif request.param == "regions":
        monkeypatch.setenv("TK_PARTITION", "1")
        monkeypatch.setenv("TK_REGION_MB", "1")
    else:
        monkeypatch.setenv("TK_PARTITION", "0")
        monkeypatch.setenv("TK_ORDER", "1")
        monkeypatch.setenv("TK_ORDER_MIN_READS", "0")
        if request.param == "ordered":
            monkeypatch.setenv("TK_RESIDENT", "0")
        else:
            monkeypatch.setenv("TK_RESIDENT", "1")
            monkeypatch.setenv("TK_REGION_KB", "4")
    return request.param


@pytest.mark.parametrize("canonical", [True, False])
def test_count_dump_small_se(api, ctx, small_se, canonical, layout):
    say("count_dump_small_se canonical", canonical, layout)
    r, t, seq, off = _count(api, ctx, small_se, 25, canonical)
    if layout == "regions":
        assert t.info()["capacity"] // 8 > (1 << 14)  # spans several regions
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
def test_count_edge_cases_all_k(api, ctx, k, layout):
    say("count_edge_cases k", k, layout)
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


def test_count_overflow_high_multiplicity(api, ctx, layout):
    """Counts far past the 13-bit slot field (k=25) fold through the side table."""
    say("count_overflow", layout)
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


def test_table_growth(api, ctx, layout):
    say("table_growth", layout)
    seqs = random_reads(9, 20000, 100, 150, n_rate=0.0)
    r, t, seq, off = _count(api, ctx, seqs, 25, True, expected=64)
    cap = t.info()["capacity"]
    ok, oc = O.count(seq, off, 25, True)
    assert cap * 0.5 >= len(ok)
    gk, gc = t.dump()
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)


@pytest.mark.parametrize("tile", ["2048", "4096", "8192", "16384"])
@pytest.mark.parametrize("region_kb", ["1", "64"])
def test_count_resident_crowded_regions(api, ctx, region_kb, tile, monkeypatch):
    """Region-resident count into a table grown at load 0.9: with 128-slot
    regions some regions receive more keys than they hold, and their words
    are added after the table grows; with 64 KB regions the table is one
    region or a few.  A second count adds onto the non-empty table.  Every
    scatter tile size of the sort passes."""
    say("count_resident_crowded", region_kb, "tile", tile)
    monkeypatch.setenv("TK_RES_TILE", tile)
    monkeypatch.setenv("TK_PARTITION", "0")
    monkeypatch.setenv("TK_ORDER", "1")
    monkeypatch.setenv("TK_ORDER_MIN_READS", "0")
    monkeypatch.setenv("TK_REGION_KB", region_kb)
    monkeypatch.setenv("TK_RESIDENT", "1")
    seqs = random_reads(17, 20000, 100, 150, n_rate=0.001) + repetitive_reads(25, 3000)
    r, t, seq, off = _count(api, ctx, seqs, 25, True, expected=64, max_load=0.9)
    ok, oc = O.count(seq, off, 25, True)
    # the merged words of the resident passes: at least one per distinct key,
    # at most one per counted window
    assert len(ok) <= ctx.count_words() <= int(oc.sum())
    gk, gc = t.dump()
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
    t.count(r)
    gk, gc = t.dump()
    assert np.array_equal(gk, ok) and np.array_equal(gc, 2 * oc)
    # table atomics stream nothing
    monkeypatch.setenv("TK_RESIDENT", "0")
    t.count(r)
    assert ctx.count_words() == 0
    gk, gc = t.dump()
    assert np.array_equal(gk, ok) and np.array_equal(gc, 3 * oc)


@pytest.mark.parametrize("mode", ["reads", "words"])
def test_count_resident_in_chunks(api, ctx, mode, monkeypatch):
    """Region-resident count in read chunks (the form of counts whose word
    streams do not fit beside the table): each chunk merges, sorts and adds
    its words onto the regions the earlier chunks wrote.  "reads": chunks of
    2,500 reads, every word in the streams; "words": streams of 200,000
    words, shorter than a chunk's words, so flushes past their end are added
    with table atomics; on the new table, not yet cleared, the first chunk's
    overflowing flushes are dropped and the chunk is merged again after the
    table is cleared (one merge launch more than region passes).  A second
    count adds onto the non-empty table."""
    say("count_resident_in_chunks", mode)
    monkeypatch.setenv("TK_PARTITION", "0")
    monkeypatch.setenv("TK_ORDER", "1")
    monkeypatch.setenv("TK_ORDER_MIN_READS", "0")
    monkeypatch.setenv("TK_RESIDENT", "1")
    monkeypatch.setenv("TK_REGION_KB", "4")
    if mode == "reads":
        monkeypatch.setenv("TK_RES_CHUNK_READS", "2500")
    else:
        monkeypatch.setenv("TK_RES_STREAM_WORDS", "200000")
    seqs = random_reads(17, 20000, 100, 150, n_rate=0.001) + repetitive_reads(25, 3000)
    ctx.timing(True)
    ctx.timing_reset()
    r, t, seq, off = _count(api, ctx, seqs, 25, True)
    ctx.sync()
    merges = ctx.kernel_time("count_merge")[1]
    regions = ctx.kernel_time("count_regions")[1]
    ctx.timing(False)
    assert regions > 1
    assert merges == regions + (1 if mode == "words" else 0), (merges, regions)
    ok, oc = O.count(seq, off, 25, True)
    assert ctx.count_words() > 0
    gk, gc = t.dump()
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
    t.count(r)
    gk, gc = t.dump()
    assert np.array_equal(gk, ok) and np.array_equal(gc, 2 * oc)


@pytest.mark.parametrize("wpr", ["1", "2"])
def test_count_resident_merge_windows_per_round(api, ctx, wpr, monkeypatch):
    """The region-resident merge with one or two windows per thread between
    its block barriers: homopolymer reads put a block's every window on one
    key (the merge's per-flush occurrence bound, and counts past the slot's
    count field), random ones fill its table to the load bound."""
    say("count_resident_merge wpr", wpr)
    monkeypatch.setenv("TK_MERGE_WPR", wpr)
    monkeypatch.setenv("TK_PARTITION", "0")
    monkeypatch.setenv("TK_ORDER", "1")
    monkeypatch.setenv("TK_ORDER_MIN_READS", "0")
    monkeypatch.setenv("TK_RESIDENT", "1")
    monkeypatch.setenv("TK_REGION_KB", "4")
    seqs = (random_reads(41, 30000, 25, 150, n_rate=0.002) + repetitive_reads(25, 6000)
            + ["A" * 150] * 3000 + ["ACG" * 50] * 500)
    r, t, seq, off = _count(api, ctx, seqs, 25, True)
    ok, oc = O.count(seq, off, 25, True)
    assert oc.max() > 8064
    gk, gc = t.dump()
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)


@pytest.mark.parametrize("resident", ["1", "0"])
def test_count_reused_blocks(api, ctx, resident, monkeypatch):
    """With the thread's block reuse on, each table takes the block of the
    table freed before it, slots and all: the region-resident passes write
    every region of such an uncleared table, table atomics clear it first.
    Three read sets in a row, each equal to the oracle's count."""
    say("count_reused_blocks resident", resident)
    monkeypatch.setenv("TK_PARTITION", "0")
    monkeypatch.setenv("TK_ORDER", "1")
    monkeypatch.setenv("TK_ORDER_MIN_READS", "0")
    monkeypatch.setenv("TK_RESIDENT", resident)
    monkeypatch.setenv("TK_REGION_KB", "4")
    sets = [random_reads(31, 8000, 100, 150, n_rate=0.001), random_reads(32, 5000, 60, 150, n_rate=0.01)]
    api.thread_cache(1 << 30)
    try:
        for seqs in (sets[0], sets[1], sets[0]):
            r, t, seq, off = _count(api, ctx, seqs, 25, True, expected=1 << 20)
            ok, oc = O.count(seq, off, 25, True)
            gk, gc = t.dump()
            assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
            t.free()
            r.free()
    finally:
        api.thread_cache(0)


def test_hll_estimate(api, ctx, small_se):
    say("hll_estimate")
    seq, off = small_se
    r = api.Reads(ctx, seq, off)
    est = api.estimate_distinct(ctx, r, 25, True)
    ok, _ = O.count(seq, off, 25, True)
    assert abs(est - len(ok)) / len(ok) < 0.08


def test_hll_estimate_sampled(api, ctx):
    """Above 64M windows the estimate samples 1/16 of the keys by hash; it
    stays within 4% of the exact distinct count (and the table built from it
    counts exactly)."""
    say("hll_estimate_sampled")
    from trinityrnaseq_amd import synth
    tx = synth.large_transcriptome(1003, 5000)
    seq, off = synth.se_reads(tx, 560_000, 150, seed=9, sub_rate=0.01)
    seq = seq.tobytes()
    r = api.Reads(ctx, seq, off)
    assert r.windows(25) >= (64 << 20)
    est = api.estimate_distinct(ctx, r, 25, True)
    ok, oc = O.count(seq, off, 25, True)
    assert abs(est - len(ok)) / len(ok) < 0.04, (est, len(ok))
    t = api.Table(ctx, 25, True, int(est * 1.05) + 4096, 0.5)
    t.count(r)
    gk, gc = t.dump()
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)


@pytest.mark.parametrize("median_policy", [0, 1, 2])
@pytest.mark.parametrize("nwin_policy", [0, 1])
def test_read_stats_parity(api, ctx, small_se, median_policy, nwin_policy, monkeypatch):
    say("read_stats", median_policy, nwin_policy)
    monkeypatch.setenv("TK_REGION_MB", "1")
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


def test_locality_order_changes_nothing(api, ctx, small_se, monkeypatch):
    """The table passes take large read sets in minimizer order; tables and
    per-read stats are identical to input order and to the oracle."""
    say("locality_order")
    seq, off = small_se
    ok, oc = O.count(seq, off, 25, True)
    want = O.read_stats(seq, off, ok, oc, 25, True, 2, 1, 0, 0)
    for flag in ("0", "stats", "1"):
        monkeypatch.setenv("TK_ORDER", flag)
        r, t, _, _ = _count(api, ctx, (seq, off))
        gk, gc = t.dump()
        assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
        got = api.read_cov_stats(ctx, t, r, 2, 1, 0, 0)
        assert np.array_equal(got, want.astype(got.dtype))
        t.free()
        r.free()


@pytest.mark.parametrize("grab,pipe", [("0", "1"), ("1", "1"), ("4", "1"), ("37", "1"), ("4", "0")])
def test_per_read_passes_ticket_modes(api, ctx, small_se, grab, pipe, monkeypatch):
    """The scoring and vote passes with each XCD's waves taking runs of
    `grab` reads from a per-XCD counter (0: a fixed stride per wave), the
    scoring of reads of at most 128 windows software-pipelined over a wave's
    reads or one read at a time, in locality and input order, short and
    long reads, every median / N-window policy, against the oracle."""
    say("ticket", grab, "pipe", pipe)
    monkeypatch.setenv("TK_READ_TICKET", grab)
    monkeypatch.setenv("TK_STATS_PIPE", pipe)
    monkeypatch.setenv("TK_VOTE_PIPE", pipe)
    seq0, off0 = small_se
    short = edge_reads() + random_reads(92, 300, 0, 152, n_rate=0.01) + ["ACGTTGCA" * 19]  # 152 bases: 128 windows
    long_ = random_reads(91, 30, 200, 3000, n_rate=0.002)
    for ordered, extra in ((True, short), (False, long_), (False, short)):
        monkeypatch.setenv("TK_ORDER_MIN_READS", "0" if ordered else "100000000")
        eseq, eoff = concat(extra)
        seq = seq0 + eseq
        off = np.concatenate([off0, off0[-1] + eoff[1:]])
        r, t, _, _ = _count(api, ctx, (seq, off))
        ok, oc = O.count(seq, off, 25, True)
        for median_policy, nwin_policy, min_count in ((0, 0, 2), (1, 1, 1), (2, 0, 3), (2, 1, 2)):
            got = api.read_cov_stats(ctx, t, r, min_count, 1, nwin_policy, median_policy)
            want = O.read_stats(seq, off, ok, oc, 25, True, min_count, 1, nwin_policy, median_policy)
            assert np.array_equal(got, want.astype(got.dtype)), (ordered, median_policy, nwin_policy, min_count)
        # vote against components made of the first reads
        n_c = 200
        comps = np.arange(n_c, dtype=np.uint32) // 3
        cseq = seq[: int(off[n_c])]
        ct = api.CTable(ctx, api.Reads(ctx, cseq, off[: n_c + 1]), comps, 25, True)
        keys, cv = O.r2t_build(cseq, off[: n_c + 1], comps, 25, True)
        gv = ct.vote(r)
        wv = O.r2t_vote(seq, off, keys, cv, 25, True)
        assert np.array_equal(gv, wv.astype(gv.dtype)), ordered
        t.free()
        r.free()


@pytest.mark.parametrize("k,canonical", [(31, True), (31, False), (13, True), (1, True), (20, False)])
def test_pipelined_scoring_k_and_strand(api, ctx, k, canonical, monkeypatch):
    """The software-pipelined scoring kernel (reads of at most 128 windows,
    their code words spread over lanes 0-5) for other k and strand-specific
    counts, reads of every length up to 127 + k bases (128 windows) with
    Ns, in locality order, against the oracle."""
    say("pipelined scoring k", k, "canonical", canonical)
    monkeypatch.setenv("TK_STATS_PIPE", "1")
    monkeypatch.setenv("TK_READ_TICKET", "4")
    monkeypatch.setenv("TK_ORDER_MIN_READS", "0")
    seqs = random_reads(300 + k, 3000, 0, 127 + k, n_rate=0.01) + ["ACGTTGCAT"[: (k % 9) + 1] * 40][:1]
    seqs = [x[: 127 + k] for x in seqs] + ["G" * (127 + k), "T" * (126 + k)]
    seq, off = concat(seqs)
    r, t, _, _ = _count(api, ctx, (seq, off), k, canonical)
    ok, oc = O.count(seq, off, k, canonical)
    for median_policy, nwin_policy, min_count in ((0, 0, 1), (1, 1, 2), (2, 0, 2)):
        got = api.read_cov_stats(ctx, t, r, min_count, 1, nwin_policy, median_policy)
        want = O.read_stats(seq, off, ok, oc, k, canonical, min_count, 1, nwin_policy, median_policy)
        assert np.array_equal(got, want.astype(got.dtype)), (median_policy, nwin_policy, min_count)


@pytest.mark.parametrize("cache", ["2048", "4096"])
def test_read_stats_cached_policies(api, ctx, small_se, cache, monkeypatch):
    """The ordered scoring pass with the block count cache (reads of at most
    128 windows), every median/N-window policy and min_count, against the
    oracle; N-rich and short reads mixed in."""
    say("read_stats_cached", cache)
    monkeypatch.setenv("TK_STATS_CACHE", cache)
    monkeypatch.setenv("TK_ORDER_MIN_READS", "0")
    seq0, off0 = small_se
    extra = edge_reads() + random_reads(5, 400, 0, 150, n_rate=0.05)
    eseq, eoff = concat(extra)
    seq = seq0 + eseq
    off = np.concatenate([off0, off0[-1] + eoff[1:]])
    r, t, _, _ = _count(api, ctx, (seq, off))
    ok, oc = O.count(seq, off, 25, True)
    for median_policy in (0, 1, 2):
        for nwin_policy in (0, 1):
            for min_count in (1, 2, 3):
                got = api.read_cov_stats(ctx, t, r, min_count, 1, nwin_policy, median_policy)
                want = O.read_stats(seq, off, ok, oc, 25, True, min_count, 1, nwin_policy, median_policy)
                assert np.array_equal(got, want.astype(got.dtype)), (median_policy, nwin_policy, min_count)


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


def test_select_pairs_edge_rules(api, ctx):
    """Pairs whose reads have no window are skipped without a rand() draw
    (merged mean "0.0"), and the draws after them keep Perl's sequence; stats
    whose mean is below 1 (missing k-mers counted as 0) are refused, and so is
    a drawn pair whose merged median is 0 (the script divides by it)."""
    say("select_pairs_edge_rules")
    rng = np.random.default_rng(8)
    n = 5000
    st = np.zeros(2 * n, api.READ_STAT_DTYPE)
    nwin = rng.integers(1, 127, 2 * n).astype(np.uint32)
    nwin[rng.random(2 * n) < 0.05] = 0
    both = rng.random(n) < 0.03
    nwin[:n][both] = 0
    nwin[n:][both] = 0
    med = rng.integers(1, 300, 2 * n).astype(np.uint32)
    med[nwin == 0] = 0
    st["nwin"] = nwin
    st["median"] = med
    st["sum"] = nwin.astype(np.uint64) * med.astype(np.uint64)
    st["sumsq_lo"] = st["sum"] * med.astype(np.uint64)
    sl, sr = st[:n], st[n:]
    for min_cov in (0, 1, 5):
        sel, ns = api.select_pairs(ctx, sl, sr, 50, min_cov, 10000)
        row = lambda s: O.format_stats_row("x", s).split("\t")
        lf = [row(s) for s in sl]
        rf = [row(s) for s in sr]
        want = O.select_pairs([float(x[1]) for x in lf], [float(x[1]) for x in rf],
                              [float(x[2]) for x in lf], [float(x[2]) for x in rf],
                              [float(x[3]) for x in lf], [float(x[3]) for x in rf], 50, min_cov, 10000)
        assert np.array_equal(sel, want), min_cov
        assert ns == int(want.sum())
        assert not sel[both].any()
    low = st.copy()
    low["sum"][0] = 0  # a read with windows and mean 0
    with pytest.raises(api.TkError, match="below 1"):
        api.select_pairs(ctx, low[:n], low[n:], 50, 0, 10000)
    zero = st.copy()
    zero["median"][0] = 0
    zero["median"][n] = 0
    with pytest.raises(api.TkError, match="median 0"):
        api.select_pairs(ctx, zero[:n], zero[n:], 50, 0, 10000)


@pytest.mark.parametrize("max_cov,min_cov", [(50, 0), (50, 1), (200, 0), (5, 3)])
def test_select_pairs_equals_reference_scripts(api, ctx, max_cov, min_cov):
    """tk_select_pairs on the integer aggregates behind the committed stats
    files selects exactly the pairs nbkc_normalize.pl selected (after
    nbkc_merge_left_right_stats.pl --sorted) on their text."""
    from data import nbkc_aggregates, nbkc_pairs, nbkc_selected
    say("select_pairs_equals_reference_scripts", max_cov, min_cov)
    la, ls = nbkc_aggregates("left", api.READ_STAT_DTYPE)
    ra, rs = nbkc_aggregates("right", api.READ_STAT_DTYPE)
    li = {a: i for i, a in enumerate(la)}
    ri = {a: i for i, a in enumerate(ra)}
    cores = [p[0] for p in nbkc_pairs()]
    sl = ls[[li[c + "/1"] for c in cores]]
    sr = rs[[ri[c + "/2"] for c in cores]]
    sel, ns = api.select_pairs(ctx, sl, sr, max_cov, min_cov, 10000)
    want = nbkc_selected("pairs", max_cov, min_cov, 10000)
    assert [c for c, s in zip(cores, sel) if s] == want
    assert ns == len(want)


@pytest.mark.parametrize("chunk", ["", "1000", "65536"])
def test_dump_in_key_range_chunks(api, ctx, small_se, chunk, monkeypatch):
    """A dump larger than one sort call takes goes out in key-range chunks,
    each sorted on the device; a small chunk limit forces many chunks.  The
    concatenation is the oracle's ascending dump, for -L 1 and -L 2."""
    say("dump_in_key_range_chunks", chunk or "default")
    if chunk:
        monkeypatch.setenv("TK_DUMP_CHUNK", chunk)
    seq, off = small_se
    r, t, _, _ = _count(api, ctx, (seq, off))
    ok, oc = O.count(seq, off, 25, True)
    gk, gc = t.dump()
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
    k2, c2 = t.dump(2)
    m = oc >= 2
    assert np.array_equal(k2, ok[m]) and np.array_equal(c2, oc[m])
    # short keys: the bins cover the whole 2k-bit key space
    seqs = random_reads(5, 3000, 20, 60, n_rate=0.01)
    r5, t5, s5, o5 = _count(api, ctx, seqs, 5, True, expected=64)
    gk, gc = t5.dump()
    ok5, oc5 = O.count(s5, o5, 5, True)
    assert np.array_equal(gk, ok5) and np.array_equal(gc, oc5)


@pytest.mark.parametrize("ordered", [False, True])
@pytest.mark.parametrize("canonical", [True, False])
def test_r2t_parity(api, ctx, canonical, ordered, monkeypatch):
    say("r2t canonical", canonical, "ordered", ordered)
    if ordered:  # reads voted in the locality order even though the set is small
        monkeypatch.setenv("TK_ORDER_MIN_READS", "0")
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


@pytest.mark.parametrize("canonical", [True, False])
def test_sample_reads_parity(api, ctx, canonical):
    """The reference's own sample reads (30,575 pairs of 76 nt, 1,190 + 627
    reads with an N): count, dump, per-read stats (min_count 1 and 2) and the
    vote against the 11 transcripts they were drawn from, each equal to the
    oracle."""
    from data import ref_seqs, sample_fasta
    say("sample_reads_parity canonical", canonical)
    seqs = sample_fasta("left")[1] + sample_fasta("right")[1]
    seq, off = concat(seqs)
    r, t, _, _ = _count(api, ctx, (seq, off), 25, canonical)
    gk, gc = t.dump()
    ok, oc = O.count(seq, off, 25, canonical)
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
    total, valid = O.window_totals(seq, off, 25)
    assert int(gc.sum()) == valid and total == 61150 * 52
    for mc in (1, 2):
        st = api.read_cov_stats(ctx, t, r, min_count=mc)
        want = O.read_stats(seq, off, ok, oc, 25, canonical, mc, 1, 0, 0)
        assert np.array_equal(st, want.astype(st.dtype)), mc
    refs = ref_seqs()
    bseq, boff = concat(refs)
    comps = np.arange(len(refs), dtype=np.uint32)
    ct = api.CTable(ctx, api.Reads(ctx, bseq, boff), comps, 25, canonical)
    got = ct.vote(r)
    keys, cv = O.r2t_build(bseq, boff, comps, 25, canonical)
    want = O.r2t_vote(seq, off, keys, cv, 25, canonical)
    assert np.array_equal(got, want.astype(got.dtype))
    if canonical:  # about half of the reads come from these 11 transcripts
        assert (got["hits"] > 0).mean() > 0.4


@pytest.mark.parametrize("shift", [0, 1, 3])
def test_reads_from_device_unaligned(api, ctx, shift):
    """Packing from device-resident ASCII at any byte alignment (and ending at
    the very end of the allocation) equals packing from host memory."""
    say("reads_from_device_unaligned", shift)
    seqs = random_reads(55 + shift, 3000, 0, 400, n_rate=0.02) + edge_reads()
    seq, off = concat(seqs)
    buf = api.DeviceBuffer(ctx, len(seq) + shift)
    buf.upload(np.frombuffer(seq, np.uint8), offset=shift)
    doff = api.DeviceBuffer(ctx, off.nbytes)
    doff.upload(off)
    r = api.Reads.from_device(ctx, buf.ptr + shift, doff.ptr, len(seqs))
    t = api.Table(ctx, 25, True, 1 << 16)
    t.count(r)
    gk, gc = t.dump()
    ok, oc = O.count(seq, off, 25, True)
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
    st = api.read_cov_stats(ctx, t, r)
    assert np.array_equal(st, O.read_stats(seq, off, ok, oc, 25, True, 1, 1, 0, 0).astype(st.dtype))


def test_upload_pageable_pinned_device(api, ctx):
    """Reads larger than the staging buffers (pageable source: staged in
    chunks through pinned buffers) pack exactly as from page-locked memory
    and from device memory."""
    say("upload_pageable_pinned_device")
    from trinityrnaseq_amd import synth
    tx = synth.small_transcriptome(1001)
    seq, off = synth.se_reads(tx, 330_001, 151, seed=77)  # ~50 MB: three and a bit 16 MB chunks
    ok, oc = O.count(seq.tobytes(), off, 25, True)
    pin = api.HostBuffer(seq.nbytes)
    pin.array[:] = seq
    dseq = api.DeviceBuffer(ctx, seq.nbytes)
    dseq.upload(seq)
    doff = api.DeviceBuffer(ctx, off.nbytes)
    doff.upload(off)
    for how in ("pageable", "pinned", "device"):
        if how == "pageable":
            r = api.Reads(ctx, seq, off)
        elif how == "pinned":
            r = api.Reads(ctx, pin.array, off)
        else:
            r = api.Reads.from_device(ctx, dseq.ptr, doff.ptr, len(off) - 1)
        t = api.Table(ctx, 25, True, len(ok) + 1024)
        t.count(r)
        gk, gc = t.dump()
        assert np.array_equal(gk, ok) and np.array_equal(gc, oc), how
        t.free()
        r.free()
    pin.free()


def test_pack_every_byte_value(api, ctx):
    """Every byte value 0..255 at every position of a 32-base word: only
    A/C/G/T/a/c/g/t are bases, anything else breaks the windows over it."""
    say("pack_every_byte_value")
    rng = np.random.default_rng(2024)
    seqs = []
    for v in range(256):
        for pos in (0, 7, 31, 32, 45):
            b = bytearray(rng.choice(np.frombuffer(b"ACGTacgt", np.uint8), 80).tobytes())
            b[pos] = v
            seqs.append(bytes(b))
    off = np.zeros(len(seqs) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    seq = b"".join(seqs)
    for k in (5, 25):
        r = api.Reads(ctx, seq, off)
        t = api.Table(ctx, k, True, 1 << 16)
        t.count(r)
        gk, gc = t.dump()
        ok, oc = O.count(seq, off, k, True)
        assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
        st = api.read_cov_stats(ctx, t, r)
        assert np.array_equal(st, O.read_stats(seq, off, ok, oc, k, True, 1, 1, 0, 0).astype(st.dtype))


def _components(seed, n_comp, max_reads=400):
    """Small Chrysalis-like components: transcripts, reads drawn from them
    (both strands, 1% errors), and contig-like pieces of the transcripts."""
    from trinityrnaseq_amd import synth
    rng = np.random.default_rng(seed)
    comps = []
    for c in range(n_comp):
        ntx = int(rng.integers(1, 4))
        tx = synth.Transcriptome(seed * 1000 + c, ntx, 0, 300, 1500)
        nreads = int(min(max_reads, 10 * (1 + rng.pareto(0.8))))
        s, o = synth.se_reads(tx, nreads, 150, seed=seed * 7 + c, sub_rate=0.01)
        reads = [s[int(o[i]):int(o[i + 1])].tobytes().decode() for i in range(nreads)]
        contigs = []
        for j in range(ntx):
            t = tx.seq(j)
            cut = int(rng.integers(100, max(101, len(t) - 100)))
            contigs += [t[:cut], t[cut - 30:]]
        comps.append((reads, contigs))
    comps[3] = ([], comps[3][1])          # a component without reads
    comps[5] = (comps[5][0], [])          # a component without contigs
    return comps


@pytest.mark.parametrize("agg,ct_lds", [("1", "1"), ("0", "0"), ("1", "0")])
@pytest.mark.parametrize("packed", ["1", "0"])
@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_components_match_per_component_oracle(api, ctx, use_graph, packed, agg, ct_lds, monkeypatch):
    """Packed: one-word component-table slots (the default at k = 25); else
    the two-word slots of the plain component table.  agg: the count merges
    each block's group of one component's reads in LDS (default at k <= 25);
    else one table atomic per window.  ct_lds: small components' tables are
    built in LDS and written out whole (default); else every contig window
    takes a memory-side atomic."""
    say("batch_components graph", use_graph, "packed", packed, "agg", agg, "ct_lds", ct_lds)
    monkeypatch.setenv("TK_BATCH_CT_PACKED", packed)
    monkeypatch.setenv("TK_BATCH_AGG", agg)
    monkeypatch.setenv("TK_BATCH_CT_LDS", ct_lds)
    comps = _components(11, 40)
    all_reads = [r for rs, _ in comps for r in rs]
    all_contigs = [c for _, cs in comps for c in cs]
    ro = np.cumsum([0] + [len(rs) for rs, _ in comps]).astype(np.uint64)
    co = np.cumsum([0] + [len(cs) for _, cs in comps]).astype(np.uint64)
    reads = api.Reads.from_list(ctx, all_reads)
    contigs = api.Reads.from_list(ctx, all_contigs)
    b = api.Batch(ctx, reads, ro, contigs, co, use_graph=use_graph)
    n = len(all_reads)
    d_st = api.DeviceBuffer(ctx, n * 32)
    d_vo = api.DeviceBuffer(ctx, n * 12)
    for rep in range(3):  # replays must agree with the first run
        b.run_d(d_st.ptr, d_vo.ptr, min_count=2)
        st = np.zeros(n, api.READ_STAT_DTYPE)
        vo = np.zeros(n, api.R2T_VOTE_DTYPE)
        d_st.download(st)
        d_vo.download(vo)
        for c, (rs, cs) in enumerate(comps):
            if not rs:
                continue
            seq, off = concat(rs)
            ok, oc = O.count(seq, off, 25, True)
            a, z = int(ro[c]), int(ro[c + 1])
            want = O.read_stats(seq, off, ok, oc, 25, True, 2, 1, 0, 0)
            assert np.array_equal(st[a:z], want.astype(st.dtype)), (rep, c)
            if cs:
                bseq, boff = concat(cs)
                keys, cv = O.r2t_build(bseq, boff, np.arange(len(cs)), 25, True)
                wv = O.r2t_vote(seq, off, keys, cv, 25, True)
            else:
                wv = np.zeros(len(rs), O.R2T_VOTE_DTYPE)
                wv["comp"] = 0xFFFFFFFF
                wv["nwin"] = [max(0, len(r) - 24) for r in rs]
            assert np.array_equal(vo[a:z], wv.astype(vo.dtype)), (rep, c)


@pytest.mark.parametrize("agg", ["1", "0"])
def test_batch_count_runs_past_the_slot_field(api, ctx, agg, monkeypatch):
    """Components whose reads repeat one k-mer thousands of times (counts past
    the 13-bit slot field go to the side table; with the block merge, each
    group's run is cut at the merge's occurrence bound) next to ordinary
    ones: every component's stats equal the oracle's."""
    say("batch_count_runs agg", agg)
    monkeypatch.setenv("TK_BATCH_AGG", agg)
    comps = _components(13, 6)
    comps.insert(2, (["A" * 150] * 2000, ["A" * 200]))
    comps.insert(4, (["ACG" * 50] * 700 + ["T" * 90] * 130, ["ACG" * 60]))
    all_reads = [r for rs, _ in comps for r in rs]
    all_contigs = [c for _, cs in comps for c in cs]
    ro = np.cumsum([0] + [len(rs) for rs, _ in comps]).astype(np.uint64)
    co = np.cumsum([0] + [len(cs) for _, cs in comps]).astype(np.uint64)
    reads = api.Reads.from_list(ctx, all_reads)
    contigs = api.Reads.from_list(ctx, all_contigs)
    b = api.Batch(ctx, reads, ro, contigs, co, use_graph=True)
    _check_batch(api, ctx, comps, b, ro)
    b.free()


def _check_batch(api, ctx, comps, b, ro):
    n = int(ro[-1])
    d_st = api.DeviceBuffer(ctx, max(n, 1) * 32)
    d_vo = api.DeviceBuffer(ctx, max(n, 1) * 12)
    b.run_d(d_st.ptr, d_vo.ptr, min_count=2)
    st = np.zeros(n, api.READ_STAT_DTYPE)
    vo = np.zeros(n, api.R2T_VOTE_DTYPE)
    d_st.download(st)
    d_vo.download(vo)
    for c, (rs, cs) in enumerate(comps):
        if not rs:
            continue
        seq, off = concat(rs)
        ok, oc = O.count(seq, off, 25, True)
        a, z = int(ro[c]), int(ro[c + 1])
        assert np.array_equal(st[a:z], O.read_stats(seq, off, ok, oc, 25, True, 2, 1, 0, 0).astype(st.dtype)), c
        if cs:
            bseq, boff = concat(cs)
            keys, cv = O.r2t_build(bseq, boff, np.arange(len(cs)), 25, True)
            assert np.array_equal(vo[a:z], O.r2t_vote(seq, off, keys, cv, 25, True).astype(vo.dtype)), c
    d_st.free()
    d_vo.free()


@pytest.mark.parametrize("window_load,waves", [(0.35, 2), (0.7, 4), (0.0, 0)])
def test_batch_layouts_autotune_and_large_component(api, ctx, window_load, waves):
    """Explicit table load and waves per component (both corners of the
    autotuner's grid), and the autotuned layout, on a batch that holds a
    component of 60,000 reads beside small ones: every component equal to
    the oracle run on it alone."""
    from trinityrnaseq_amd import synth
    say("batch layouts", window_load, waves)
    comps = _components(23, 12)
    tx = synth.Transcriptome(777, 8, 0, 1500, 4000)
    s, o = synth.se_reads(tx, 60000, 150, seed=778, sub_rate=0.01)
    big = [s[int(o[i]):int(o[i + 1])].tobytes().decode() for i in range(60000)]
    comps.insert(4, (big, [tx.seq(j) for j in range(tx.n)]))
    all_reads = [r for rs, _ in comps for r in rs]
    all_contigs = [c for _, cs in comps for c in cs]
    ro = np.cumsum([0] + [len(rs) for rs, _ in comps]).astype(np.uint64)
    co = np.cumsum([0] + [len(cs) for _, cs in comps]).astype(np.uint64)
    reads = api.Reads.from_list(ctx, all_reads)
    contigs = api.Reads.from_list(ctx, all_contigs)
    b = api.Batch(ctx, reads, ro, contigs, co, window_load=window_load, stats_waves=waves, use_graph=True)
    if window_load == 0.0:
        n = len(all_reads)
        d_st = api.DeviceBuffer(ctx, n * 32)
        d_vo = api.DeviceBuffer(ctx, n * 12)
        ms = b.autotune(d_st.ptr, d_vo.ptr, min_count=2)
        say("  autotuned", b.info(), f"{ms:.3f} ms")
        d_st.free()
        d_vo.free()
    _check_batch(api, ctx, comps, b, ro)
    b.free()


def _plain_job(api, ctx, seq, off, n_pairs, max_cov=50):
    r = api.Reads(ctx, seq.tobytes(), off)
    t = api.Table(ctx, 25, True, int(api.estimate_distinct(ctx, r, 25, True) * 1.05) + 4096)
    t.count(r)
    st = api.read_cov_stats(ctx, t, r, min_count=2)
    sel, ns = api.select_pairs(ctx, st[:n_pairs], st[n_pairs:], max_cov, 1, 10000)
    return sel, ns


def test_server_general_job_selections(api, ctx):
    """tk_server_submit_job: all reads counted and scored (strand-specific
    here, plus unpaired reads that count but are never selected), then a pair
    selection over units given as read indices in a shuffled order and a
    single-read selection; each equals tk_select_pairs on the stats of the
    same job run alone, gathered in the same order."""
    say("server_general_job")
    from trinityrnaseq_amd import synth
    tx = synth.large_transcriptome(3100, 800)
    n = 40000
    buf = np.empty(2 * n * 150, np.uint8)
    synth.pe_reads(tx, n, 150, seed=3100, out=buf)
    extra, _ = synth.se_reads(tx, 3000, 120, seed=3101)
    seq = np.concatenate([buf, extra])
    off = np.concatenate([np.arange(2 * n + 1, dtype=np.uint64) * np.uint64(150),
                          np.uint64(2 * n * 150) + np.arange(1, 3001, dtype=np.uint64) * np.uint64(120)])
    rng = np.random.default_rng(5)
    perm = rng.permutation(n - 100)  # 100 pairs never selected
    left, right = perm.astype(np.uint64), (perm + n).astype(np.uint64)
    singles = rng.permutation(2 * n + 3000)[:50000].astype(np.uint64)
    r = api.Reads(ctx, seq.tobytes(), off)
    t = api.Table(ctx, 25, False, int(api.estimate_distinct(ctx, r, 25, False) * 1.05) + 4096)
    t.count(r)
    st = api.read_cov_stats(ctx, t, r, min_count=2)
    want_p, nw_p = api.select_pairs(ctx, st[left], st[right], 30, 1, 10000)
    want_s, nw_s = api.select_pairs(ctx, st[singles], st[singles], 20, 1, 10000)
    t.free()
    r.free()
    srv = api.Server(slots=2, hbm_budget_bytes=1 << 30)
    try:
        a = srv.submit_job(seq, off, [(left, right)], canonical=False, max_cov=30)
        b = srv.submit_job(seq, off, [(singles, None), (left, right)], canonical=False, max_cov=20)
        res, (sel,) = srv.wait(a)
        assert res.n_selected == nw_p and np.array_equal(sel, want_p)
        res, (sel_s, sel_p2) = srv.wait(b)
        assert np.array_equal(sel_s, want_s)
        want_p2, _ = api.select_pairs(ctx, st[left], st[right], 20, 1, 10000)
        assert np.array_equal(sel_p2, want_p2)
        assert res.n_selected == nw_s + int(want_p2.sum())
        with pytest.raises(api.TkError):  # a unit index past the reads
            srv.submit_job(seq, off, [(np.array([len(off) - 1], np.uint64), None)])
    finally:
        srv.close()


def test_server_job_out_of_device_memory(api, ctx, monkeypatch):
    """A job whose device allocations fail (forced from the host: the
    library's bytes in use may not pass TK_DEVICE_ALLOC_LIMIT) reports
    TK_ERR_OOM through wait() and gives its admission back; a small job
    submitted beside it, and the large job submitted again once the limit is
    lifted, equal the same jobs run alone."""
    say("server_oom")
    from trinityrnaseq_amd import synth
    samples = []
    for j, n in enumerate([4000, 150000]):
        tx = synth.large_transcriptome(3100 + j, 800)
        buf = np.empty(2 * n * 150, np.uint8)
        synth.pe_reads(tx, n, 150, seed=3100 + j, out=buf)
        off = np.arange(2 * n + 1, dtype=np.uint64) * np.uint64(150)
        samples.append((buf, off, n))
    alone = [_plain_job(api, ctx, b, o, n) for b, o, n in samples]
    import gc
    gc.collect()  # the plain jobs' device blocks are freed before the limit is set
    srv = api.Server(slots=2, hbm_budget_bytes=8 << 30)
    try:
        # room for the small job (about 16 MB) but not for the large one's
        # 45 MB of reads, whatever the other holds when it asks
        monkeypatch.setenv("TK_DEVICE_ALLOC_LIMIT", str(api.device_bytes_in_use() + (40 << 20)))
        big = srv.submit(*samples[1])
        small = srv.submit(*samples[0])
        with pytest.raises(api.TkError) as err:
            srv.wait(big)
        assert err.value.code == -3  # TK_ERR_OOM
        res, sel = srv.wait(small)
        assert res.n_selected == alone[0][1] and np.array_equal(sel, alone[0][0])
        monkeypatch.delenv("TK_DEVICE_ALLOC_LIMIT")
        again = srv.submit(*samples[1])
        res, sel = srv.wait(again)
        assert res.n_selected == alone[1][1] and np.array_equal(sel, alone[1][0])
    finally:
        srv.close()


def test_server_results_equal_alone_and_in_the_mix(api, ctx):
    """Jobs admitted concurrently (tight budget forces queueing) give exactly
    the selections of the same jobs run alone."""
    say("server_mix")
    from trinityrnaseq_amd import synth
    samples = []
    for j, n in enumerate([20000, 60000, 5000, 90000, 30000, 45000]):
        tx = synth.large_transcriptome(3000 + j, 800)
        buf = np.empty(2 * n * 150, np.uint8)
        synth.pe_reads(tx, n, 150, seed=3000 + j, out=buf)
        off = np.arange(2 * n + 1, dtype=np.uint64) * np.uint64(150)
        samples.append((buf, off, n))
    alone = [_plain_job(api, ctx, b, o, n) for b, o, n in samples]
    srv = api.Server(slots=3, hbm_budget_bytes=96 << 20)
    try:
        ids = [srv.submit(b, o, n) for b, o, n in samples]
        for (sel_a, ns_a), jid in zip(alone, ids):
            res, sel = srv.wait(jid)
            assert res.n_selected == ns_a
            assert np.array_equal(sel, sel_a)
            assert res.latency_s >= res.run_s >= 0
    finally:
        srv.close()
    # the smallest sample also matches the oracle's Perl-semantics selection
    b, o, n = samples[2]
    seq = b.tobytes()
    ok, oc = O.count(seq, o, 25, True)
    st = O.read_stats(seq, o, ok, oc, 25, True, 2, 1, 0, 0)
    rows = [O.format_stats_row("x", s).split("\t") for s in st]
    want = O.select_pairs([float(x[1]) for x in rows[:n]], [float(x[1]) for x in rows[n:]],
                          [float(x[2]) for x in rows[:n]], [float(x[2]) for x in rows[n:]],
                          [float(x[3]) for x in rows[:n]], [float(x[3]) for x in rows[n:]], 50, 1, 10000)
    assert np.array_equal(alone[2][0], want)


def _iworm_check(api, ctx, t, k, ds, msc, keep=None, keep_thr=None):
    gk, gc = t.dump()
    contigs, sums, seeds, info = api.inchworm(ctx, t, msc, keep)
    want_c, want_s, want_seed = O.inchworm(gk, gc, k, ds, msc, keep)
    assert len(contigs) == len(want_c)
    assert contigs == want_c
    assert np.array_equal(sums, want_s) and np.array_equal(seeds, want_seed)
    return info


@pytest.mark.parametrize("canonical", [True, False])
def test_inchworm_parity_small_se(api, ctx, small_se, canonical):
    """Greedy contigs over the 200k-read single-end table (the small
    workload) equal the oracle's serial greedy, contig for contig, with seed
    coverage 1 and 2."""
    say("inchworm_parity_small_se canonical", canonical)
    r, t, _, _ = _count(api, ctx, small_se, 25, canonical)
    for msc in (2, 1):
        info = _iworm_check(api, ctx, t, 25, canonical, msc)
        say("  min_seed_cov", msc, info)


@pytest.mark.parametrize("walks,lmax", [(1, 4096), (64, 16), (3, 8)])
def test_inchworm_parity_forced_rounds(api, ctx, monkeypatch, walks, lmax):
    """Tiny rounds (one walk: the serial order itself), and path buffers so
    short that walks overflow and the first walk of a round regrows them."""
    say("inchworm_parity_forced_rounds", walks, lmax)
    monkeypatch.setenv("TK_IWORM_WALKS", str(walks))
    monkeypatch.setenv("TK_IWORM_LMAX", str(lmax))
    seqs = repetitive_reads(61, 3000, genome_len=2500, read_len=90, err=0.01) + edge_reads()
    r, t, _, _ = _count(api, ctx, seqs, 25, True)
    info = _iworm_check(api, ctx, t, 25, True, 2)
    say("  ", info)


@pytest.mark.parametrize("k", [13, 25, 31])
def test_inchworm_parity_pruned_with_seed_entropy(api, ctx, k):
    """After Inchworm's prune (min count 2, entropy 1.0), seeds filtered by a
    seed entropy of 1.5; repeats and low-complexity reads make branches."""
    say("inchworm_parity_pruned k", k)
    seqs = repetitive_reads(70 + k, 4000, genome_len=1500, read_len=100, err=0.02) + \
        ["A" * 60, "AC" * 40, "ACGTTGCA" * 12] * 20 + random_reads(71, 300, 30, 120)
    r, t, _, _ = _count(api, ctx, seqs, k, True)
    t.prune(2, api.entropy_keep_table(k, 1.0))
    keep = api.entropy_keep_table(k, 1.5)
    _iworm_check(api, ctx, t, k, True, 2, keep)
    _iworm_check(api, ctx, t, k, True, 1, None)


def test_inchworm_sample_reads_rebuild_transcripts(api, ctx):
    """The reference's sample reads: contigs equal the oracle's, and every
    25-mer of the 11 transcripts the reads came from lies in some contig."""
    from data import ref_seqs, sample_fasta
    say("inchworm_sample_reads")
    seq, off = concat(sample_fasta("left")[1] + sample_fasta("right")[1])
    r, t, _, _ = _count(api, ctx, (seq, off), 25, True)
    _iworm_check(api, ctx, t, 25, True, 2)
    contigs, sums, seeds, info = api.inchworm(ctx, t, 2)
    in_contigs = set()
    for c in contigs:
        for i in range(len(c) - 24):
            w = c[i:i + 25]
            in_contigs.add(min(w, brute.rc(w)))
    for ref in ref_seqs():
        ks = [min(ref[i:i + 25], brute.rc(ref[i:i + 25])) for i in range(len(ref) - 24)]
        assert all(k in in_contigs for k in ks)


def _kset_keys(strings, tags=None):
    out = np.zeros((len(strings), 2), np.uint64)
    for i, s in enumerate(strings):
        v = brute.kmer_value(s)
        out[i] = (v, tags[i]) if tags is not None else (v & ((1 << 64) - 1), v >> 64)
    return out


@pytest.mark.parametrize("kk,canonical", [(48, True), (48, False), (63, True), (40, True), (25, False)])
def test_kset_count_parity(api, ctx, kk, canonical):
    """Counts of a fixed set of long k-mers (weld k-mers are 48-mers) in the
    reads equal the oracle's; keys are a mix of read windows and misses."""
    say("kset_count kk", kk, "canonical", canonical)
    import random
    rnd = random.Random(kk)
    seqs = repetitive_reads(kk, 20000, genome_len=5000, read_len=150, err=0.01) + edge_reads() + \
        random_reads(kk + 1, 500, 40, 300, n_rate=0.03)
    cand = set()
    for s in seqs[:3000]:
        u = s.upper()
        for i in range(0, max(0, len(u) - kk + 1), 11):
            w = u[i:i + kk]
            if all(ch in "ACGT" for ch in w):
                cand.add(min(w, brute.rc(w)) if canonical else w)
    for _ in range(2000):
        cand.add("".join(rnd.choice("ACGT") for _ in range(kk)))
    keys = sorted(cand)
    rnd.shuffle(keys)
    kv = _kset_keys(keys)
    seq, off = concat(seqs)
    ks = api.KSet(ctx, kv)
    got = ks.count(api.Reads(ctx, seq, off), kk, canonical)
    want = O.kset_count(kv, seq, off, kk, canonical)
    assert np.array_equal(got, want) and int(want.sum()) > 0
    ks.free()


def test_kset_count_tagged_and_errors(api, ctx):
    say("kset_count tagged")
    k = 24
    seqs = repetitive_reads(5, 30000, genome_len=4000, read_len=150, err=0.01)
    tags = np.array([i % 7 for i in range(len(seqs))], np.uint32)
    keys = set()
    for t in range(7):
        for s in seqs[t:4000:7]:
            for i in range(0, len(s) - k + 1, 9):
                w = s[i:i + k]
                keys.add((t, min(w, brute.rc(w))))
    keys = sorted(keys)
    kv = _kset_keys([w for _, w in keys], [t for t, _ in keys])
    seq, off = concat(seqs)
    ks = api.KSet(ctx, kv)
    got = ks.count(api.Reads(ctx, seq, off), k, True, tags)
    assert np.array_equal(got, O.kset_count(kv, seq, off, k, True, tags))
    ks.free()
    with pytest.raises(api.TkError, match="not distinct"):
        api.KSet(ctx, np.concatenate([kv[:50], kv[10:11]]))
    empty = api.KSet(ctx, np.zeros((0, 2), np.uint64))
    assert len(empty.count(api.Reads(ctx, seq, off), k, True)) == 0
