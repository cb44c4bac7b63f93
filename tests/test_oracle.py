"""Pins of the C oracle (CPU): brute force, closed forms and Perl's own rand()."""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

import brute
from data import concat, edge_reads, random_reads, repetitive_reads
from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"


def _as_dict(keys, counts, k):
    return {brute.kmer_of(int(a), k): int(b) for a, b in zip(keys, counts)}


@pytest.mark.parametrize("k", [1, 5, 13, 25, 31])
def test_revcomp_matches_strings(k):
    rng = np.random.default_rng(k)
    for _ in range(200):
        v = int(rng.integers(0, 1 << (2 * k), dtype=np.uint64))
        s = brute.kmer_of(v, k)
        assert brute.kmer_of(O.revcomp(v, k), k) == brute.rc(s)
        assert O.revcomp(O.revcomp(v, k), k) == v


@pytest.mark.parametrize("canonical", [True, False])
@pytest.mark.parametrize("k", [5, 25])
def test_count_matches_brute_force(k, canonical):
    seqs = random_reads(7 + k, 300) + edge_reads() + repetitive_reads(3, 200)
    seq, off = concat(seqs)
    keys, counts = O.count(seq, off, k, canonical)
    assert np.all(np.diff(keys.astype(np.int64)) > 0) if len(keys) > 1 else True
    assert _as_dict(keys, counts, k) == dict(brute.count(seqs, k, canonical))


def test_count_closed_forms():
    k = 25
    seqs = repetitive_reads(11, 500) + random_reads(12, 200)
    seq, off = concat(seqs)
    keys, counts = O.count(seq, off, k, True)
    total, valid = O.window_totals(seq, off, k)
    assert int(counts.sum()) == valid
    assert total == sum(max(0, len(s) - k + 1) for s in seqs)
    # canonical counting is invariant under reverse-complementing every read
    seq2, off2 = concat([brute.rc(s) for s in seqs])
    k2, c2 = O.count(seq2, off2, k, True)
    assert np.array_equal(keys, k2) and np.array_equal(counts, c2)
    # every canonical key is <= its reverse complement
    for v in keys[:2000]:
        assert int(v) <= O.revcomp(int(v), k)
    h = O.histo(counts)
    assert int(h.sum()) == len(keys)


@pytest.mark.parametrize("low,high,inc", [(1, 10000, 1), (2, 50, 1), (1, 20, 3), (5, 40, 2), (3, 9, 5)])
def test_histo_matches_brute_force(low, high, inc):
    rng = np.random.default_rng(low * 100 + high)
    counts = rng.integers(1, 80, size=3000).astype(np.uint32)
    counts[:5] = [1, 10000, 10001, 123456, 2]
    assert O.histo(counts, low, high, inc).tolist() == brute.histo(counts.tolist(), low, high, inc)


def test_entropy_closed_forms():
    k = 25
    assert O.entropy_counts(25, 0, 0, 0, k) == 0.0
    assert math.isclose(O.entropy_counts(12, 13, 0, 0, k),
                        -(12 / 25) * math.log2(12 / 25) - (13 / 25) * math.log2(13 / 25))
    # entropy of a composition is symmetric in the bases (DS-mode invariance)
    assert O.entropy_counts(3, 5, 7, 10, k) == O.entropy_counts(10, 7, 5, 3, k)
    for w in ["A" * 25, "ACGT" * 6 + "A", "GATTACA" * 3 + "GATT"]:
        a, c, g, t = (w.count(b) for b in "ACGT")
        assert math.isclose(O.entropy_counts(a, c, g, t, k), brute.entropy(w), rel_tol=1e-12)


@pytest.mark.parametrize("thr", [0.0, 1.0, 1.5, 1.9])
def test_entropy_keep_table_is_robust(thr):
    """No 25-mer composition lies within 1e-6 of the thresholds Trinity uses, so
    float vs double entropy (an open detail of the reference) cannot flip a decision."""
    k = 25
    keep = O.entropy_keep_table(k, thr)
    K1 = k + 1
    for a in range(K1):
        for c in range(K1 - a):
            for g in range(K1 - a - c):
                t = k - a - c - g
                h = O.entropy_counts(a, c, g, t, k)
                if thr > 0:
                    assert abs(h - thr) > 1e-6
                assert keep[(a * K1 + c) * K1 + g] == (h >= thr)
    # float32 evaluation agrees everywhere
    for a in range(K1):
        for c in range(K1 - a):
            for g in range(K1 - a - c):
                t = k - a - c - g
                hf = np.float32(0)
                for x in (a, c, g, t):
                    if x:
                        p = np.float32(x) / np.float32(k)
                        hf -= p * np.float32(np.log2(p))
                assert (float(hf) >= thr) == bool(keep[(a * K1 + c) * K1 + g])


def test_prune_matches_filter():
    k = 25
    seqs = repetitive_reads(5, 400, genome_len=300) + ["A" * 40, "AC" * 30, "AAAAAAAAAAAACCCCCCCCCCCCCC"]
    seq, off = concat(seqs)
    keys, counts = O.count(seq, off, k)
    for mc, me in [(1, 0.0), (2, 1.0), (3, 1.5), (1, 1.9)]:
        pk, pc = O.prune(keys, counts, k, mc, me)
        exp = [(int(a), int(b)) for a, b in zip(keys, counts)
               if b >= mc and brute.entropy(brute.kmer_of(int(a), k)) >= me]
        assert list(zip(pk.tolist(), pc.tolist())) == exp


@pytest.mark.parametrize("median_policy", [0, 1, 2])
@pytest.mark.parametrize("nwin_policy", [0, 1])
@pytest.mark.parametrize("min_count", [1, 2])
def test_read_stats_matches_brute_force(median_policy, nwin_policy, min_count):
    k = 25
    seqs = repetitive_reads(21, 300) + edge_reads() + random_reads(22, 40, n_rate=0.05)
    seq, off = concat(seqs)
    keys, counts = O.count(seq, off, k)
    d = _as_dict(keys, counts, k)
    got = O.read_stats(seq, off, keys, counts, k, True, min_count, 1, nwin_policy, median_policy)
    for s, g in zip(seqs, got):
        e = brute.read_stats(s, d, k, True, min_count, 1, nwin_policy, median_policy)
        assert (int(g["median"]), int(g["nwin"]), int(g["sum"])) == (e["median"], e["nwin"], e["sum"])
        assert int(g["sumsq_lo"]) + (int(g["sumsq_hi"]) << 64) == e["sumsq"]


def test_stats_row_format():
    st = np.zeros(1, O.READ_STAT_DTYPE)[0]
    st["median"], st["nwin"], st["sum"], st["sumsq_lo"] = 4, 3, 12, 4 * 4 + 3 * 3 + 5 * 5
    row = O.format_stats_row("r1/1", st)
    mean = 4.0
    sd = math.sqrt((3 * 50 - 144)) / 3
    assert row == "r1/1\t4\t%g\t%g\t%g\n" % (mean, sd, sd / mean * 100)


def test_drand48_matches_perl_rand():
    want = [float(x) for x in (GOLDEN / "perl_rand_12345.txt").read_text().split()]
    got = O.drand48_stream(12345, len(want))
    assert got == want
    py = brute.Drand48(12345)
    assert [py() for _ in range(len(want))] == want


def test_perl_avg1_rounding():
    assert O.perl_avg1(3, 4) == 3.5
    assert O.perl_avg1(0.25, 0.0) == float("%.1f" % 0.125)
    assert O.perl_avg1(12.3456, 7.1) == float("%.1f" % ((12.3456 + 7.1) / 2))


def test_select_matches_perl_semantics():
    rng = np.random.default_rng(5)
    n = 5000
    lm = rng.integers(0, 300, n)
    rm = rng.integers(0, 300, n)
    lu = np.round(rng.random(n) * 200, 4)
    ru = np.round(rng.random(n) * 200, 4)
    lu[:20] = 0
    ru[:20] = 0
    ls = np.round(rng.random(n) * 50, 4)
    rs = np.round(rng.random(n) * 50, 4)
    rows = [(str(a), repr(float(b)), repr(float(c)), str(d), repr(float(e)), repr(float(f)))
            for a, b, c, d, e, f in zip(lm, lu, ls, rm, ru, rs)]
    for max_cov, max_cv in [(50, 10000), (200, 10000), (30, 1)]:
        want = brute.perl_select(rows, max_cov, 1, max_cv)
        got = O.select_pairs(lm, rm, lu, ru, ls, rs, max_cov, 1, max_cv)
        assert np.flatnonzero(got).tolist() == want


@pytest.mark.parametrize("canonical", [True, False])
def test_r2t_matches_brute_force(canonical):
    k = 25
    bundles = ["X".join(random_reads(40 + i, 3, 60, 300, n_rate=0.0, lower_rate=0.0)) for i in range(30)]
    comps = list(range(100, 130))
    # shared k-mers between components, and a duplicated bundle to exercise the min rule
    bundles[5] = bundles[5] + "X" + bundles[9][:80]
    bundles.append(bundles[3])
    comps.append(7)
    reads = []
    rnd = np.random.default_rng(3)
    for i in range(400):
        b = bundles[int(rnd.integers(0, len(bundles)))]
        s = int(rnd.integers(0, max(1, len(b) - 76)))
        r = b[s:s + 76].replace("X", "N")
        reads.append(brute.rc(r) if (canonical and rnd.random() < 0.5) else r)
    reads += edge_reads()
    bseq, boff = concat(bundles)
    keys, cv = O.r2t_build(bseq, boff, comps, k, canonical)
    rseq, roff = concat(reads)
    got = O.r2t_vote(rseq, roff, keys, cv, k, canonical)
    want = brute.r2t(bundles, comps, reads, k, canonical)
    assert [(int(g["comp"]), int(g["hits"]), int(g["nwin"])) for g in got] == want


def test_cpuref_matches_oracle():
    """The multi-threaded CPU baseline computes what the oracle computes."""
    import numpy as np
    from data import concat, edge_reads, random_reads
    from oracle import oracle as O
    seqs = random_reads(31, 3000, 0, 300, n_rate=0.01) + edge_reads()
    seq, off = concat(seqs)
    ok, oc = O.count(seq, off, 25, True)
    want = O.read_stats(seq, off, ok, oc, 25, True, 2, 1, 0, 0)
    seqs2 = repetitive_reads(9, 2000, genome_len=500)
    s2, o2 = concat(seqs2)
    k2, c2 = O.count(s2, o2, 25, True)
    assert int(c2.max()) > 1
    for threads in (1, 4):
        got = O.cpuref_count_stats(s2, o2, 25, True, 2, threads, summary=True)
        assert got["digest"] == O.table_digest(k2, c2)
        assert np.array_equal(got["histo"], O.histo(c2).astype(np.uint64))
        # the digest tells tables apart: one count off, or one key dropped
        c3 = c2.copy()
        c3[7] += 1
        assert O.table_digest(k2, c3) != got["digest"]
        assert O.table_digest(k2[1:], c2[1:]) != got["digest"]
    for threads in (1, 4):
        got = O.cpuref_count_stats(seq, off, 25, True, 2, threads)
        assert got["distinct"] == len(ok)
        assert np.array_equal(got["median"], want["median"])
        assert np.array_equal(got["nwin"], want["nwin"])
        assert np.array_equal(got["sum"], want["sum"])


# ------------------------------------------- golden vectors of the reference scripts

from data import (NBKC, NBKC_SETTINGS, nbkc_aggregates, nbkc_pairs,  # noqa: E402
                  nbkc_selected)


def test_nbkc_fixture_text_is_the_aggregates_formatted():
    """The stats text the reference scripts read is the oracle's row format
    of the committed integer aggregates (which the GPU selection reads)."""
    for side in ("left", "right"):
        accs, st = nbkc_aggregates(side, O.READ_STAT_DTYPE)
        rows = (NBKC / f"{side}.stats.sort").read_text().splitlines(keepends=True)[1:]
        assert rows == [O.format_stats_row(a, s) for a, s in zip(accs, st)]
        # C-locale order, as `sort -k1,1` leaves the files for --sorted
        assert [a.encode() for a in accs] == sorted(a.encode() for a in accs)


@pytest.mark.parametrize("max_cov,min_cov,max_cv", NBKC_SETTINGS)
def test_select_matches_reference_scripts(max_cov, min_cov, max_cv):
    """or_select_pairs on the left/right values of every pair the reference
    merge kept selects exactly the accessions nbkc_normalize.pl printed."""
    pairs = nbkc_pairs()
    f = lambda i: [float(p[i]) for p in pairs]  # noqa: E731
    sel = O.select_pairs(f(2), f(6), f(3), f(7), f(4), f(8), max_cov, min_cov, max_cv)
    got = [p[0] for p, s in zip(pairs, sel) if s]
    assert got == nbkc_selected("pairs", max_cov, min_cov, max_cv)


def test_single_end_select_matches_reference_script():
    """Single-end: nbkc_normalize.pl on one stats file (no merge).  Its
    median is an integer and max_CV = 10000 leaves only mean > 0, so the pair
    rule with the read on both sides makes the same decisions."""
    rows = [ln.split("\t") for ln in (NBKC / "left.stats.sort").read_text().splitlines()[1:]]
    f = lambda i: [float(r[i]) for r in rows]  # noqa: E731
    sel = O.select_pairs(f(1), f(1), f(2), f(2), f(3), f(3), 50, 0, 10000)
    got = [r[0][:-2] for r, s in zip(rows, sel) if s]
    assert got == nbkc_selected("left", 50, 0, 10000)


def test_merge_matches_reference_script():
    """The pairs the reference merge keeps are the accessions present on both
    sides; its median/mean/stdev columns are %.1f of the two sides' average."""
    la, _ = nbkc_aggregates("left", O.READ_STAT_DTYPE)
    ra, _ = nbkc_aggregates("right", O.READ_STAT_DTYPE)
    both = sorted(set(a[:-2] for a in la) & set(a[:-2] for a in ra), key=lambda s: (s + "/1").encode())
    pairs = nbkc_pairs()
    assert [p[0] for p in pairs] == both
    for p in pairs:
        for c, (a, b) in zip((9, 10, 11), ((2, 6), (3, 7), (4, 8))):
            assert float(p[c]) == O.perl_avg1(float(p[a]), float(p[b]))


def test_stats_row_matches_textbook_mean_and_pstdev():
    """mean = statistics.mean and stdev = statistics.pstdev (population) of
    explicit per-window values, printed %g; median = upper middle value."""
    import statistics
    rng = np.random.default_rng(77)
    for trial in range(400):
        n = int(rng.integers(1, 127))
        v = np.maximum(1, rng.poisson(float(np.exp(rng.normal(2, 2))), n)).astype(np.int64)
        if trial % 7 == 0:
            v[:] = int(v[0])  # constant: stdev exactly 0
        st = np.zeros(1, O.READ_STAT_DTYPE)[0]
        s = np.sort(v)
        st["median"], st["nwin"], st["sum"] = int(s[n // 2]), n, int(v.sum())
        sq = int((v * v).sum())
        st["sumsq_lo"], st["sumsq_hi"] = sq & ((1 << 64) - 1), sq >> 64
        f = O.format_stats_row("r", st).rstrip("\n").split("\t")
        mean = statistics.mean(v.tolist())
        sd = statistics.pstdev(v.tolist())
        assert f[1] == str(int(s[n // 2]))
        assert f[2] == "%g" % mean
        assert f[3] == "%g" % sd
        assert f[4] == "%g" % (sd / mean * 100)


@pytest.mark.parametrize("name", ["all", "NormMaxKCov50"])
def test_histo_default_shape_matches_reference_witness(name):
    """jellyfish histo's default shape, witnessed by the reference's
    kmer_histo files: labels from 1, the top label 10001 collecting every
    count above 10000, and (NormMaxKCov50) bins with no k-mer omitted."""
    from data import counts_for_histo_witness, histo_witness
    w = histo_witness(name)
    assert w[0][0] == 1 and all(n > 0 for _, n in w)
    labels = [a for a, _ in w]
    if name == "all":  # every bin filled, the last one the >10000 bin
        assert labels == list(range(1, 10002))
    else:  # after normalization: empty bins (and the empty top bin) omitted
        assert labels[-1] < 10001 and len(labels) < labels[-1]
    counts, want = counts_for_histo_witness(name)
    bins = O.histo(counts)
    base = O.histo_base(1, 1)
    got = [(base + i, int(b)) for i, b in enumerate(bins) if b]
    assert got == want


# ------------------------------------------- Inchworm greedy contigs

def _dict_from_reads(seqs, k, ds):
    seq, off = concat(seqs)
    keys, counts = O.count(seq, off, k, ds)
    return keys, counts


@pytest.mark.parametrize("k,ds", [(5, True), (11, True), (25, True), (11, False), (25, False)])
def test_inchworm_matches_brute_force(k, ds):
    """or_inchworm equals the string brute force on dictionaries with
    branches (repeats, errors) and low-count tails; seed coverage 1 and 2."""
    seqs = repetitive_reads(40 + k, 600, genome_len=700, read_len=60, err=0.01) + random_reads(41, 60, 30, 80)
    keys, counts = _dict_from_reads(seqs, k, ds)
    d = _as_dict(keys, counts, k)
    for msc in (1, 2):
        contigs, sums, seeds = O.inchworm(keys, counts, k, ds, msc)
        want = brute.inchworm(d, k, ds, msc)
        assert [c for c, _, _ in want] == contigs
        assert [t for _, t, _ in want] == sums.tolist()
        assert [s for _, _, s in want] == [brute.kmer_of(int(x), k) for x in seeds]


def test_inchworm_seed_entropy_and_pruned_entries():
    k = 11
    seqs = repetitive_reads(3, 400, genome_len=300, read_len=50) + ["A" * 40, "AC" * 25] * 5
    keys, counts = _dict_from_reads(seqs, k, True)
    counts = counts.copy()
    counts[::7] = 0  # pruned entries stay in the table with count 0
    d = {brute.kmer_of(int(a), k): int(b) for a, b in zip(keys, counts)}
    keep = O.entropy_keep_table(k, 1.5)
    contigs, sums, seeds = O.inchworm(keys, counts, k, True, 2, keep)
    want = brute.inchworm(d, k, True, 2, 1.5)
    assert [c for c, _, _ in want] == contigs and [t for _, t, _ in want] == sums.tolist()
    for s in seeds:
        assert brute.entropy(brute.kmer_of(int(s), k)) >= 1.5


def test_inchworm_reassembles_a_transcript():
    """Error-free reads tiling one repeat-free transcript at depth >= 2: one
    contig, the transcript or its reverse complement (DS), every k-mer used
    once, cov = total count / k-mers."""
    import random
    rnd = random.Random(5)
    tx = "".join(rnd.choice("ACGT") for _ in range(3000))
    k = 25
    reads = [tx[i:i + 100] for i in range(0, len(tx) - 100 + 1, 7)] + [tx[-100:], tx[-50:]]
    reads = [r if j % 2 else brute.rc(r) for j, r in enumerate(reads)]
    keys, counts = _dict_from_reads(reads, k, True)
    assert len(keys) == len(tx) - k + 1
    contigs, sums, _ = O.inchworm(keys, counts, k, True, 2)
    assert len([c for c in contigs if len(c) > k]) == 1
    main = max(contigs, key=len)
    # k-mers at the transcript's ends are covered once and cannot seed but are reached by extension
    assert main in (tx, brute.rc(tx))
    assert int(sums[contigs.index(main)]) == int(counts.sum())


def test_inchworm_uses_each_kmer_once():
    k = 25
    seqs = repetitive_reads(77, 2000, genome_len=3000, read_len=100, err=0.005)
    keys, counts = _dict_from_reads(seqs, k, True)
    contigs, sums, seeds = O.inchworm(keys, counts, k, True, 1)
    seen = set()
    d = _as_dict(keys, counts, k)
    for c, t in zip(contigs, sums):
        ks = [min(c[i:i + k], brute.rc(c[i:i + k])) for i in range(len(c) - k + 1)]
        assert not (set(ks) & seen) and len(set(ks)) == len(ks)
        seen |= set(ks)
        assert sum(d[x] for x in ks) == int(t)
    assert seen == set(d)  # seed coverage 1: every k-mer ends in some contig


# ------------------------------------------- counts of a fixed k-mer set

def _keys_of(strings, tags=None):
    out = np.zeros((len(strings), 2), np.uint64)
    for i, s in enumerate(strings):
        v = brute.kmer_value(s)
        if tags is not None:
            out[i] = (v, tags[i])
        else:
            out[i] = (v & ((1 << 64) - 1), v >> 64)
    return out


@pytest.mark.parametrize("kk,canonical", [(48, True), (48, False), (64, True), (33, True), (25, True)])
def test_kset_count_matches_brute_force(kk, canonical):
    import random
    rnd = random.Random(kk)
    seqs = repetitive_reads(kk, 800, genome_len=600, read_len=120, err=0.01) + edge_reads() + \
        random_reads(kk + 1, 50, 40, 150, n_rate=0.03)
    # keys: windows of the genome-like reads (hits) and random strings (misses), canonical forms
    cand = set()
    for s in seqs[:300]:
        u = s.upper()
        for i in range(0, max(0, len(u) - kk + 1), 7):
            w = u[i:i + kk]
            if all(ch in "ACGT" for ch in w):
                cand.add(min(w, brute.rc(w)) if canonical else w)
    for _ in range(200):
        cand.add("".join(rnd.choice("ACGT") for _ in range(kk)))
    keys = sorted(cand)
    rnd.shuffle(keys)
    seq, off = concat(seqs)
    got = O.kset_count(_keys_of(keys), seq, off, kk, canonical)
    want = brute.kset_count(keys, seqs, kk, canonical)
    assert got.tolist() == want
    assert sum(want) > 0


def test_kset_count_tagged_matches_brute_force():
    k = 24
    seqs = repetitive_reads(4, 900, genome_len=500, read_len=100, err=0.01)
    tags = [i % 3 for i in range(len(seqs))]
    keys = []
    for t in range(3):
        for s in seqs[t:300:3]:
            for i in range(0, len(s) - k + 1, 5):
                w = s[i:i + k]
                keys.append((t, min(w, brute.rc(w))))
    keys = sorted(set(keys))
    seq, off = concat(seqs)
    got = O.kset_count(_keys_of([w for _, w in keys], [t for t, _ in keys]), seq, off, k, True, tags)
    assert got.tolist() == brute.kset_count(keys, seqs, k, True, tags)
