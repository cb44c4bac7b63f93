"""The drop-in CLI tools keep the reference command lines and formats.

Host-only paths (jellyfish dump/histo over a database, nbkc_normalize) run on
the CPU; paths that count or score run the HIP library (marked gpu).
"""
from __future__ import annotations

import os
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

import brute
from data import concat, edge_reads, random_reads, repetitive_reads
from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "trinityrnaseq_amd" / "bin"


def run(args, cwd=None, check=True, **kw):
    r = subprocess.run([str(a) for a in args], cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, **kw)
    if check and r.returncode != 0:
        raise AssertionError(f"{args} -> {r.returncode}\n{r.stderr}")
    return r


def write_db(path, k, canonical, keys, counts):
    with open(path, "wb") as f:
        f.write(b"TKJF1\n")
        f.write(struct.pack("<IIQ", k, canonical, len(keys)))
        f.write(np.ascontiguousarray(keys, np.uint64).tobytes())
        f.write(np.ascontiguousarray(counts, np.uint32).tobytes())


def write_fasta(path, names, seqs):
    with open(path, "w") as f:
        for n, s in zip(names, seqs):
            f.write(f">{n}\n{s}\n")


def dump_text(keys, counts, k, lo=1, hi=0xFFFFFFFF):
    out = []
    for a, c in zip(keys.tolist(), counts.tolist()):
        if lo <= c <= hi:
            out.append(f">{c}\n{brute.kmer_of(a, k)}\n")
    return "".join(out)


@pytest.fixture(scope="module")
def sample(tmp_path_factory):
    d = tmp_path_factory.mktemp("tools")
    seqs = repetitive_reads(31, 3000, genome_len=2000, read_len=76) + random_reads(32, 300) + edge_reads()
    names = [f"read{i:06d}/1" for i in range(len(seqs))]
    fa = d / "reads.fa"
    write_fasta(fa, names, seqs)
    seq, off = concat(seqs)
    keys, counts = O.count(seq, off, 25, True)
    return d, fa, names, seqs, seq, off, keys, counts


def test_version_string():
    r = run([BIN / "jellyfish", "--version"])
    name, version = r.stdout.split()[:2]
    assert version.startswith("2.")  # Trinity:4070-4074


def test_dump_and_histo_from_database(sample):
    d, fa, names, seqs, seq, off, keys, counts = sample
    db = d / "oracle.jf"
    write_db(db, 25, 1, keys, counts)
    assert run([BIN / "jellyfish", "dump", db]).stdout == dump_text(keys, counts, 25)
    assert run([BIN / "jellyfish", "dump", "-L", "2", db]).stdout == dump_text(keys, counts, 25, 2)
    col = run([BIN / "jellyfish", "dump", "-c", "-L", "3", "-U", "9", db]).stdout
    want = "".join(f"{brute.kmer_of(a, 25)} {c}\n" for a, c in zip(keys.tolist(), counts.tolist()) if 3 <= c <= 9)
    assert col == want
    run([BIN / "jellyfish", "histo", "-t", "4", "-o", d / "h.histo", db])
    bins = O.histo(counts)
    base = O.histo_base(1, 1)
    want = "".join(f"{base + i} {int(b)}\n" for i, b in enumerate(bins) if b)
    assert (d / "h.histo").read_text() == want
    h = run([BIN / "jellyfish", "histo", "-l", "2", "-h", "20", "-i", "3", "-f", db]).stdout
    bins = O.histo(counts, 2, 20, 3)
    base = O.histo_base(2, 3)
    assert h == "".join(f"{base + 3 * i} {int(b)}\n" for i, b in enumerate(bins))


def test_tools_reject_unknown_options():
    r = run([BIN / "jellyfish", "count", "-m", "25", "--bogus", "x.fa"], check=False)
    assert r.returncode != 0 and "unknown option" in r.stderr
    r = run([BIN / "ReadsToTranscripts", "-i", "a", "-f", "b", "-o", "c", "-zzz"], check=False)
    assert r.returncode != 0


def _stats_rows(rng, n):
    rows = []
    for i in range(n):
        med = int(rng.integers(0, 400))
        nw = 126
        vals = rng.integers(max(med // 2, 1), med * 2 + 2, nw)
        st = np.zeros(1, O.READ_STAT_DTYPE)[0]
        st["median"], st["nwin"], st["sum"] = med, nw, int(vals.sum())
        sq = int((vals.astype(object) ** 2).sum())
        st["sumsq_lo"], st["sumsq_hi"] = sq & ((1 << 64) - 1), sq >> 64
        if i % 97 == 0:
            st["median"], st["nwin"], st["sum"], st["sumsq_lo"], st["sumsq_hi"] = 0, 0, 0, 0, 0
        rows.append(st)
    return rows


def test_nbkc_normalize_matches_perl_semantics(tmp_path):
    rng = np.random.default_rng(8)
    n = 4000
    left = _stats_rows(rng, n)
    right = _stats_rows(rng, n)
    # sorted accessions, with a few unpaired reads on either side
    lnames = [f"p{i:07d}/1" for i in range(n)]
    rnames = [f"p{i:07d}/2" for i in range(n)]
    drop_l = set(range(5, n, 611))
    drop_r = set(range(7, n, 509))
    hdr = "acc\tmedian_cov\tmean_cov\tstdev\tpct_stdev_of_avg\n"
    with open(tmp_path / "L", "w") as f:
        f.write(hdr)
        f.writelines(O.format_stats_row(lnames[i], left[i]) for i in range(n) if i not in drop_l)
    with open(tmp_path / "R", "w") as f:
        f.write(hdr)
        f.writelines(O.format_stats_row(rnames[i], right[i]) for i in range(n) if i not in drop_r)
    for max_cov in (30, 50, 200):
        r = run([BIN / "nbkc_normalize", "--left", tmp_path / "L", "--right", tmp_path / "R", "--sorted",
                 "--max_cov", max_cov, "--min_cov", 1, "--max_CV", 10000, "--merged_out", tmp_path / "M"])
        rows, cores = [], []
        for i in range(n):
            if i in drop_l or i in drop_r:
                continue
            lf = O.format_stats_row("x", left[i]).split("\t")
            rf = O.format_stats_row("x", right[i]).split("\t")
            rows.append((lf[1], lf[2], lf[3], rf[1], rf[2], rf[3]))
            cores.append(f"p{i:07d}")
        want = [cores[j] for j in brute.perl_select(rows, max_cov, 1, 10000)]
        assert r.stdout.split() == want
        # the merged file alone gives the same selection (nbkc_normalize.pl --stats_file)
        r2 = run([BIN / "nbkc_normalize", "--stats_file", tmp_path / "M", "--max_cov", max_cov,
                  "--min_cov", 1, "--max_CV", 10000])
        assert r2.stdout == r.stdout
    merged = (tmp_path / "M").read_text().splitlines()
    assert merged[0].split("\t")[-3:] == ["median_cov", "mean_cov", "stdev"]
    assert len(merged) == 1 + len(cores)


# ---------------------------------------------------------------- GPU paths

@pytest.mark.gpu
def test_jellyfish_count_dump_histo(sample):
    d, fa, names, seqs, seq, off, keys, counts = sample
    run([BIN / "jellyfish", "count", "-t", "4", "-m", "25", "-s", "100000000", "-o", d / "m.jf", "--canonical", fa])
    assert run([BIN / "jellyfish", "dump", "-L", "1", d / "m.jf"]).stdout == dump_text(keys, counts, 25)
    run([BIN / "jellyfish", "count", "-t", "4", "-m", "25", "-s", "1000", fa], cwd=d)  # default mer_counts.jf
    assert (d / "mer_counts.jf").exists()
    fk, fc = O.count(seq, off, 25, False)
    assert run([BIN / "jellyfish", "dump", d / "mer_counts.jf"]).stdout == dump_text(fk, fc, 25)


@pytest.mark.gpu
def test_fastaToKmerCoverageStats_matches_oracle(sample):
    d, fa, names, seqs, seq, off, keys, counts = sample
    dump = d / "kmers.min2.fa"
    dump.write_text(dump_text(keys, counts, 25, 2))
    out = run([BIN / "fastaToKmerCoverageStats", "--reads", fa, "--kmers", dump, "--kmer_size", "25",
               "--num_threads", "4", "--DS"]).stdout.splitlines(True)
    assert out[0] == "acc\tmedian_cov\tmean_cov\tstdev\tpct_stdev_of_avg\n"
    m = counts >= 2
    st = O.read_stats(seq, off, keys[m], counts[m], 25, True, 1, 1, 0, 0)
    want = [O.format_stats_row(n, s) for n, s in zip(names, st)]
    assert out[1:] == want
    # --kmers_from_reads counts the dictionary itself
    out2 = run([BIN / "fastaToKmerCoverageStats", "--kmers_from_reads", fa, "--reads", fa, "--DS",
                "--num_threads", "1"]).stdout.splitlines(True)
    st2 = O.read_stats(seq, off, keys, counts, 25, True, 1, 1, 0, 0)
    assert out2[1:] == [O.format_stats_row(n, s) for n, s in zip(names, st2)]


def iworm_text(contigs, sums, k, min_len=25, min_cov=1):
    out, n = [], 0
    for c, t in zip(contigs, sums.tolist()):
        cov = t // (len(c) - k + 1)
        if len(c) < min_len or cov < min_cov:
            continue
        n += 1
        out.append(f">a{n};{cov} K: {k} length: {len(c)}\n{c}\n")
    return "".join(out)


@pytest.mark.gpu
def test_inchworm_load_prune(sample):
    d, fa, names, seqs, seq, off, keys, counts = sample
    dump = d / "jellyfish.kmers.fa"
    dump.write_text(dump_text(keys, counts, 25))
    r = run([BIN / "inchworm", "--kmers", dump, "--run_inchworm", "-K", "25", "--monitor", "1", "--DS",
             "--num_threads", "6", "--PARALLEL_IWORM", "--min_any_entropy", "1.5", "-L", "25",
             "--no_prune_error_kmers", "--minKmerCount", "2", "--dump_pruned", d / "pruned.fa"], cwd=d)
    pk, pc = O.prune(keys, counts, 25, 2, 1.5)
    assert (d / "inchworm.kmer_count").read_text().strip() == str(len(pk))
    assert (d / "pruned.fa").read_text() == dump_text(pk, pc, 25)
    contigs, sums, _ = O.inchworm(pk, pc, 25, True, 2, O.entropy_keep_table(25, 1.5))
    assert r.stdout == iworm_text(contigs, sums, 25)
    assert r.stdout.startswith(">a1;")
    # -L and --min_assembly_coverage filter the report; Inchworm's error pruning is refused
    r2 = run([BIN / "inchworm", "--kmers", dump, "--run_inchworm", "-K", "25", "--DS", "-L", "60",
              "--min_assembly_coverage", "3", "--min_seed_entropy", "0", "--min_seed_coverage", "1",
              "--no_prune_error_kmers"], cwd=d)
    contigs, sums, _ = O.inchworm(keys, counts, 25, True, 1, None)
    assert r2.stdout == iworm_text(contigs, sums, 25, 60, 3)
    r3 = run([BIN / "inchworm", "--kmers", dump, "--run_inchworm", "-K", "25", "--DS"], cwd=d, check=False)
    assert r3.returncode != 0 and "no_prune_error_kmers" in r3.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("strand", [False, True])
def test_reads_to_transcripts(tmp_path, strand):
    from trinityrnaseq_amd import synth
    tx = synth.large_transcriptome(91, 300)
    bundles, comps = [], []
    i, cid = 0, 0
    rng = np.random.default_rng(2)
    while i < tx.n:
        m = int(rng.integers(1, 4))
        bundles.append("X".join(tx.seq(j) for j in range(i, min(i + m, tx.n))))
        comps.append(cid * 3 + 1)
        cid += 1
        i += m
    write_fasta(tmp_path / "bundles.fa", [f"s_{c}" for c in comps], bundles)
    s, o = synth.se_reads(tx, 20000, 100, seed=9, sub_rate=0.02)
    reads = [s[int(o[j]):int(o[j + 1])].tobytes().decode() for j in range(len(o) - 1)]
    reads = [r.lower() if j % 50 == 0 else r for j, r in enumerate(reads)] + edge_reads()
    names = [f"q{j}/1" for j in range(len(reads))]
    write_fasta(tmp_path / "reads.fa", names, reads)
    args = [BIN / "ReadsToTranscripts", "-i", tmp_path / "reads.fa", "-f", tmp_path / "bundles.fa",
            "-o", tmp_path / "r2c.out", "-t", "4", "-max_mem_reads", "5000", "-p", "10"]
    if strand:
        args.append("-strand")
    run(args)
    bseq, boff = concat(bundles)
    keys, cv = O.r2t_build(bseq, boff, comps, 25, not strand)
    rseq, roff = concat(reads)
    votes = O.r2t_vote(rseq, roff, keys, cv, 25, not strand)
    want = []
    for nm, r, v in zip(names, reads, votes):
        if v["hits"] and 100 * int(v["hits"]) >= 10 * int(v["nwin"]):
            want.append(f"{int(v['comp'])}\t>{nm}\t{100.0 * int(v['hits']) / int(v['nwin']):g}\t{r.upper()}\n")
    assert (tmp_path / "r2c.out").read_text() == "".join(want)


# ------------------------------------------- golden vectors of the reference scripts

from data import NBKC, NBKC_SETTINGS, counts_for_histo_witness, nbkc_selected, nbkc_stderr  # noqa: E402


@pytest.mark.parametrize("max_cov,min_cov,max_cv", NBKC_SETTINGS)
def test_nbkc_normalize_equals_reference_scripts(tmp_path, max_cov, min_cov, max_cv):
    """The CLI on the committed stats files writes byte for byte what
    nbkc_merge_left_right_stats.pl --sorted and nbkc_normalize.pl wrote:
    the merged pairs file, the selected accessions and the three summary
    lines on stderr."""
    r = run([BIN / "nbkc_normalize", "--left", NBKC / "left.stats.sort", "--right", NBKC / "right.stats.sort",
             "--sorted", "--max_cov", max_cov, "--min_cov", min_cov, "--max_CV", max_cv,
             "--merged_out", tmp_path / "pairs.stats"])
    assert (tmp_path / "pairs.stats").read_bytes() == (NBKC / "pairs.stats").read_bytes()
    assert r.stdout.split("\n")[:-1] == nbkc_selected("pairs", max_cov, min_cov, max_cv)
    assert r.stderr.splitlines()[-3:] == nbkc_stderr("pairs", max_cov, min_cov, max_cv).splitlines()
    r2 = run([BIN / "nbkc_normalize", "--stats_file", NBKC / "pairs.stats", "--max_cov", max_cov,
              "--min_cov", min_cov, "--max_CV", max_cv])
    assert r2.stdout == r.stdout
    assert r2.stderr == nbkc_stderr("pairs", max_cov, min_cov, max_cv)


def test_nbkc_normalize_single_end_equals_reference_script():
    r = run([BIN / "nbkc_normalize", "--stats_file", NBKC / "left.stats.sort", "--max_cov", 50,
             "--min_cov", 0, "--max_CV", 10000])
    assert r.stdout.split("\n")[:-1] == nbkc_selected("left", 50, 0, 10000)
    assert r.stderr == nbkc_stderr("left", 50, 0, 10000)


@pytest.mark.parametrize("name", ["all", "NormMaxKCov50"])
def test_jellyfish_histo_default_shape(tmp_path, name):
    """`jellyfish histo` with default bounds prints the shape of the
    reference's histogram witnesses (labels from 1, empty bins omitted,
    10001 for counts above 10000)."""
    counts, want = counts_for_histo_witness(name)
    keys = np.arange(len(counts), dtype=np.uint64) * 7919
    write_db(tmp_path / "w.jf", 25, 1, keys, counts)
    run([BIN / "jellyfish", "histo", "-t", "2", "-o", tmp_path / "w.histo", tmp_path / "w.jf"])
    assert (tmp_path / "w.histo").read_text() == "".join(f"{a} {b}\n" for a, b in want)


# ------------------------------------------- read ingest (FASTQ -> FASTA) and extraction

import gzip  # noqa: E402
import hashlib  # noqa: E402

INGEST = Path(__file__).resolve().parent / "golden" / "ingest"
SAMPLE = Path(__file__).resolve().parent / "golden" / "sample_data"


@pytest.mark.parametrize("args,want", [
    (["-I", "edge.fq"], "edge.fa"),
    (["-I", "edge.fq", "-a", "1"], "edge.a1.fa"),
    (["-I", "edge.fq", "--rev"], "edge.rev.fa"),
    (["-I", "edge.fq, edge.fq"], "edge.twice.fa"),
    (["-I", "edge_dirty.fq", "--ignore_dirty"], "edge_dirty.ignore.fa"),
])
def test_fastq_to_fasta_equals_reference_script(args, want):
    """fastQ_to_fastA writes byte for byte what fastQ_to_fastA.pl wrote on
    the committed edge-case FASTQ (Casava fields, /1 /2 suffixes, \\r\\n,
    IUPAC reverse complement, dirty records skipped)."""
    r = run([BIN / "fastQ_to_fastA", *args], cwd=INGEST)
    assert r.stdout.encode() == (INGEST / want).read_bytes()


def test_fastq_to_fasta_dirty_record_is_fatal():
    r = run([BIN / "fastQ_to_fastA", "-I", "edge_dirty.fq"], cwd=INGEST, check=False)
    assert r.returncode == int((INGEST / "edge_dirty.fatal.rc").read_text())
    assert r.stdout.encode() == (INGEST / "edge_dirty.fatal.fa").read_bytes()
    assert "improperly formatted entry" in r.stderr


def test_sample_reads_to_fasta_equal_the_shipped_fasta(tmp_path):
    """The reference's sample reads: the FASTA written from reads.*.fq.gz is
    the reads.*.fa.gz shipped beside them, by both tool names; 30,575 records
    per file and 61,150 in both.fa (test_prep.py:26-100); --rev and -a 2
    outputs hash to what fastQ_to_fastA.pl wrote."""
    both = b""
    for side, rt in (("left", "1"), ("right", "2")):
        want = gzip.decompress((SAMPLE / f"reads.{side}.fa.gz").read_bytes())
        r1 = subprocess.run([BIN / "fastQ_to_fastA", "-I", SAMPLE / f"reads.{side}.fq.gz"], stdout=subprocess.PIPE,
                            check=True)
        assert r1.stdout == want
        r2 = subprocess.run([BIN / "seqtk-trinity", "seq", "-A", "-R", rt, SAMPLE / f"reads.{side}.fq.gz"],
                            stdout=subprocess.PIPE, check=True)
        assert r2.stdout == want
        assert want.count(b">") == 30575
        both += want
    assert both.count(b">") == 61150
    sums = {}
    for ln in (INGEST / "sample_sha256.txt").read_text().splitlines():
        h, rest = ln.split("  ", 1)
        f, *args = rest.split()
        sums[f] = (h, args)
    for f, (h, args) in sums.items():
        r = subprocess.run([BIN / "fastQ_to_fastA", "-I", SAMPLE / f, *args], stdout=subprocess.PIPE, check=True)
        assert hashlib.sha256(r.stdout).hexdigest() == h, f


def test_seqtk_trinity_read_type_rule(tmp_path):
    """seqtk-trinity -R N: names carrying /1, /2 or a Casava field keep that
    read type, others get /N; -r reverse-complements (IUPAC table)."""
    fq = tmp_path / "x.fq"
    fq.write_text("@a/1\nACGTN\n+\nIIIII\n@b 2:N:0\nAACCG\n+\nIIIII\n@c\nRYKMB\n+\nIIIII\n")
    r = run([BIN / "seqtk-trinity", "seq", "-A", "-R", "1", fq])
    assert r.stdout == ">a/1\nACGTN\n>b/2\nAACCG\n>c/1\nRYKMB\n"
    r = run([BIN / "seqtk-trinity", "seq", "-A", "-r", "-R", "2", fq])
    assert r.stdout == ">a/1\nNACGT\n>b/2\nCGGTT\n>c/2\nVKMRY\n"


def _fq_records(data: bytes):
    lines = data.split(b"\n")
    return [b"\n".join(lines[i:i + 4]) + b"\n" for i in range(0, len(lines) - 1, 4)]


def test_extract_normalized_reads_fastq(tmp_path):
    """make_normalized_reads_file on FASTQ: every record whose core name
    (no '@', no /1 /2, no _forward/_reverse) is selected, as its 4 lines, in
    file order; selected names may carry /1 or /2; a name no record has is
    an error listed in <out>.missing_accs."""
    left = gzip.decompress((SAMPLE / "reads.left.fq.gz").read_bytes())
    recs = _fq_records(left)
    names = [r.split(b"\n")[0][1:].split()[0].decode()[:-2] for r in recs]
    rng = np.random.default_rng(4)
    pick = sorted(set(rng.choice(len(names), 3000, replace=False).tolist()))
    sel = tmp_path / "sel.accs"
    sel.write_text("".join(names[i] + ("/1\n" if i % 5 == 0 else "\n") for i in reversed(pick)) + "\n  \n")
    out = tmp_path / "left.norm.fq"
    run([BIN / "extract_normalized_reads", "--seqType", "fq", "--selected", sel, "--output", out,
         SAMPLE / "reads.left.fq.gz"])
    assert out.read_bytes() == b"".join(recs[i] for i in pick)
    # both sources in order: left then right records of the same pairs
    right = _fq_records(gzip.decompress((SAMPLE / "reads.right.fq.gz").read_bytes()))
    run([BIN / "extract_normalized_reads", "--seqType", "fq", "--selected", sel, "--output", out,
         SAMPLE / "reads.left.fq.gz", SAMPLE / "reads.right.fq.gz"])
    assert out.read_bytes() == b"".join(recs[i] for i in pick) + b"".join(right[i] for i in pick)
    sel.write_text(names[0] + "\nnot_a_read\nalso_missing/2\n")
    r = run([BIN / "extract_normalized_reads", "--seqType", "fq", "--selected", sel, "--output", out,
             SAMPLE / "reads.left.fq.gz"], check=False)
    assert r.returncode != 0 and "missing 2" in r.stderr
    assert (tmp_path / "left.norm.fq.missing_accs").read_text() == "also_missing\nnot_a_read\n"


def test_extract_normalized_reads_fasta(tmp_path):
    fa = tmp_path / "x.fa"
    fa.write_bytes(b"\n\n>a/1 first read\nACGT\nAC GT\n>b/2\t\nGGGG\n>c\nTT\x01TT\n>d/1\n\n")
    sel = tmp_path / "s"
    sel.write_text("a\nc\nd\n")
    out = tmp_path / "o.fa"
    run([BIN / "extract_normalized_reads", "--seqType", "fa", "--selected", sel, "--output", out, fa])
    assert out.read_bytes() == b">a/1 first read\nACGTACGT\n>c\nTTTT\n>d/1\n\n"


def _rc_iupac(s: str) -> str:
    """seqtk-trinity -r / revcomp_fasta.pl: reverse, IUPAC complement."""
    return s.translate(str.maketrans("ACGTUMRWSYKVHDBNacgtumrwsykvhdbn", "TGCAAKYWSRMBDHVNtgcaakywsrmbdhvn"))[::-1]


def _oracle_normalization(mode: str, max_cov: int, min_cov: int = 0, max_cv: int = 10000, ss: str = ""):
    """insilico_read_normalization.pl on the reference's sample reads, step by
    step through the oracle: mode "pairs" (--pairs_together), "separate"
    (each side selected on its own, :917-944) or "single" (the left reads as
    single-end).  Returns the expected normalized FASTQ bytes per side and
    the selected core names per side."""
    from data import sample_fasta
    sides = ("left",) if mode == "single" else ("left", "right")
    canonical = not ss
    kind = {"left": ss[:1], "right": ss[1:2]} if mode != "single" else {"left": ""}
    names, seqs = {}, {}
    for side in sides:
        n, q = sample_fasta(side)
        names[side], seqs[side] = n, [_rc_iupac(x) for x in q] if kind[side] == "R" else q
    seq, off = concat(sum((seqs[x] for x in sides), []))
    keys, counts = O.count(seq, off, 25, canonical)
    m = counts >= 2  # dump -L 2
    rows = {}
    for side in sides:
        s1, o1 = concat(seqs[side])
        st = O.read_stats(s1, o1, keys[m], counts[m], 25, canonical, 1, 1, 0, 0)
        rows[side] = sorted(((n, O.format_stats_row(n, x).split("\t")) for n, x in zip(names[side], st)),
                            key=lambda t: t[0].encode())  # sort -k1,1 in the C locale
    f = lambda rs, i: [float(r[1][i]) for r in rs]  # noqa: E731
    wanted = {}
    if mode == "pairs":
        L, R = rows["left"], rows["right"]
        assert [a[:-2] for a, _ in L] == [b[:-2] for b, _ in R]  # the sample is fully paired
        sel = O.select_pairs(f(L, 1), f(R, 1), f(L, 2), f(R, 2), f(L, 3), f(R, 3), max_cov, min_cov, max_cv)
        w = [a[:-2] for (a, _), x in zip(L, sel) if x]
        wanted = {"left": w, "right": w}
    else:
        for side in sides:
            S = rows[side]
            sel = O.select_pairs(f(S, 1), f(S, 1), f(S, 2), f(S, 2), f(S, 3), f(S, 3), max_cov, min_cov, max_cv)
            wanted[side] = [a[:-2] for (a, _), x in zip(S, sel) if x]
    out = {}
    for side in sides:
        recs = _fq_records(gzip.decompress((SAMPLE / f"reads.{side}.fq.gz").read_bytes()))
        ws = set(wanted[side])
        out[side] = b"".join(x for x in recs if x.split(b"\n")[0][1:].split()[0].decode()[:-2] in ws)
    return out, wanted


@pytest.mark.gpu
def test_normalization_chain_on_sample_reads(tmp_path):
    """insilico_read_normalization.pl's PE chain on the reference's sample
    reads, tool by tool (FASTQ->FASTA, jellyfish count + dump -L 2, per-read
    stats of each side, C-locale sort, pair merge + selection at max_cov 30,
    extraction of the selected FASTQ records), against the oracle's count,
    stats and selection."""
    d = tmp_path
    for side in ("left", "right"):
        with open(d / f"{side}.fa", "wb") as f:
            subprocess.run([BIN / "fastQ_to_fastA", "-I", SAMPLE / f"reads.{side}.fq.gz"], stdout=f, check=True)
    (d / "both.fa").write_bytes((d / "left.fa").read_bytes() + (d / "right.fa").read_bytes())
    run([BIN / "jellyfish", "count", "-t", "4", "-m", "25", "-s", "10000000", "--canonical", d / "both.fa"], cwd=d)
    dump = d / "jellyfish.K25.min2.kmers.fa"
    dump.write_text(run([BIN / "jellyfish", "dump", "-L", "2", d / "mer_counts.jf"]).stdout)
    for side in ("left", "right"):
        out = run([BIN / "fastaToKmerCoverageStats", "--reads", d / f"{side}.fa", "--kmers", dump, "--kmer_size",
                   "25", "--num_threads", "4", "--DS"]).stdout.splitlines(True)
        (d / f"{side}.stats.sort").write_text(out[0] + "".join(sorted(out[1:], key=lambda x: x.encode())))
    r = run([BIN / "nbkc_normalize", "--left", d / "left.stats.sort", "--right", d / "right.stats.sort",
             "--sorted", "--max_cov", "30", "--min_cov", "0", "--max_CV", "10000"])
    (d / "pairs.accs").write_text(r.stdout)
    for side in ("left", "right"):
        run([BIN / "extract_normalized_reads", "--seqType", "fq", "--selected", d / "pairs.accs", "--output",
             d / f"{side}.norm.fq", SAMPLE / f"reads.{side}.fq.gz"])
    want, wanted = _oracle_normalization("pairs", 30)
    assert r.stdout.split() == wanted["left"]
    assert 0 < len(wanted["left"]) < 30575
    for side in ("left", "right"):
        assert (d / f"{side}.norm.fq").read_bytes() == want[side]


# ------------------------------------------- insilico_read_normalization in one process

IRN = BIN / "insilico_read_normalization"


def _irn_args(mode: str, max_cov: int, ss: str = ""):
    args = ["--seqType", "fq", "--JM", "2G", "--CPU", "4", "--max_cov", max_cov, "--min_cov", "0",
            "--max_CV", "10000"]
    if mode == "single":
        args += ["--single", SAMPLE / "reads.left.fq.gz"]
    else:
        args += ["--left", SAMPLE / "reads.left.fq.gz", "--right", SAMPLE / "reads.right.fq.gz"]
    if mode == "pairs":
        args += ["--pairs_together", "--PARALLEL_STATS"]
    if ss:
        args += ["--SS_lib_type", ss]
    return [str(a) for a in args]


def _check_irn_outputs(out: Path, mode: str, max_cov: int, want):
    for side, exp in want.items():
        base = f"reads.{side}.fq.gz.normalized_K25_maxC{max_cov}_minC0_maxCV10000.fq"
        assert (out / base).read_bytes() == exp, (mode, side)
        link = out / f"{'single' if mode == 'single' else side}.norm.fq"
        assert link.is_symlink() and link.resolve() == (out / base).resolve()
        assert (out / (base + ".ok")).exists()
    assert not (out / "tmp_normalized_reads").exists()


def test_insilico_normalization_plan_matches_reference_merge(tmp_path):
    """The selection units of --pairs_together are the rows
    nbkc_merge_left_right_stats.pl --sorted wrote for the golden stats files
    (unpaired accessions on either side skipped), whatever the reads' input
    order; without --pairs_together each side is its own sorted list."""
    import random
    accs = {x: [ln.split("\t")[0] for ln in (NBKC / f"{x}.stats.sort").read_text().splitlines()[1:]]
            for x in ("left", "right")}
    rnd = random.Random(1)
    for side, a in accs.items():
        a = a[:]
        rnd.shuffle(a)
        (tmp_path / f"{side}.fa").write_text("".join(f">{x}\nACGTACGTAC\n" for x in a))
    base = [IRN, "--seqType", "fa", "--JM", "1G", "--max_cov", "30", "--left", tmp_path / "left.fa", "--right",
            tmp_path / "right.fa", "--plan_only"]
    got = run(base + ["--pairs_together"]).stdout
    rows = [ln.split("\t") for ln in (NBKC / "pairs.stats").read_text().splitlines()[1:]]
    assert got == "# selection 0\n" + "".join(f"{r[1]}\t{r[5]}\n" for r in rows)
    got = run(base).stdout.split("# selection 1\n")
    assert got[0] == "# selection 0\n" + "".join(x + "\n" for x in sorted(accs["left"], key=str.encode))
    assert got[1] == "".join(x + "\n" for x in sorted(accs["right"], key=str.encode))


def test_insilico_normalization_service_protocol(tmp_path):
    """The service's socket protocol without device work: concurrent
    --plan_only jobs come back with the plan the one-process run prints,
    option errors come back as the job's failure, --shutdown ends the
    service and removes its socket."""
    import threading
    import time
    (tmp_path / "l.fa").write_text("".join(f">r{i}/1\nACGTACGTAC\n" for i in (3, 1, 2)))
    (tmp_path / "r.fa").write_text("".join(f">r{i}/2\nACGTACGTAC\n" for i in (2, 3, 9)))
    job = ["--seqType", "fa", "--JM", "1G", "--max_cov", "30", "--left", "l.fa", "--right", "r.fa",
           "--pairs_together", "--plan_only"]
    want = run([IRN, *job], cwd=tmp_path).stdout
    assert want == "# selection 0\nr2/1\tr2/2\nr3/1\tr3/2\n"
    sock = tmp_path / "s.sock"
    srv = subprocess.Popen([str(IRN), "--serve", str(sock), "--slots", "1", "--hbm_budget_gb", "1"],
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    try:
        t0 = time.time()
        while not sock.exists():
            assert srv.poll() is None, srv.communicate()[1]
            assert time.time() - t0 < 60
            time.sleep(0.1)
        got = []
        th = [threading.Thread(target=lambda: got.append(run([IRN, "--server", sock, *job], cwd=tmp_path,
                                                             timeout=60))) for _ in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert len(got) == 3 and all(r.returncode == 0 and want in r.stderr for r in got)
        bad = run([IRN, "--server", sock, "--seqType", "fx", "--JM", "1G", "--max_cov", "30", "--single", "l.fa"],
                  cwd=tmp_path, check=False, timeout=60)
        assert bad.returncode == 1 and "set --seqType" in bad.stderr
        run([IRN, "--server", sock, "--shutdown"], timeout=60)
        assert srv.wait(timeout=60) == 0
        assert not sock.exists()
    finally:
        if srv.poll() is None:
            srv.kill()
            srv.wait()


@pytest.mark.parametrize("args,msg", [
    (["--seqType", "fx", "--JM", "1G", "--max_cov", "30", "--single", "x.fa"], "set --seqType to 'fq' or 'fa'"),
    (["--seqType", "fa", "--JM", "1", "--max_cov", "30", "--single", "x.fa"], "must set --JM"),
    (["--seqType", "fa", "--JM", "1.5Gb", "--max_cov", "30", "--single", "x.fa"], "must set --JM"),
    (["--seqType", "fa", "--JM", "10Gb", "--max_cov", "30", "--single", "x.fa"], "cannot parse max_memory"),
    (["--seqType", "fa", "--JM", "1G", "--max_cov", "1", "--single", "x.fa"], "--max_cov at least 2"),
    (["--seqType", "fa", "--JM", "1G", "--max_cov", "9", "--left", "x.fa"], "need either options 'left' and"),
    (["--seqType", "fa", "--JM", "1G", "--max_cov", "9", "--single", "x.fa", "--pairs_together"],
     "must use the --left and --right"),
    (["--seqType", "fa", "--JM", "1G", "--max_cov", "9", "--single", "x.fa", "--SS_lib_type", "RR"],
     "unrecognized SS_lib_type"),
    (["--seqType", "fa", "--JM", "1G", "--max_cov", "9", "--single", "x.fa", "--bogus"], "do not understand"),
    (["--server", "/nonexistent/norm.sock", "--seqType", "fa", "--JM", "1G", "--max_cov", "9", "--single", "x.fa"],
     "no normalization service"),
])
def test_insilico_normalization_option_errors(tmp_path, args, msg):
    """The script's option checks (:184-256), before any device work; an
    explicit --server that nobody serves is an error, not a local run."""
    r = run([IRN, *args], cwd=tmp_path, check=False)
    assert r.returncode != 0 and msg in r.stderr, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode,ss,max_cov", [("pairs", "", 30), ("separate", "", 20), ("single", "", 20),
                                             ("pairs", "RF", 30)])
def test_insilico_normalization_cli(tmp_path, mode, ss, max_cov):
    """The whole per-sample job in one process (count, stats and selection on
    the device through the normalization service) writes the normalized FASTQ
    the oracle's step-by-step chain gives: pairs together, each side on its
    own, single-end, and a strand-specific RF library (left reads
    reverse-complemented, k-mers not canonical); outputs named and linked as
    the Perl script does, its tmp directory removed."""
    want, wanted = _oracle_normalization(mode, max_cov, 0, 10000, ss)
    assert all(0 < len(w) < 30575 for w in wanted.values())
    r = run([IRN, *_irn_args(mode, max_cov, ss), "--output", tmp_path / "norm"])
    assert "Normalization complete" in r.stderr
    _check_irn_outputs(tmp_path / "norm", mode, max_cov, want)


@pytest.mark.gpu
def test_insilico_normalization_service(tmp_path):
    """The resident service: jobs sent by concurrent clients (relative
    --output resolved in each client's directory) give the outputs of the
    one-process runs; a job with bad options fails alone; --shutdown stops
    the service."""
    import threading
    import time
    sock = tmp_path / "norm.sock"
    srv = subprocess.Popen([str(IRN), "--serve", str(sock), "--slots", "2", "--hbm_budget_gb", "16"],
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    try:
        t0 = time.time()
        while not sock.exists():
            assert srv.poll() is None, srv.communicate()[1]
            assert time.time() - t0 < 120, "service did not start"
            time.sleep(0.2)
        cases = [("pairs", 30), ("separate", 20), ("single", 20)]
        results = {}

        def client(mode, max_cov):
            d = tmp_path / mode
            d.mkdir()
            results[mode] = run([IRN, "--server", sock, *_irn_args(mode, max_cov), "--output", "out"], cwd=d,
                                check=False, timeout=600)

        th = [threading.Thread(target=client, args=c) for c in cases]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for mode, max_cov in cases:
            r = results[mode]
            assert r.returncode == 0, r.stderr
            assert "Normalization complete" in r.stderr
            want, _ = _oracle_normalization(mode, max_cov)
            _check_irn_outputs(tmp_path / mode / "out", mode, max_cov, want)
        bad = run([IRN, "--server", sock, "--seqType", "fq", "--JM", "1G", "--max_cov", "9", "--single",
                   tmp_path / "missing.fq", "--output", tmp_path / "bad"], check=False, timeout=120)
        assert bad.returncode != 0 and "missing.fq" in bad.stderr
        # a job the device part refuses (max_CV below the bound where the CV
        # test is exact) fails alone too
        bad = run([IRN, "--server", sock, *_irn_args("single", 20), "--max_CV", "1", "--output", tmp_path / "bad2"],
                  check=False, timeout=300)
        assert bad.returncode != 0 and "max_cv" in bad.stderr, bad.stderr
        assert srv.poll() is None
        run([IRN, "--server", sock, "--shutdown"], timeout=120)
        assert srv.wait(timeout=120) == 0
        assert not sock.exists()
    finally:
        if srv.poll() is None:
            srv.kill()
            srv.wait()


# ------------------------------------------- Chrysalis: welds, de Bruijn graphs, node weights

def _debruijn_expected(records, k, ss=False):
    out = []
    for name, s in records:
        out.append(f"Component {name[2:] if name.startswith('s_') else name}\n")
        ids, order, preds = {}, [], {}
        for part in "".join(ch if ch.upper() in "ACGT" else " " for ch in s).split():
            part = part.upper()
            prev = None
            for i in range(len(part) - k + 1):
                w = part[i:i + k]
                key = w if ss else min(w, brute.rc(w))
                if key not in ids:
                    ids[key] = len(order)
                    order.append(key)
                    preds[key] = set()
                if prev is not None:
                    preds[key].add(ids[prev])
                prev = key
        for key in order:
            p = ",".join(str(x) for x in sorted(preds[key])) or "-1"
            out.append(f"{ids[key]}\t{key}\t{p}\n")
    return "".join(out)


@pytest.mark.parametrize("ss", [False, True])
def test_fasta_to_debruijn(tmp_path, ss):
    import random
    rnd = random.Random(3)
    recs = []
    for c in range(6):
        seqs = ["".join(rnd.choice("ACGT") for _ in range(rnd.randint(30, 300))) for _ in range(rnd.randint(1, 4))]
        if c == 2:
            seqs.append(seqs[0][10:80])  # repeated k-mers: shared nodes, several predecessors
        recs.append((f"s_{c * 7}", "X".join(seqs)))
    recs.append(("s_99", "ACGTN" * 3))  # no k-mer
    write_fasta(tmp_path / "b.fa", [n for n, _ in recs], [s for _, s in recs])
    args = [BIN / "FastaToDeBruijn", "--fasta", tmp_path / "b.fa", "-K", "24", "--graph_per_record", "--threads", "2"]
    if ss:
        args.append("--SS")
    assert run(args).stdout == _debruijn_expected(recs, 24, ss)


def _weld_expected(contigs, reads, k, kk, min_glue, strand):
    flank = (kk - k) // 2
    canon = (lambda s: s) if strand else (lambda s: min(s, brute.rc(s)))
    occ = {}
    for i, c in enumerate(contigs):
        for p in range(len(c) - k + 1):
            w = c[p:p + k]
            if strand or w <= brute.rc(w):
                occ.setdefault(w, []).append((i, c, p))
            else:
                rc = brute.rc(c)
                occ.setdefault(brute.rc(w), []).append((i, rc, len(c) - p - k))
    cands = []
    for w, v in occ.items():
        for (ia, sa, pa) in v:
            for (ib, sb, pb) in v:
                if ia == ib or pa < flank or pb + k + flank > len(sb):
                    continue
                weld = sa[pa - flank:pa] + w + sb[pb + k:pb + k + flank]
                cands.append((min(ia, ib), max(ia, ib), canon(weld), w))
    welds = sorted({c[2] for c in cands})
    counts = brute.kset_count(welds, reads, kk, not strand)
    sup = dict(zip(welds, counts))
    best = {}
    for a, b, weld, w in cands:
        s = sup[weld]
        if s < min_glue:
            continue
        cur = best.get((a, b))
        if cur is None or (s, [-ord(ch) for ch in weld], [-ord(ch) for ch in w]) > \
                (cur[0], [-ord(ch) for ch in cur[1]], [-ord(ch) for ch in cur[2]]):
            best[(a, b)] = (s, weld, w)
    return "".join(f"{a} -> {b} weld: {best[(a, b)][1]} k-mer: {best[(a, b)][2]} support: {best[(a, b)][0]}\n"
                   for a, b in sorted(best))


@pytest.mark.gpu
@pytest.mark.parametrize("strand", [False, True])
def test_graph_from_fasta_weld_support(tmp_path, strand):
    """Contigs cut from transcripts with overlaps of 30-60 bases (so pairs
    share 24-mers), some reverse-complemented; reads tiled over the
    transcripts: the welded pairs and their best weld k-mer equal the
    restatement's, with the weld counts taken from the reads."""
    import random
    rnd = random.Random(11 + strand)
    contigs, reads = [], []
    for t in range(12):
        tx = "".join(rnd.choice("ACGT") for _ in range(rnd.randint(400, 900)))
        cut = rnd.randint(150, len(tx) - 150)
        ov = rnd.randint(30, 60)
        a, b = tx[:cut + ov], tx[cut:]
        if t % 3 == 0 and not strand:
            b = brute.rc(b)
        contigs += [a, b]
        for i in range(0, len(tx) - 100, rnd.randint(3, 9)):
            r = tx[i:i + 100]
            reads.append(brute.rc(r) if (i % 2 and not strand) else r)
    contigs.append("ACGT" * 10)
    write_fasta(tmp_path / "iw.fa", [f"a{i + 1};10" for i in range(len(contigs))], contigs)
    write_fasta(tmp_path / "r.fa", [f"r{i}" for i in range(len(reads))], reads)
    args = [BIN / "GraphFromFasta", "-i", tmp_path / "iw.fa", "-r", tmp_path / "r.fa", "-min_contig_length", "1",
            "-min_glue", "2", "-glue_factor", "0.05", "-min_iso_ratio", "0.05", "-t", "4", "-k", "24", "-kk", "48"]
    if strand:
        args.append("-strand")
    out = run(args).stdout
    assert out == _weld_expected(contigs, reads, 24, 48, 2, strand)
    assert len(out.splitlines()) >= 10


@pytest.mark.gpu
def test_quantify_graph_single_and_batched(tmp_path):
    """Node weights of each component = the windows of its reads equal to the
    node (canonical), one component per call and all components of a listing
    in one call; both equal the brute-force count."""
    import random
    rnd = random.Random(5)
    recs, comp_reads = [], {}
    for c in (3, 8, 21):
        tx = "".join(rnd.choice("ACGT") for _ in range(rnd.randint(300, 700)))
        recs.append((f"s_{c}", tx[:200] + "X" + tx[150:]))
        rs = []
        for i in range(0, len(tx) - 76, 4):
            r = tx[i:i + 76]
            rs.append(brute.rc(r) if i % 8 else r)
        comp_reads[c] = rs
    write_fasta(tmp_path / "b.fa", [n for n, _ in recs], [s for _, s in recs])
    graph = run([BIN / "FastaToDeBruijn", "--fasta", tmp_path / "b.fa", "-K", "24", "--graph_per_record"]).stdout
    blocks = graph.split("Component ")[1:]
    listing = []
    for blk in blocks:
        cid = int(blk.split("\n", 1)[0])
        base = tmp_path / f"c{cid}"
        (tmp_path / f"c{cid}.graph.tmp").write_text("Component " + blk)
        (tmp_path / f"c{cid}.reads.tmp").write_text("".join(f">r{j} 100\n{r}\n" for j, r in enumerate(comp_reads[cid])))
        listing.append(f"{cid}\t{base}\n")
    (tmp_path / "listing.txt").write_text("".join(listing))

    def expected(cid):
        lines = (tmp_path / f"c{cid}.graph.tmp").read_text().splitlines()
        kmers = [ln.split("\t")[1] for ln in lines[1:]]
        cnt = brute.kset_count(kmers, comp_reads[cid], 24, True)
        return lines[0] + "\n" + "".join(f"{ln}\t{n}\n" for ln, n in zip(lines[1:], cnt))

    for cid in (3, 8, 21):
        run([BIN / "QuantifyGraph", "-g", tmp_path / f"c{cid}.graph.tmp", "-i", tmp_path / f"c{cid}.reads.tmp",
             "-o", tmp_path / f"c{cid}.graph.out", "-max_reads", "1000000", "-k", "24"])
        assert (tmp_path / f"c{cid}.graph.out").read_text() == expected(cid)
        (tmp_path / f"c{cid}.graph.out").unlink()
    run([BIN / "QuantifyGraph", "--batch", tmp_path / "listing.txt", "-k", "24"])
    for cid in (3, 8, 21):
        assert (tmp_path / f"c{cid}.graph.out").read_text() == expected(cid)
