"""Small seeded inputs with the edge cases the reference tools meet."""
from __future__ import annotations

import random

import numpy as np


def random_reads(seed: int, n: int, lo: int = 0, hi: int = 120, n_rate: float = 0.01,
                 lower_rate: float = 0.02, other: str = "NXRY.-") -> list[str]:
    rnd = random.Random(seed)
    out = []
    for _ in range(n):
        L = rnd.randint(lo, hi)
        s = []
        for _ in range(L):
            r = rnd.random()
            if r < n_rate:
                s.append(rnd.choice(other))
            else:
                b = rnd.choice("ACGT")
                s.append(b.lower() if rnd.random() < lower_rate else b)
        out.append("".join(s))
    return out


def repetitive_reads(seed: int, n: int, genome_len: int = 400, read_len: int = 60,
                     err: float = 0.01) -> list[str]:
    """Reads tiled from a short random genome so k-mers recur many times."""
    rnd = random.Random(seed)
    g = "".join(rnd.choice("ACGT") for _ in range(genome_len))
    comp = str.maketrans("ACGT", "TGCA")
    out = []
    for _ in range(n):
        s = rnd.randint(0, genome_len - read_len)
        r = list(g[s:s + read_len])
        for i in range(read_len):
            if rnd.random() < err:
                r[i] = rnd.choice("ACGT")
        r = "".join(r)
        if rnd.random() < 0.5:
            r = r.translate(comp)[::-1]
        out.append(r)
    return out


def edge_reads() -> list[str]:
    return [
        "",                                   # empty
        "ACGT",                               # shorter than k
        "A" * 25,                             # exactly one window, low entropy
        "N" * 40,                             # no valid window
        "ACGTACGTACGTACGTACGTACGTA",          # k long
        "acgtacgtacgtacgtacgtacgtacgtNacgtacgtacgtacgtacgtacgtac",  # lowercase + N in middle
        "GATTACA" * 20,
        "TTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTT",
        "CCCCCGGGGGAAAAATTTTTCCCCCGGGGGAAAAATTTTT",
    ]


def concat(seqs: list[str]) -> tuple[bytes, np.ndarray]:
    bs = [s.encode() for s in seqs]
    off = np.zeros(len(bs) + 1, dtype=np.uint64)
    if bs:
        off[1:] = np.cumsum([len(b) for b in bs])
    return b"".join(bs), off


GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"
NBKC = GOLDEN / "nbkc"
NBKC_SETTINGS = [(50, 0, 10000), (50, 1, 10000), (200, 0, 10000), (5, 3, 10000), (50, 0, 1), (30, 10, 2)]


def nbkc_aggregates(side: str, dtype) -> tuple[list[str], np.ndarray]:
    """Accessions and integer stats behind tests/golden/nbkc/<side>.stats.sort."""
    lines = (NBKC / f"{side}.aggregates.tsv").read_text().splitlines()[1:]
    accs = []
    st = np.zeros(len(lines), dtype)
    for i, ln in enumerate(lines):
        f = ln.split("\t")
        accs.append(f[0])
        st[i] = (int(f[1]), int(f[2]), int(f[3]), int(f[4]), int(f[5]))
    return accs, st


def nbkc_pairs() -> list[list[str]]:
    """Rows of the reference merge's output (pairs.stats), header dropped."""
    return [ln.split("\t") for ln in (NBKC / "pairs.stats").read_text().splitlines()[1:]]


def nbkc_selected(prefix: str, max_cov: int, min_cov: int, max_cv: int) -> list[str]:
    return (NBKC / f"{prefix}.C{max_cov}.minC{min_cov}.maxCV{max_cv}.accs").read_text().split()


def nbkc_stderr(prefix: str, max_cov: int, min_cov: int, max_cv: int) -> str:
    return (NBKC / f"{prefix}.C{max_cov}.minC{min_cov}.maxCV{max_cv}.stderr").read_text()


def histo_witness(name: str) -> list[tuple[int, int]]:
    """(label, n_kmers) lines of a reference jellyfish-histo witness file."""
    out = []
    for ln in (GOLDEN / "histo" / f"kmer_histo.{name}.txt").read_text().splitlines():
        a, b = ln.split()
        out.append((int(a), int(b)))
    return out


def counts_for_histo_witness(name: str, seed: int = 3) -> tuple[np.ndarray, list[tuple[int, int]]]:
    """A k-mer count multiset whose default histo has exactly the witness's
    labels: n' = 1 + n % 5 k-mers per label (label 10001 taking counts above
    10000, up to 2^32 - 1).  Returns the counts and the expected lines."""
    rng = np.random.default_rng(seed)
    vals, want = [], []
    for lab, n in histo_witness(name):
        m = 1 + n % 5
        if lab == 10001:
            vals += [int(x) for x in rng.choice([10001, 10002, 65535, 1 << 20, 2**32 - 1], m)]
        else:
            vals += [lab] * m
        want.append((lab, m))
    counts = np.array(vals, np.uint32)
    rng.shuffle(counts)
    return counts, want


def sample_fasta(side: str) -> tuple[list[str], list[str]]:
    """Names and sequences of the reference's sample reads (FASTA form)."""
    import gzip
    txt = gzip.decompress((GOLDEN / "sample_data" / f"reads.{side}.fa.gz").read_bytes()).decode()
    lines = txt.split("\n")
    return [ln[1:] for ln in lines[0::2] if ln], [ln for ln in lines[1::2]][:len(lines) // 2]


def ref_seqs() -> list[str]:
    """The 11 sample transcripts (refSeqs.fa), one sequence each."""
    out, cur = [], []
    for ln in (GOLDEN / "sample_data" / "refSeqs.fa").read_text().splitlines():
        if ln.startswith(">"):
            if cur:
                out.append("".join(cur))
            cur = []
        else:
            cur.append(ln.strip())
    if cur:
        out.append("".join(cur))
    return out
