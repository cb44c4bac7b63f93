"""Seeded synthetic transcriptomes and reads for the workloads of the k-mer path.

Workloads (sizes as named by the benchmark):
  small_se  : 500 transcripts, lengths U[500,3000], expression lognormal(2,1.5),
              76-nt single-end reads sampled by length x expression, 50%
              reverse-complemented, 0.5% substitutions (seed 1001).
  large_pe  : 50k transcripts, lengths lognormal (median 1.2 kb, sigma 0.6)
              clipped to [200,15000], 10% of transcripts share a 100-500 nt
              segment with a partner, expression lognormal(2,1.5); 2x150
              paired-end reads from fragments N(300,50), 1% substitutions,
              0.05% N (seed 1002).
Generation is vectorised numpy in chunks; every array is a pure function of
the seed and the chunk index, so a prefix of the reads can be regenerated
without the rest.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

ASCII = np.frombuffer(b"ACGTN", dtype=np.uint8)


@dataclass
class Transcriptome:
    codes: np.ndarray     # uint8 0..3, all transcripts concatenated
    starts: np.ndarray    # int64 [n]
    lengths: np.ndarray   # int64 [n]
    weights: np.ndarray   # float64 [n] expression

    @property
    def n(self) -> int:
        return len(self.lengths)

    def seq(self, i: int) -> str:
        s = self.codes[self.starts[i]: self.starts[i] + self.lengths[i]]
        return ASCII[s].tobytes().decode()


def make_transcriptome(seed: int, n: int, lengths: np.ndarray, expr_mu: float = 2.0,
                       expr_sigma: float = 1.5, share_frac: float = 0.0,
                       share_len=(100, 500)) -> Transcriptome:
    rng = np.random.default_rng([seed, 0])
    lengths = np.asarray(lengths, dtype=np.int64)
    starts = np.zeros(n, dtype=np.int64)
    starts[1:] = np.cumsum(lengths)[:-1]
    codes = rng.integers(0, 4, size=int(lengths.sum()), dtype=np.uint8)
    if share_frac > 0:
        m = int(n * share_frac) // 2 * 2
        who = rng.permutation(n)[:m].reshape(-1, 2)
        for a, b in who:
            L = int(rng.integers(share_len[0], share_len[1] + 1))
            L = min(L, int(lengths[a]), int(lengths[b]))
            sa = int(rng.integers(0, lengths[a] - L + 1))
            sb = int(rng.integers(0, lengths[b] - L + 1))
            codes[starts[b] + sb: starts[b] + sb + L] = codes[starts[a] + sa: starts[a] + sa + L]
    weights = rng.lognormal(expr_mu, expr_sigma, size=n)
    return Transcriptome(codes, starts, lengths, weights)


def small_transcriptome(seed: int = 1001, n: int = 500) -> Transcriptome:
    rng = np.random.default_rng([seed, 1])
    return make_transcriptome(seed, n, rng.integers(500, 3001, size=n))


def large_transcriptome(seed: int = 1002, n: int = 50_000) -> Transcriptome:
    rng = np.random.default_rng([seed, 1])
    L = np.clip(np.round(rng.lognormal(np.log(1200.0), 0.6, size=n)), 200, 15000)
    return make_transcriptome(seed, n, L.astype(np.int64), share_frac=0.10)


def _rc(x: np.ndarray) -> np.ndarray:
    return (3 - x[..., ::-1]).astype(np.uint8)


def _mutate(rng, reads: np.ndarray, sub_rate: float, n_rate: float) -> np.ndarray:
    flat = reads.reshape(-1)
    ne = rng.binomial(flat.size, sub_rate)
    if ne:
        pos = rng.integers(0, flat.size, size=ne)
        flat[pos] = (flat[pos] + rng.integers(1, 4, size=ne, dtype=np.uint8)) % 4
    nn = rng.binomial(flat.size, n_rate) if n_rate > 0 else 0
    if nn:
        flat[rng.integers(0, flat.size, size=nn)] = 4
    return reads


def _pick(rng, tx: Transcriptome, n: int) -> np.ndarray:
    p = tx.weights * tx.lengths
    cdf = np.cumsum(p)
    cdf /= cdf[-1]
    return np.minimum(np.searchsorted(cdf, rng.random(n), side="right"), tx.n - 1)


def se_reads(tx: Transcriptome, n_reads: int, read_len: int = 76, seed: int = 1001,
             sub_rate: float = 0.005, n_rate: float = 0.0, chunk: int = 1 << 20):
    """Single-end reads -> (ascii uint8 [n*read_len], uint64 offsets[n+1])."""
    out = np.empty(n_reads * read_len, dtype=np.uint8)
    for c0 in range(0, n_reads, chunk):
        m = min(chunk, n_reads - c0)
        rng = np.random.default_rng([seed, 2, c0 // chunk])
        t = _pick(rng, tx, m)
        s = (rng.random(m) * (tx.lengths[t] - read_len + 1)).astype(np.int64)
        idx = (tx.starts[t] + s)[:, None] + np.arange(read_len)[None, :]
        r = tx.codes[idx]
        flip = rng.random(m) < 0.5
        r[flip] = _rc(r[flip])
        _mutate(rng, r, sub_rate, n_rate)
        out[c0 * read_len:(c0 + m) * read_len] = ASCII[r].reshape(-1)
    off = np.arange(n_reads + 1, dtype=np.uint64) * np.uint64(read_len)
    return out, off


def pe_reads(tx: Transcriptome, n_pairs: int, read_len: int = 150, seed: int = 1002,
             frag_mu: float = 300.0, frag_sd: float = 50.0, sub_rate: float = 0.01,
             n_rate: float = 0.0005, chunk: int = 1 << 19, first_pair: int = 0):
    """Paired-end reads -> (left ascii, right ascii, offsets) with pair i = left i, right i.

    first_pair lets a caller regenerate only pairs [first_pair, first_pair+n_pairs)
    (first_pair must be a multiple of chunk)."""
    assert first_pair % chunk == 0
    left = np.empty(n_pairs * read_len, dtype=np.uint8)
    right = np.empty(n_pairs * read_len, dtype=np.uint8)
    ar = np.arange(read_len)[None, :]
    for c0 in range(0, n_pairs, chunk):
        m = min(chunk, n_pairs - c0)
        rng = np.random.default_rng([seed, 3, (first_pair + c0) // chunk])
        t = _pick(rng, tx, m)
        L = tx.lengths[t]
        f = np.clip(np.round(rng.normal(frag_mu, frag_sd, size=m)), read_len, None).astype(np.int64)
        f = np.minimum(f, L)
        s = (rng.random(m) * (L - f + 1)).astype(np.int64)
        a = tx.codes[(tx.starts[t] + s)[:, None] + ar]
        b = _rc(tx.codes[(tx.starts[t] + s + f - read_len)[:, None] + ar])
        flip = rng.random(m) < 0.5
        l_ = np.where(flip[:, None], b, a)
        r_ = np.where(flip[:, None], a, b)
        _mutate(rng, l_, sub_rate, n_rate)
        _mutate(rng, r_, sub_rate, n_rate)
        left[c0 * read_len:(c0 + m) * read_len] = ASCII[l_].reshape(-1)
        right[c0 * read_len:(c0 + m) * read_len] = ASCII[r_].reshape(-1)
    off = np.arange(n_pairs + 1, dtype=np.uint64) * np.uint64(read_len)
    return left, right, off


def read_names(n: int, suffix: str = "", prefix: str = "r", start: int = 0) -> list[str]:
    return [f"{prefix}{i:09d}{suffix}" for i in range(start, start + n)]


def write_fasta(path, names, seq: np.ndarray, off: np.ndarray, width: int = 0) -> None:
    """Write reads as FASTA (one sequence line per record unless width > 0)."""
    buf = seq.tobytes() if isinstance(seq, np.ndarray) else seq
    with open(path, "w") as fh:
        for i, nm in enumerate(names):
            s = buf[int(off[i]):int(off[i + 1])].decode()
            if width:
                s = "\n".join(s[j:j + width] for j in range(0, len(s), width)) or ""
            fh.write(f">{nm}\n{s}\n")
