"""Brute-force, string-based restatements used to pin the C oracle.

Independent of the oracle's 2-bit arithmetic: k-mers are Python strings,
reverse complement is str.translate, canonical is min() on strings (the
lexicographic order of ACGT equals the 2-bit numeric order).  Slow; for
small inputs only.
"""
from __future__ import annotations

import math
from collections import Counter

COMP = str.maketrans("ACGTacgt", "TGCAtgca")


def rc(s: str) -> str:
    return s.translate(COMP)[::-1]


def key_of(s: str) -> int:
    v = 0
    for ch in s:
        v = v * 4 + "ACGT".index(ch)
    return v


def kmer_of(v: int, k: int) -> str:
    return "".join("ACGT"[(v >> (2 * (k - 1 - i))) & 3] for i in range(k))


def windows(seq: str, k: int):
    """Yield (kmer_upper or None) for each window; None when a non-ACGT base is inside."""
    up = seq.upper()
    for i in range(len(up) - k + 1):
        w = up[i:i + k]
        yield w if all(c in "ACGT" for c in w) else None


def canon(w: str, canonical: bool) -> str:
    return min(w, rc(w)) if canonical else w


def count(seqs, k: int, canonical: bool = True) -> Counter:
    c = Counter()
    for s in seqs:
        for w in windows(s, k):
            if w is not None:
                c[canon(w, canonical)] += 1
    return c


def histo(counts, low=1, high=10000, inc=1):
    base = (1 if inc >= low else low - inc) if low > 1 else 1
    ceil = high + inc
    nb = (ceil + inc - base) // inc
    bins = [0] * nb
    for v in counts:
        if v < base:
            bins[0] += 1
        elif v >= ceil:
            bins[-1] += 1
        else:
            bins[(v - base) // inc] += 1
    return bins


def entropy(w: str) -> float:
    n = len(w)
    h = 0.0
    for b, c in Counter(w).items():
        p = c / n
        h -= p * math.log2(p)
    return h


def read_stats(seq: str, dict_counts: dict, k: int, canonical=True, min_count=1, missing=1,
               nwin_policy=0, median_policy=0):
    vals = []
    for w in windows(seq, k):
        if w is None:
            if nwin_policy == 0:
                vals.append(missing)
            continue
        c = dict_counts.get(canon(w, canonical), 0)
        vals.append(c if (c and c >= min_count) else missing)
    n = len(vals)
    if n == 0:
        return dict(median=0, nwin=0, sum=0, sumsq=0)
    s = sorted(vals)
    if median_policy == 1:
        med = s[(n - 1) // 2]
    elif median_policy == 2 and n % 2 == 0:
        med = (s[n // 2 - 1] + s[n // 2]) // 2
    else:
        med = s[n // 2]
    return dict(median=med, nwin=n, sum=sum(vals), sumsq=sum(v * v for v in vals))


class Drand48:
    """drand48 as Perl's rand(): x <- (0x5DEECE66D*x + 11) mod 2^48, srand(s): x = s<<16 | 0x330E."""

    def __init__(self, seed: int):
        self.x = ((seed << 16) | 0x330E) & ((1 << 48) - 1)

    def __call__(self) -> float:
        self.x = (0x5DEECE66D * self.x + 0xB) & ((1 << 48) - 1)
        return self.x / float(1 << 48)


def perl_select(rows, max_cov, min_cov=1, max_cv=10000):
    """nbkc_merge_left_right_stats.pl + nbkc_normalize.pl on already-paired rows.

    rows: list of (l_med, l_mean, l_sd, r_med, r_mean, r_sd) as the text the
    stats tool printed (strings).  Returns list of selected indices."""
    rnd = Drand48(12345)
    out = []
    for i, (lm, lu, ls, rm, ru, rs) in enumerate(rows):
        med = float("%.1f" % ((float(lm) + float(rm)) / 2))
        mean = float("%.1f" % ((float(lu) + float(ru)) / 2))
        sd = float("%.1f" % ((float(ls) + float(rs)) / 2))
        if med < min_cov:
            continue
        if mean <= 0:
            continue
        if sd / mean > max_cv:
            continue
        if rnd() <= max_cov / med:
            out.append(i)
    return out


def r2t(bundles, comps, reads, k: int, canonical=True):
    table = {}
    for seq, c in zip(bundles, comps):
        for w in windows(seq, k):
            if w is None:
                continue
            key = canon(w, canonical)
            table[key] = min(c, table.get(key, c))
    res = []
    for s in reads:
        hits = Counter()
        nwin = max(0, len(s) - k + 1)
        for w in windows(s, k):
            if w is None:
                continue
            c = table.get(canon(w, canonical))
            if c is not None:
                hits[c] += 1
        if not hits:
            res.append((0xFFFFFFFF, 0, nwin))
        else:
            best = min(hits.items(), key=lambda kv: (-kv[1], kv[0]))
            res.append((best[0], best[1], nwin))
    return res


def inchworm(d: dict, k: int, ds: bool, min_seed_cov: int, min_seed_entropy: float | None = None):
    """Greedy contigs over a string dictionary {kmer: count} (canonical keys
    when ds): seeds by count descending then string, each extended right then
    left by the most abundant unused overlapping k-mer (ties: A<C<G<T)."""
    canon = (lambda s: min(s, rc(s))) if ds else (lambda s: s)
    used = set()
    seeds = sorted((s for s, c in d.items()
                    if c >= min_seed_cov and c > 0 and (min_seed_entropy is None or entropy(s) >= min_seed_entropy)),
                   key=lambda s: (-d[s], s))
    out = []
    for s in seeds:
        if s in used:
            continue
        used.add(s)
        contig, total = s, d[s]
        for side in ("right", "left"):
            while True:
                best = None
                for b in "ACGT":
                    y = canon(contig[-(k - 1):] + b if side == "right" else b + contig[:k - 1])
                    c = d.get(y, 0)
                    if c > 0 and y not in used and (best is None or c > best[0]):
                        best = (c, b, y)
                if best is None:
                    break
                used.add(best[2])
                total += best[0]
                contig = contig + best[1] if side == "right" else best[1] + contig
        out.append((contig, total, s))
    return out


def kmer_value(s: str) -> int:
    v = 0
    for ch in s:
        v = v * 4 + "ACGT".index(ch)
    return v


def kset_count(keys: list, reads: list, kk: int, canonical: bool, tags=None) -> list:
    """Occurrences of each key string (or (tag, kmer string)) among the
    reads' all-ACGT windows, by string slicing."""
    index = {k: i for i, k in enumerate(keys)}
    out = [0] * len(keys)
    for r, s in enumerate(reads):
        u = s.upper()
        for i in range(len(u) - kk + 1):
            w = u[i:i + kk]
            if any(ch not in "ACGT" for ch in w):
                continue
            if canonical:
                w = min(w, rc(w))
            key = (tags[r], w) if tags is not None else w
            j = index.get(key)
            if j is not None:
                out[j] += 1
    return out
