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
