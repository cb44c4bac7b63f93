"""Seeded synthetic transcriptomes and reads (wrapper of lib/libtk_synth.so).

Workloads:
  small_se : 500 transcripts, lengths U[500,3000], expression lognormal(2,1.5),
             76-nt single-end reads sampled by length x expression, 50%
             reverse-complemented, 0.5% substitutions (seed 1001).
  large_pe : 50k transcripts, lengths lognormal (median 1.2 kb, sigma 0.6)
             clipped to [200,15000], 10% of transcripts share a 100-500 nt
             segment with a partner, expression lognormal(2,1.5); 2x150
             paired-end reads from fragments N(300,50), 1% substitutions,
             0.05% N (seed 1002).
Every read is a function of (seed, index): any slice regenerates alone.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

SO = Path(__file__).resolve().parent / "lib" / "libtk_synth.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not SO.exists():
            raise RuntimeError(f"{SO} missing: run __graft_entry__.build()")
        l = C.CDLL(str(SO))
        l.tks_transcriptome.restype = C.c_void_p
        l.tks_transcriptome.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_double, C.c_double]
        l.tks_tx_info.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        l.tks_tx_seq.restype = C.c_uint64
        l.tks_tx_seq.argtypes = [C.c_void_p, C.c_uint64, C.c_char_p]
        l.tks_tx_length.restype = C.c_uint64
        l.tks_tx_length.argtypes = [C.c_void_p, C.c_uint64]
        l.tks_tx_free.argtypes = [C.c_void_p]
        l.tks_reads.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_uint64, C.c_uint64, C.c_int,
                                C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                C.c_void_p, C.c_void_p]
        _lib = l
    return _lib


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


class Transcriptome:
    def __init__(self, seed: int, n: int, len_model: int, p1: float, p2: float, lo: float = 0,
                 hi: float = 0, share_frac: float = 0.0, share_lo: int = 100, share_hi: int = 500,
                 expr_mu: float = 2.0, expr_sigma: float = 1.5):
        self.h = C.c_void_p(lib().tks_transcriptome(seed, n, len_model, p1, p2, lo, hi, share_frac,
                                                    share_lo, share_hi, expr_mu, expr_sigma))
        a, b = C.c_uint64(), C.c_uint64()
        lib().tks_tx_info(self.h, C.byref(a), C.byref(b))
        self.n, self.total = a.value, b.value

    def length(self, i: int) -> int:
        return lib().tks_tx_length(self.h, i)

    def seq(self, i: int) -> str:
        buf = C.create_string_buffer(self.length(i) + 1)
        L = lib().tks_tx_seq(self.h, i, buf)
        return buf.raw[:L].decode()

    def __del__(self):
        try:
            if self.h:
                lib().tks_tx_free(self.h)
                self.h = None
        except Exception:
            pass


def small_transcriptome(seed: int = 1001, n: int = 500) -> Transcriptome:
    return Transcriptome(seed, n, 0, 500, 3000)


def large_transcriptome(seed: int = 1002, n: int = 50_000) -> Transcriptome:
    return Transcriptome(seed, n, 1, 1200.0, 0.6, 200, 15000, share_frac=0.10)


def se_reads(tx: Transcriptome, n_reads: int, read_len: int = 76, seed: int = 1001,
             sub_rate: float = 0.005, n_rate: float = 0.0, first: int = 0):
    """Single-end reads -> (ascii uint8 [n*read_len], uint64 offsets[n+1])."""
    out = np.empty(max(n_reads * read_len, 1), dtype=np.uint8)
    lib().tks_reads(tx.h, seed, 0, first, n_reads, read_len, 0.0, 0.0, sub_rate, n_rate, 0.5,
                    _threads(), out.ctypes.data, None)
    off = np.arange(n_reads + 1, dtype=np.uint64) * np.uint64(read_len)
    return out[: n_reads * read_len], off


def pe_reads(tx: Transcriptome, n_pairs: int, read_len: int = 150, seed: int = 1002,
             frag_mu: float = 300.0, frag_sd: float = 50.0, sub_rate: float = 0.01,
             n_rate: float = 0.0005, first: int = 0, out: np.ndarray | None = None):
    """Paired-end reads; pair i = left i, right i.

    Returns (left, right, offsets).  With `out` (uint8 of 2*n*read_len) the
    left reads fill its first half and the right reads its second half (the
    layout of Trinity's both.fa: left file then right file)."""
    if out is None:
        out = np.empty(max(2 * n_pairs * read_len, 1), dtype=np.uint8)
    half = n_pairs * read_len
    left, right = out[:half], out[half:2 * half]
    lib().tks_reads(tx.h, seed, 1, first, n_pairs, read_len, frag_mu, frag_sd, sub_rate, n_rate,
                    0.5, _threads(), left.ctypes.data, right.ctypes.data)
    off = np.arange(n_pairs + 1, dtype=np.uint64) * np.uint64(read_len)
    return left, right, off


def read_names(n: int, suffix: str = "", prefix: str = "r", start: int = 0) -> list[str]:
    return [f"{prefix}{i:09d}{suffix}" for i in range(start, start + n)]


def write_fasta(path, names, seq, off) -> None:
    buf = seq.tobytes() if isinstance(seq, np.ndarray) else seq
    with open(path, "w") as fh:
        for i, nm in enumerate(names):
            fh.write(f">{nm}\n{buf[int(off[i]):int(off[i + 1])].decode()}\n")
