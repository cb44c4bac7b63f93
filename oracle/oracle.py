"""ctypes wrapper of the CPU restatement (oracle/tk_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "build" / "libtk_oracle.so"

READ_STAT_DTYPE = np.dtype([("median", "<u4"), ("nwin", "<u4"), ("sum", "<u8"),
                            ("sumsq_lo", "<u8"), ("sumsq_hi", "<u8")])
R2T_VOTE_DTYPE = np.dtype([("comp", "<u4"), ("hits", "<u4"), ("nwin", "<u4")])

_lib = None
_P = C.c_void_p


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return SO


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not SO.exists():
            build()
        l = C.CDLL(str(SO))
        sig = {
            "or_base_code": (C.c_int, [C.c_int]),
            "or_revcomp": (C.c_uint64, [C.c_uint64, C.c_int]),
            "or_nwin": (C.c_uint64, [C.c_uint64, C.c_int]),
            "or_count": (C.c_int64, [C.c_char_p, _P, C.c_uint64, C.c_int, C.c_int, _P, _P, C.c_uint64]),
            "or_window_totals": (None, [C.c_char_p, _P, C.c_uint64, C.c_int, _P, _P]),
            "or_histo_nbins": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "or_histo_base": (C.c_uint64, [C.c_uint64, C.c_uint64]),
            "or_histo": (None, [_P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, _P]),
            "or_entropy_counts": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "or_entropy_keep_table": (None, [C.c_int, C.c_double, _P]),
            "or_prune": (C.c_uint64, [_P, _P, C.c_uint64, C.c_int, C.c_uint32, C.c_double, _P, _P]),
            "or_read_stats": (None, [C.c_char_p, _P, C.c_uint64, C.c_int, C.c_int, _P, _P, C.c_uint64,
                                     C.c_uint32, C.c_uint32, C.c_int, C.c_int, _P]),
            "or_format_stats_row": (C.c_int, [C.c_char_p, _P, C.c_char_p, C.c_size_t]),
            "or_drand48_seed": (C.c_uint64, [C.c_uint32]),
            "or_drand48_next": (C.c_double, [C.POINTER(C.c_uint64)]),
            "or_select_pairs": (C.c_uint64, [_P, _P, _P, _P, _P, _P, C.c_uint64, C.c_int, C.c_int,
                                             C.c_int, _P]),
            "or_perl_avg1": (C.c_double, [C.c_double, C.c_double]),
            "or_r2t_build": (C.c_int64, [C.c_char_p, _P, _P, C.c_uint64, C.c_int, C.c_int, _P, _P,
                                         C.c_uint64]),
            "or_r2t_vote_reads": (None, [C.c_char_p, _P, C.c_uint64, C.c_int, C.c_int, _P, _P,
                                         C.c_uint64, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def _off(o) -> np.ndarray:
    return np.ascontiguousarray(o, dtype=np.uint64)


def revcomp(v: int, k: int) -> int:
    return lib().or_revcomp(v, k)


def window_totals(seq, off, k: int) -> tuple[int, int]:
    """(all windows, windows free of non-ACGT bytes); seq is bytes or a uint8 array."""
    off = _off(off)
    t = np.zeros(1, np.uint64)
    v = np.zeros(1, np.uint64)
    if isinstance(seq, np.ndarray):
        seq = C.cast(C.c_void_p(seq.ctypes.data), C.c_char_p)
    lib().or_window_totals(seq, _p(off), len(off) - 1, k, _p(t), _p(v))
    return int(t[0]), int(v[0])


def count(seq: bytes, off, k: int = 25, canonical: bool = True):
    off = _off(off)
    total, _ = window_totals(seq, off, k)
    cap = max(total, 1)
    keys = np.zeros(cap, np.uint64)
    counts = np.zeros(cap, np.uint32)
    n = lib().or_count(seq, _p(off), len(off) - 1, k, int(canonical), _p(keys), _p(counts), cap)
    assert n >= 0
    return keys[:n].copy(), counts[:n].copy()


def histo(counts: np.ndarray, low: int = 1, high: int = 10000, inc: int = 1) -> np.ndarray:
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    nb = lib().or_histo_nbins(low, high, inc)
    bins = np.zeros(nb, np.uint64)
    lib().or_histo(_p(counts), len(counts), low, high, inc, _p(bins))
    return bins


def histo_base(low: int, inc: int) -> int:
    return lib().or_histo_base(low, inc)


def entropy_counts(a: int, c: int, g: int, t: int, k: int) -> float:
    return lib().or_entropy_counts(a, c, g, t, k)


def entropy_keep_table(k: int, min_entropy: float) -> np.ndarray:
    keep = np.zeros((k + 1) ** 3, np.uint8)
    lib().or_entropy_keep_table(k, float(min_entropy), _p(keep))
    return keep


def prune(keys, counts, k: int, min_count: int, min_entropy: float):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    ko = np.zeros(max(len(keys), 1), np.uint64)
    co = np.zeros(max(len(keys), 1), np.uint32)
    n = lib().or_prune(_p(keys), _p(counts), len(keys), k, min_count, float(min_entropy), _p(ko), _p(co))
    return ko[:n].copy(), co[:n].copy()


def read_stats(seq: bytes, off, keys, counts, k: int = 25, canonical: bool = True, min_count: int = 1,
               missing_value: int = 1, nwin_policy: int = 0, median_policy: int = 0) -> np.ndarray:
    off = _off(off)
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    out = np.zeros(len(off) - 1, READ_STAT_DTYPE)
    lib().or_read_stats(seq, _p(off), len(off) - 1, k, int(canonical),
                        _p(keys) if len(keys) else None, _p(counts) if len(counts) else None,
                        len(keys), min_count, missing_value, nwin_policy, median_policy, _p(out))
    return out


def format_stats_row(acc: str, st) -> str:
    buf = C.create_string_buffer(512)
    one = np.array([st], dtype=READ_STAT_DTYPE)
    lib().or_format_stats_row(acc.encode(), _p(one), buf, 512)
    return buf.value.decode()


def drand48_stream(seed: int, n: int) -> list[float]:
    st = C.c_uint64(lib().or_drand48_seed(seed))
    return [lib().or_drand48_next(C.byref(st)) for _ in range(n)]


def perl_avg1(a: float, b: float) -> float:
    return lib().or_perl_avg1(a, b)


def select_pairs(l_med, r_med, l_mean, r_mean, l_sd, r_sd, max_cov: int, min_cov: int = 1,
                 max_cv: int = 10000) -> np.ndarray:
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (l_med, r_med, l_mean, r_mean, l_sd, r_sd)]
    n = len(arrs[0])
    sel = np.zeros(max(n, 1), np.uint8)
    lib().or_select_pairs(*[_p(a) for a in arrs], n, max_cov, min_cov, max_cv, _p(sel))
    return sel[:n].copy()


def r2t_build(seq: bytes, off, comps, k: int = 25, canonical: bool = True):
    off = _off(off)
    comps = np.ascontiguousarray(comps, dtype=np.uint32)
    total, _ = window_totals(seq, off, k)
    cap = max(total, 1)
    keys = np.zeros(cap, np.uint64)
    cv = np.zeros(cap, np.uint32)
    n = lib().or_r2t_build(seq, _p(off), _p(comps), len(off) - 1, k, int(canonical), _p(keys), _p(cv), cap)
    assert n >= 0
    return keys[:n].copy(), cv[:n].copy()


def r2t_vote(seq: bytes, off, keys, comps, k: int = 25, canonical: bool = True) -> np.ndarray:
    off = _off(off)
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    comps = np.ascontiguousarray(comps, dtype=np.uint32)
    out = np.zeros(len(off) - 1, R2T_VOTE_DTYPE)
    lib().or_r2t_vote_reads(seq, _p(off), len(off) - 1, k, int(canonical),
                            _p(keys) if len(keys) else None, _p(comps) if len(comps) else None,
                            len(keys), _p(out))
    return out


# ---------------------------------------------------------------------------
# Multi-threaded CPU restatement of count + coverage stats (tk_cpuref.cpp):
# the CPU baseline bench.py times on the host cores; checked against the
# single-threaded functions above by tests/test_oracle.py.
_cpuref = None


def cpuref() -> C.CDLL:
    global _cpuref
    if _cpuref is None:
        so = HERE / "build" / "libtk_cpuref.so"
        if not so.exists():
            build()
        l = C.CDLL(str(so))
        l.cr_count_stats.restype = C.c_int
        l.cr_count_stats.argtypes = [C.c_char_p, _P, C.c_uint64, C.c_int, C.c_int, C.c_uint32, C.c_int,
                                     C.c_uint64, _P, _P, _P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_uint64), _P, _P]
        _cpuref = l
    return _cpuref


def cpuref_count_stats(seq, off, k: int = 25, canonical: bool = True, min_count: int = 2,
                       threads: int = 1, expected_distinct: int = 0, summary: bool = False) -> dict:
    """Count, then per-read (median, nwin, sum) with missing = 1, N-windows
    counted as missing and the upper median; also the wall-clock seconds of
    the two passes and the distinct-key count.  With `summary`, also the
    table digest (table_digest) and the default-bin histogram."""
    off = _off(off)
    n = len(off) - 1
    med = np.zeros(max(n, 1), np.uint32)
    nwin = np.zeros(max(n, 1), np.uint32)
    sm = np.zeros(max(n, 1), np.uint64)
    tc, ts, nd = C.c_double(), C.c_double(), C.c_uint64()
    if isinstance(seq, np.ndarray):
        sp = C.cast(C.c_void_p(seq.ctypes.data), C.c_char_p)
    else:
        sp = seq
    exp = expected_distinct or int(off[-1]) + 1024
    dg = np.zeros(1, np.uint64)
    hist = np.zeros(10001, np.uint64)
    rc = cpuref().cr_count_stats(sp, _p(off), n, k, int(canonical), min_count, threads, exp, _p(med),
                                 _p(nwin), _p(sm), C.byref(tc), C.byref(ts), C.byref(nd),
                                 _p(dg) if summary else None, _p(hist) if summary else None)
    if rc:
        raise MemoryError("cr_count_stats: out of memory")
    out = {"median": med[:n], "nwin": nwin[:n], "sum": sm[:n], "count_s": tc.value,
           "stats_s": ts.value, "distinct": nd.value}
    if summary:
        out["digest"] = int(dg[0])
        out["histo"] = hist
    return out


def table_digest(keys, counts) -> int:
    """Order-independent digest of a (key, count) table: sum mod 2^64 of
    mix64(key ^ count * 0x9E3779B97F4A7C15), as cr_count_stats computes it."""
    with np.errstate(over="ignore"):
        x = np.asarray(keys, np.uint64) ^ (np.asarray(counts, np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xFF51AFD7ED558CCD)
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xC4CEB9FE1A85EC53)
        x ^= x >> np.uint64(33)
        return int(x.sum(dtype=np.uint64))


def inchworm(keys, counts, k: int = 25, ds: bool = True, min_seed_cov: int = 2,
             seed_keep: np.ndarray | None = None):
    """or_inchworm: contigs (list of str), per-contig count sums and seed keys, in seed order."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    n = len(keys)
    cap = n * (k + 1) + 64
    seq = C.create_string_buffer(cap)
    off = np.zeros(n + 2, np.uint64)
    sums = np.zeros(n + 1, np.uint64)
    seeds = np.zeros(n + 1, np.uint64)
    nc = C.c_uint64()
    l = lib()
    l.or_inchworm.restype = C.c_int64
    l.or_inchworm.argtypes = [_P, _P, C.c_uint64, C.c_int, C.c_int, C.c_uint32, _P, C.c_char_p, C.c_uint64,
                              _P, _P, _P, C.c_uint64, C.POINTER(C.c_uint64)]
    keep = None if seed_keep is None else np.ascontiguousarray(seed_keep, np.uint8)
    rc = l.or_inchworm(_p(keys) if n else None, _p(counts) if n else None, n, k, int(ds), min_seed_cov,
                       _p(keep) if keep is not None else None, seq, cap, _p(off), _p(sums), _p(seeds), n + 1,
                       C.byref(nc))
    assert rc >= 0, rc
    m = nc.value
    raw = seq.raw
    contigs = [raw[int(off[i]):int(off[i + 1])].decode() for i in range(m)]
    return contigs, sums[:m].copy(), seeds[:m].copy()


def kset_count(keys, seq: bytes, off, kk: int, canonical: bool = True, tags=None) -> np.ndarray:
    """or_kset_count: occurrences of each (lo, hi) key among the reads' windows."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, 2)
    off = _off(off)
    n = len(keys)
    out = np.zeros(max(n, 1), np.uint64)
    l = lib()
    l.or_kset_count.restype = C.c_int
    l.or_kset_count.argtypes = [_P, C.c_uint64, C.c_char_p, _P, C.c_uint64, C.c_int, C.c_int, _P, _P]
    t = None if tags is None else np.ascontiguousarray(tags, dtype=np.uint32)
    rc = l.or_kset_count(_p(keys) if n else None, n, seq, _p(off), len(off) - 1, kk, int(canonical),
                         _p(t) if t is not None else None, _p(out))
    assert rc == 0, rc
    return out[:n]
