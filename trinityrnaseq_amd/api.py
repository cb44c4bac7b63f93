"""ctypes binding of libtrinity_gpu.so (include/trinity_gpu.h) for tests and the bench.

The command-line tools in trinityrnaseq_amd/bin are the drop-in replacements
for the reference binaries; this module exposes the same entry points to
Python.  Loading fails loudly when the HIP library has not been built.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libtrinity_gpu.so"

READ_STAT_DTYPE = np.dtype([("median", "<u4"), ("nwin", "<u4"), ("sum", "<u8"),
                            ("sumsq_lo", "<u8"), ("sumsq_hi", "<u8")])
R2T_VOTE_DTYPE = np.dtype([("comp", "<u4"), ("hits", "<u4"), ("nwin", "<u4")])

MEDIAN_UPPER, MEDIAN_LOWER, MEDIAN_MEAN_FLOOR = 0, 1, 2
NWIN_AS_MISSING, NWIN_SKIP = 0, 1

STATUS = {0: "TK_OK", -1: "TK_ERR_ARG", -2: "TK_ERR_HIP", -3: "TK_ERR_OOM",
          -4: "TK_ERR_FULL", -5: "TK_ERR_STATE", -6: "TK_ERR_CAP"}

# name -> (restype, argtypes); the full exported surface of trinity_gpu.h
_P = C.c_void_p
_U64P = C.POINTER(C.c_uint64)
_U32P = C.POINTER(C.c_uint32)
SIGNATURES = {
    "tk_last_error": (C.c_char_p, []),
    "tk_abi_version": (C.c_int, []),
    "tk_init": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "tk_ctx_free": (C.c_int, [_P]),
    "tk_sync": (C.c_int, [_P]),
    "tk_timing_enable": (C.c_int, [_P, C.c_int]),
    "tk_timing_reset": (C.c_int, [_P]),
    "tk_timing_get": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double), _U64P]),
    "tk_device_bytes_in_use": (C.c_int, [_U64P]),
    "tk_thread_cache": (C.c_int, [C.c_uint64]),
    "tk_count_words": (C.c_int, [_P, _U64P]),
    "tk_buffer_alloc": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "tk_buffer_free": (C.c_int, [_P, _P]),
    "tk_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_uint64]),
    "tk_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_uint64]),
    "tk_memcpy_d2d": (C.c_int, [_P, _P, _P, C.c_uint64]),
    "tk_host_alloc": (C.c_int, [C.c_uint64, C.POINTER(_P)]),
    "tk_host_free": (C.c_int, [_P]),
    "tk_reads_from_host": (C.c_int, [_P, C.c_char_p, _P, C.c_uint64, C.POINTER(_P)]),
    "tk_reads_from_device_d": (C.c_int, [_P, _P, _P, C.c_uint64, C.POINTER(_P)]),
    "tk_reads_info": (C.c_int, [_P, _U64P, _U64P, _U32P]),
    "tk_reads_windows": (C.c_int, [_P, C.c_int, _U64P]),
    "tk_reads_free": (C.c_int, [_P]),
    "tk_estimate_distinct": (C.c_int, [_P, _P, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "tk_table_create": (C.c_int, [_P, C.c_int, C.c_int, C.c_uint64, C.c_double, C.POINTER(_P)]),
    "tk_table_info": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), _U64P, _U64P]),
    "tk_count": (C.c_int, [_P, _P, _P]),
    "tk_table_dump_size": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _U64P]),
    "tk_table_dump": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, _P, C.c_uint64, _U64P]),
    "tk_table_histo": (C.c_int, [_P, _P, C.c_uint64, C.c_uint64, C.c_uint64, _P, C.c_uint64]),
    "tk_inchworm": (C.c_int, [_P, _P, C.c_uint32, _P, C.POINTER(_P)]),
    "tk_kset_create": (C.c_int, [_P, _P, C.c_uint64, C.POINTER(_P)]),
    "tk_kset_count": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, _P, _P]),
    "tk_kset_free": (C.c_int, [_P]),
    "tk_contigs_info": (C.c_int, [_P, _U64P, _U64P, _U64P, _U64P]),
    "tk_contigs_get": (C.c_int, [_P, _P, _P, _P, _P]),
    "tk_contigs_free": (C.c_int, [_P]),
    "tk_table_load": (C.c_int, [_P, _P, _P, _P, C.c_uint64]),
    "tk_table_prune": (C.c_int, [_P, _P, C.c_uint32, _P, _U64P]),
    "tk_table_lookup": (C.c_int, [_P, _P, _P, C.c_uint64, _P]),
    "tk_table_free": (C.c_int, [_P]),
    "tk_read_cov_stats": (C.c_int, [_P, _P, _P, _P, _P]),
    "tk_read_cov_stats_d": (C.c_int, [_P, _P, _P, _P, _P]),
    "tk_select_pairs_d_noalloc": (C.c_int, [_P, _P, _P, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_uint32, _P,
                                           _U64P]),
    "tk_select_pairs_d": (C.c_int, [_P, _P, _P, C.c_uint64, C.c_int, C.c_int, C.c_int,
                                    C.c_uint32, _P, _U64P]),
    "tk_select_pairs": (C.c_int, [_P, _P, _P, C.c_uint64, C.c_int, C.c_int, C.c_int, _P, _U64P]),
    "tk_ctable_build": (C.c_int, [_P, C.c_int, C.c_int, _P, _P, C.POINTER(_P)]),
    "tk_ctable_info": (C.c_int, [_P, _U64P, _U64P]),
    "tk_r2t_vote_reads": (C.c_int, [_P, _P, _P, _P]),
    "tk_r2t_vote_reads_d": (C.c_int, [_P, _P, _P, _P]),
    "tk_ctable_free": (C.c_int, [_P]),
    "tk_batch_create": (C.c_int, [_P, C.c_int, C.c_int, _P, _P, _P, _P, C.c_uint64, _P, C.POINTER(_P)]),
    "tk_batch_run_d": (C.c_int, [_P, _P, _P, _P, _P]),
    "tk_batch_run": (C.c_int, [_P, _P, _P, _P, _P]),
    "tk_batch_autotune": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(C.c_double)]),
    "tk_batch_info": (C.c_int, [_P, _U64P, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "tk_batch_free": (C.c_int, [_P]),
    "tk_server_create": (C.c_int, [_P, C.POINTER(_P)]),
    "tk_server_submit": (C.c_int, [_P, _P, _P, C.c_uint64, _P, _U64P]),
    "tk_server_wait": (C.c_int, [_P, C.c_uint64, _P]),
    "tk_server_submit_job": (C.c_int, [_P, _P, _U64P]),
    "tk_server_free": (C.c_int, [_P]),
}


class StatsOpts(C.Structure):
    _fields_ = [("min_count", C.c_uint32), ("missing_value", C.c_uint32),
                ("nwin_policy", C.c_int32), ("median_policy", C.c_int32)]


class BatchOpts(C.Structure):
    _fields_ = [("window_load", C.c_double), ("stats_waves", C.c_int32), ("use_graph", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("k", C.c_int32), ("min_count", C.c_int32),
                ("max_cov", C.c_int32), ("max_cv", C.c_int32), ("max_load", C.c_double),
                ("hbm_budget_bytes", C.c_uint64)]


class JobResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_selected", C.c_uint64), ("distinct_estimate", C.c_double),
                ("wait_s", C.c_double), ("run_s", C.c_double), ("latency_s", C.c_double),
                ("message", C.c_char * 256)]


class ServerSelection(C.Structure):
    _fields_ = [("left", C.c_void_p), ("right", C.c_void_p), ("n_units", C.c_uint64), ("selected", C.c_void_p),
                ("n_selected", C.c_uint64)]


class ServerJob(C.Structure):
    _fields_ = [("seq", C.c_void_p), ("offsets", C.c_void_p), ("n_reads", C.c_uint64), ("k", C.c_int32),
                ("canonical", C.c_int32), ("max_cov", C.c_int32), ("min_cov", C.c_int32), ("max_cv", C.c_int32),
                ("n_selections", C.c_int32), ("sel", ServerSelection * 2)]


class TkError(RuntimeError):
    def __init__(self, code: int, fn: str, msg: str):
        super().__init__(f"{fn} -> {STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _check(rc: int, fn: str) -> None:
    if rc != 0:
        raise TkError(rc, fn, lib().tk_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a.size else None


class Context:
    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        _check(lib().tk_init(device, C.byref(self.h)), "tk_init")

    def close(self):
        if self.h:
            lib().tk_ctx_free(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        _check(lib().tk_sync(self.h), "tk_sync")

    def timing(self, on: bool = True):
        _check(lib().tk_timing_enable(self.h, int(on)), "tk_timing_enable")

    def timing_reset(self):
        _check(lib().tk_timing_reset(self.h), "tk_timing_reset")

    def count_words(self) -> int:
        """Merged words of the last count's region-resident passes (0: table atomics)."""
        w = C.c_uint64()
        _check(lib().tk_count_words(self.h, C.byref(w)), "tk_count_words")
        return w.value

    def kernel_time(self, name: str) -> tuple[float, int]:
        ms = C.c_double()
        n = C.c_uint64()
        _check(lib().tk_timing_get(self.h, name.encode(), C.byref(ms), C.byref(n)), "tk_timing_get")
        return ms.value, n.value


class DeviceBuffer:
    """Device memory owned by the library (tk_buffer_alloc)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib().tk_buffer_alloc(ctx.h, self.nbytes, C.byref(p)), "tk_buffer_alloc")
        self.ptr = p.value or 0

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        _check(lib().tk_memcpy_h2d(self.ctx.h, C.c_void_p(self.ptr + offset), _ptr(a), a.nbytes),
               "tk_memcpy_h2d")

    def download(self, a: np.ndarray, offset: int = 0):
        assert offset + a.nbytes <= self.nbytes and a.flags.c_contiguous
        _check(lib().tk_memcpy_d2h(self.ctx.h, _ptr(a), C.c_void_p(self.ptr + offset), a.nbytes),
               "tk_memcpy_d2h")

    def copy_from(self, other: "DeviceBuffer", nbytes: int):
        assert nbytes <= self.nbytes and nbytes <= other.nbytes
        _check(lib().tk_memcpy_d2d(self.ctx.h, C.c_void_p(self.ptr), C.c_void_p(other.ptr), nbytes),
               "tk_memcpy_d2d")

    def free(self):
        if self.ptr:
            lib().tk_buffer_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostBuffer:
    """Page-locked host memory (tk_host_alloc), viewed as a uint8 array."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib().tk_host_alloc(self.nbytes, C.byref(p)), "tk_host_alloc")
        self.ptr = p.value or 0
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(self.nbytes, 1)).from_address(self.ptr))[: self.nbytes]

    def free(self):
        if self.ptr:
            self.array = None
            lib().tk_host_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def entropy_keep_table(k: int, min_entropy: float) -> np.ndarray:
    """Inchworm's entropy filter as a composition table for Table.prune:
    keep[(a*(k+1)+c)*(k+1)+g] = 1 when the Shannon entropy (bits) of a k-mer
    with a A's, c C's, g G's and k-a-c-g T's is >= min_entropy (the rule the
    inchworm CLI applies for --min_any_entropy)."""
    K1 = k + 1
    keep = np.zeros(K1 * K1 * K1, np.uint8)
    for a in range(K1):
        for c in range(K1 - a):
            for g in range(K1 - a - c):
                h = 0.0
                for n in (a, c, g, k - a - c - g):
                    if n > 0:
                        p = n / k
                        h -= p * np.log2(p)
                keep[(a * K1 + c) * K1 + g] = h >= min_entropy
    return keep


def pack_sequences(seqs) -> tuple[bytes, np.ndarray]:
    """List of str/bytes -> (concatenated bytes, uint64 offsets[n+1])."""
    bs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
    off = np.zeros(len(bs) + 1, dtype=np.uint64)
    if bs:
        off[1:] = np.cumsum([len(b) for b in bs], dtype=np.uint64)
    return b"".join(bs), off


class Reads:
    def __init__(self, ctx: Context, seq, offsets: np.ndarray):
        """seq: bytes, or a uint8 array (pinned host memory uploads without staging)."""
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        if isinstance(seq, np.ndarray):
            seq = np.ascontiguousarray(seq, dtype=np.uint8)
            sp = C.cast(C.c_void_p(seq.ctypes.data), C.c_char_p) if seq.size else None
        else:
            sp = seq
        self.h = C.c_void_p()
        self._ctx = ctx
        _check(lib().tk_reads_from_host(ctx.h, sp, _ptr(offsets), len(offsets) - 1, C.byref(self.h)),
               "tk_reads_from_host")

    @classmethod
    def from_device(cls, ctx: Context, d_seq: int, d_offsets: int, n: int) -> "Reads":
        self = cls.__new__(cls)
        self.h = C.c_void_p()
        self._ctx = ctx
        _check(lib().tk_reads_from_device_d(ctx.h, C.c_void_p(d_seq), C.c_void_p(d_offsets), n,
                                            C.byref(self.h)), "tk_reads_from_device_d")
        return self

    @classmethod
    def from_list(cls, ctx: Context, seqs) -> "Reads":
        s, o = pack_sequences(seqs)
        return cls(ctx, s, o)

    def info(self):
        n = C.c_uint64(); b = C.c_uint64(); m = C.c_uint32()
        _check(lib().tk_reads_info(self.h, C.byref(n), C.byref(b), C.byref(m)), "tk_reads_info")
        return n.value, b.value, m.value

    def windows(self, k: int) -> int:
        t = C.c_uint64()
        _check(lib().tk_reads_windows(self.h, k, C.byref(t)), "tk_reads_windows")
        return t.value

    def free(self):
        if self.h:
            lib().tk_reads_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def estimate_distinct(ctx: Context, reads: Reads, k: int, canonical: bool = True) -> float:
    e = C.c_double()
    _check(lib().tk_estimate_distinct(ctx.h, reads.h, k, int(canonical), C.byref(e)),
           "tk_estimate_distinct")
    return e.value


class Table:
    def __init__(self, ctx: Context, k: int = 25, canonical: bool = True, expected: int = 1 << 16,
                 max_load: float = 0.5):
        self.ctx = ctx
        self.k = k
        self.h = C.c_void_p()
        _check(lib().tk_table_create(ctx.h, k, int(canonical), int(expected), float(max_load),
                                     C.byref(self.h)), "tk_table_create")

    def info(self):
        k = C.c_int(); can = C.c_int(); cap = C.c_uint64(); b = C.c_uint64()
        _check(lib().tk_table_info(self.h, C.byref(k), C.byref(can), C.byref(cap), C.byref(b)),
               "tk_table_info")
        return {"k": k.value, "canonical": can.value, "capacity": cap.value, "bytes": b.value}

    def count(self, reads: Reads):
        _check(lib().tk_count(self.ctx.h, self.h, reads.h), "tk_count")

    def dump_size(self, lo: int = 1, hi: int = 0xFFFFFFFF) -> int:
        n = C.c_uint64()
        _check(lib().tk_table_dump_size(self.ctx.h, self.h, lo, hi, C.byref(n)), "tk_table_dump_size")
        return n.value

    def dump(self, lo: int = 1, hi: int = 0xFFFFFFFF):
        n = self.dump_size(lo, hi)
        keys = np.empty(n, dtype=np.uint64)
        counts = np.empty(n, dtype=np.uint32)
        got = C.c_uint64()
        _check(lib().tk_table_dump(self.ctx.h, self.h, lo, hi, _ptr(keys), _ptr(counts), n,
                                   C.byref(got)), "tk_table_dump")
        return keys[: got.value], counts[: got.value]

    def histo(self, low: int = 1, high: int = 10000, inc: int = 1) -> np.ndarray:
        base = (1 if inc >= low else low - inc) if low > 1 else 1
        nb = (high + 2 * inc - base) // inc
        bins = np.zeros(nb, dtype=np.uint64)
        _check(lib().tk_table_histo(self.ctx.h, self.h, low, high, inc, _ptr(bins), nb),
               "tk_table_histo")
        return bins

    def load(self, keys: np.ndarray, counts: np.ndarray):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        assert keys.shape == counts.shape
        _check(lib().tk_table_load(self.ctx.h, self.h, _ptr(keys), _ptr(counts), len(keys)),
               "tk_table_load")

    def prune(self, min_count: int = 1, keep_table: np.ndarray | None = None) -> int:
        kept = C.c_uint64()
        kt = None
        if keep_table is not None:
            keep_table = np.ascontiguousarray(keep_table, dtype=np.uint8)
            assert keep_table.size == (self.k + 1) ** 3
            kt = _ptr(keep_table)
        _check(lib().tk_table_prune(self.ctx.h, self.h, min_count, kt, C.byref(kept)), "tk_table_prune")
        return kept.value

    def lookup(self, keys: np.ndarray) -> np.ndarray:
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.zeros(len(keys), dtype=np.uint32)
        _check(lib().tk_table_lookup(self.ctx.h, self.h, _ptr(keys), len(keys), _ptr(out)),
               "tk_table_lookup")
        return out

    def free(self):
        if self.h:
            lib().tk_table_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def read_cov_stats(ctx: Context, table: Table, reads: Reads, min_count: int = 1,
                   missing_value: int = 1, nwin_policy: int = NWIN_AS_MISSING,
                   median_policy: int = MEDIAN_UPPER) -> np.ndarray:
    n = reads.info()[0]
    out = np.zeros(n, dtype=READ_STAT_DTYPE)
    o = StatsOpts(min_count, missing_value, nwin_policy, median_policy)
    _check(lib().tk_read_cov_stats(ctx.h, table.h, reads.h, C.byref(o), _ptr(out)),
           "tk_read_cov_stats")
    return out


def read_cov_stats_d(ctx: Context, table: Table, reads: Reads, d_out: int, min_count: int = 1,
                     missing_value: int = 1, nwin_policy: int = NWIN_AS_MISSING,
                     median_policy: int = MEDIAN_UPPER) -> None:
    o = StatsOpts(min_count, missing_value, nwin_policy, median_policy)
    _check(lib().tk_read_cov_stats_d(ctx.h, table.h, reads.h, C.byref(o), C.c_void_p(d_out)),
           "tk_read_cov_stats_d")


def select_pairs_d(ctx: Context, d_left: int, d_right: int, n: int, max_cov: int, min_cov: int,
                   max_cv: int, max_nwin: int, d_sel: int) -> int:
    ns = C.c_uint64()
    _check(lib().tk_select_pairs_d(ctx.h, C.c_void_p(d_left), C.c_void_p(d_right), n, max_cov,
                                   min_cov, max_cv, max_nwin, C.c_void_p(d_sel), C.byref(ns)),
           "tk_select_pairs_d")
    return ns.value


def select_pairs(ctx: Context, left: np.ndarray, right: np.ndarray, max_cov: int, min_cov: int = 1,
                 max_cv: int = 10000) -> tuple[np.ndarray, int]:
    left = np.ascontiguousarray(left, dtype=READ_STAT_DTYPE)
    right = np.ascontiguousarray(right, dtype=READ_STAT_DTYPE)
    assert left.shape == right.shape
    sel = np.zeros(len(left), dtype=np.uint8)
    ns = C.c_uint64()
    _check(lib().tk_select_pairs(ctx.h, _ptr(left), _ptr(right), len(left), max_cov, min_cov, max_cv,
                                 _ptr(sel), C.byref(ns)), "tk_select_pairs")
    return sel, ns.value


class KSet:
    """A fixed set of 128-bit keys (tk_kset): keys is a (n, 2) uint64 array of (lo, hi)."""

    def __init__(self, ctx: Context, keys: np.ndarray):
        self.ctx = ctx
        keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, 2)
        self.n = len(keys)
        self.h = C.c_void_p()
        _check(lib().tk_kset_create(ctx.h, _ptr(keys) if self.n else None, self.n, C.byref(self.h)),
               "tk_kset_create")

    def count(self, reads: Reads, kk: int, canonical: bool = True, tags: np.ndarray | None = None) -> np.ndarray:
        out = np.zeros(max(self.n, 1), np.uint64)
        t = None if tags is None else np.ascontiguousarray(tags, dtype=np.uint32)
        _check(lib().tk_kset_count(self.ctx.h, self.h, reads.h, kk, int(canonical),
                                   _ptr(t) if t is not None else None, _ptr(out)), "tk_kset_count")
        return out[:self.n]

    def free(self):
        if self.h:
            lib().tk_kset_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def inchworm(ctx: Context, table: "Table", min_seed_cov: int = 2, seed_keep: np.ndarray | None = None):
    """Greedy contigs over the table (tk_inchworm): (list of str, count sums,
    seed keys, {"rounds", "walks"})."""
    h = C.c_void_p()
    keep = None if seed_keep is None else np.ascontiguousarray(seed_keep, dtype=np.uint8)
    _check(lib().tk_inchworm(ctx.h, table.h, min_seed_cov, _ptr(keep) if keep is not None else None,
                             C.byref(h)), "tk_inchworm")
    try:
        n, tot, rounds, walks = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib().tk_contigs_info(h, C.byref(n), C.byref(tot), C.byref(rounds), C.byref(walks)),
               "tk_contigs_info")
        seq = np.zeros(max(tot.value, 1), np.uint8)
        off = np.zeros(n.value + 1, np.uint64)
        sums = np.zeros(max(n.value, 1), np.uint64)
        seeds = np.zeros(max(n.value, 1), np.uint64)
        _check(lib().tk_contigs_get(h, _ptr(seq), _ptr(off), _ptr(sums), _ptr(seeds)), "tk_contigs_get")
    finally:
        lib().tk_contigs_free(h)
    raw = seq.tobytes()
    contigs = [raw[int(off[i]):int(off[i + 1])].decode() for i in range(n.value)]
    return contigs, sums[:n.value], seeds[:n.value], {"rounds": rounds.value, "walks": walks.value}


class CTable:
    def __init__(self, ctx: Context, bundles: Reads, comp_ids: np.ndarray, k: int = 25,
                 canonical: bool = True):
        self.ctx = ctx
        comp_ids = np.ascontiguousarray(comp_ids, dtype=np.uint32)
        self.h = C.c_void_p()
        _check(lib().tk_ctable_build(ctx.h, k, int(canonical), bundles.h, _ptr(comp_ids),
                                     C.byref(self.h)), "tk_ctable_build")

    def info(self):
        n = C.c_uint64(); cap = C.c_uint64()
        _check(lib().tk_ctable_info(self.h, C.byref(n), C.byref(cap)), "tk_ctable_info")
        return n.value, cap.value

    def vote(self, reads: Reads) -> np.ndarray:
        n = reads.info()[0]
        out = np.zeros(n, dtype=R2T_VOTE_DTYPE)
        _check(lib().tk_r2t_vote_reads(self.ctx.h, self.h, reads.h, _ptr(out)), "tk_r2t_vote_reads")
        return out

    def vote_d(self, reads: Reads, d_out: int) -> None:
        """Votes into a device buffer of 12 bytes per read (asynchronous)."""
        _check(lib().tk_r2t_vote_reads_d(self.ctx.h, self.h, reads.h, C.c_void_p(d_out)), "tk_r2t_vote_reads_d")

    def free(self):
        if self.h:
            lib().tk_ctable_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_bytes_in_use() -> int:
    b = C.c_uint64()
    _check(lib().tk_device_bytes_in_use(C.byref(b)), "tk_device_bytes_in_use")
    return b.value


def thread_cache(limit_bytes: int) -> None:
    """Reuse the device blocks this thread frees, up to limit_bytes held
    (0 turns reuse off and releases them); the thread must work through one
    context."""
    _check(lib().tk_thread_cache(int(limit_bytes)), "tk_thread_cache")


class Batch:
    """Many components (reads + contigs each) counted, scored and voted in one run."""

    def __init__(self, ctx: Context, reads: Reads, read_comp_off, contigs: Reads, contig_comp_off,
                 k: int = 25, canonical: bool = True, window_load: float = 0.0, stats_waves: int = 0,
                 use_graph: bool = True):
        self.ctx = ctx
        self._keep = (reads, contigs)
        self.n_reads = reads.info()[0]
        ro = np.ascontiguousarray(read_comp_off, dtype=np.uint64)
        co = np.ascontiguousarray(contig_comp_off, dtype=np.uint64)
        assert len(ro) == len(co)
        o = BatchOpts(window_load, stats_waves, int(use_graph))
        self.h = C.c_void_p()
        _check(lib().tk_batch_create(ctx.h, k, int(canonical), reads.h, _ptr(ro), contigs.h, _ptr(co),
                                     len(ro) - 1, C.byref(o), C.byref(self.h)), "tk_batch_create")

    def run(self, min_count: int = 1, missing_value: int = 1, nwin_policy: int = NWIN_AS_MISSING,
            median_policy: int = MEDIAN_UPPER):
        st = np.zeros(self.n_reads, dtype=READ_STAT_DTYPE)
        vo = np.zeros(self.n_reads, dtype=R2T_VOTE_DTYPE)
        o = StatsOpts(min_count, missing_value, nwin_policy, median_policy)
        _check(lib().tk_batch_run(self.ctx.h, self.h, C.byref(o), _ptr(st), _ptr(vo)), "tk_batch_run")
        return st, vo

    def run_d(self, d_stats: int, d_votes: int, min_count: int = 1, missing_value: int = 1,
              nwin_policy: int = NWIN_AS_MISSING, median_policy: int = MEDIAN_UPPER):
        o = StatsOpts(min_count, missing_value, nwin_policy, median_policy)
        _check(lib().tk_batch_run_d(self.ctx.h, self.h, C.byref(o), C.c_void_p(d_stats or None),
                                    C.c_void_p(d_votes or None)), "tk_batch_run_d")

    def autotune(self, d_stats: int, d_votes: int, min_count: int = 1) -> float:
        o = StatsOpts(min_count, 1, NWIN_AS_MISSING, MEDIAN_UPPER)
        ms = C.c_double()
        _check(lib().tk_batch_autotune(self.ctx.h, self.h, C.byref(o), C.c_void_p(d_stats or None),
                                       C.c_void_p(d_votes or None), C.byref(ms)), "tk_batch_autotune")
        return ms.value

    def info(self):
        b = C.c_uint64(); wl = C.c_double(); w = C.c_int()
        _check(lib().tk_batch_info(self.h, C.byref(b), C.byref(wl), C.byref(w)), "tk_batch_info")
        return {"bytes": b.value, "window_load": wl.value, "stats_waves": w.value}

    def free(self):
        if self.h:
            lib().tk_batch_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Resident normalization service (count + coverage stats + selection per sample)."""

    def __init__(self, device: int = 0, slots: int = 4, k: int = 25, min_count: int = 2, max_cov: int = 50,
                 max_cv: int = 10000, max_load: float = 0.5, hbm_budget_bytes: int = 0):
        o = ServerOpts(device, slots, k, min_count, max_cov, max_cv, max_load, hbm_budget_bytes)
        self.h = C.c_void_p()
        _check(lib().tk_server_create(C.byref(o), C.byref(self.h)), "tk_server_create")
        self._keep = {}

    def submit(self, seq: np.ndarray, offsets: np.ndarray, n_pairs: int, want_selected: bool = True) -> int:
        seq = np.ascontiguousarray(seq, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        assert len(offsets) == 2 * n_pairs + 1
        sel = np.zeros(max(n_pairs, 1), np.uint8) if want_selected else None
        jid = C.c_uint64()
        _check(lib().tk_server_submit(self.h, _ptr(seq), _ptr(offsets), n_pairs,
                                      _ptr(sel) if sel is not None else None, C.byref(jid)), "tk_server_submit")
        self._keep[jid.value] = (seq, offsets, sel)
        return jid.value

    def submit_job(self, seq: np.ndarray, offsets: np.ndarray, selections, k: int = 25, canonical: bool = True,
                   max_cov: int = 50, min_cov: int = 1, max_cv: int = 10000) -> int:
        """The general job (tk_server_submit_job): every read counted and scored,
        then one or two selections, each a (left, right) pair of read-index
        arrays in unit order (right None: single-read units)."""
        seq = np.ascontiguousarray(seq, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        assert 1 <= len(selections) <= 2
        job = ServerJob()
        job.seq, job.offsets, job.n_reads = _ptr(seq), _ptr(offsets), len(offsets) - 1
        job.k, job.canonical, job.max_cov, job.min_cov, job.max_cv = k, int(canonical), max_cov, min_cov, max_cv
        job.n_selections = len(selections)
        keep = {"buffers": [seq, offsets], "job": job, "selected": []}
        for q, (left, right) in enumerate(selections):
            left = np.ascontiguousarray(left, dtype=np.uint64)
            right = None if right is None else np.ascontiguousarray(right, dtype=np.uint64)
            sel = np.zeros(max(len(left), 1), np.uint8)
            keep["buffers"] += [left, right]
            keep["selected"].append(sel)
            job.sel[q] = ServerSelection(_ptr(left), _ptr(right) if right is not None else None, len(left),
                                         _ptr(sel), 0)
        jid = C.c_uint64()
        _check(lib().tk_server_submit_job(self.h, C.byref(job), C.byref(jid)), "tk_server_submit_job")
        self._keep[jid.value] = keep
        return jid.value

    def wait(self, job_id: int):
        """(result, selected flags); for a general job the flags are a list, one array per selection."""
        r = JobResult()
        _check(lib().tk_server_wait(self.h, job_id, C.byref(r)), "tk_server_wait")
        kept = self._keep.pop(job_id)
        if r.status:
            raise TkError(r.status, "server job", r.message.decode(errors="replace"))
        if isinstance(kept, dict):
            job = kept["job"]
            return r, [x[: job.sel[q].n_units] for q, x in enumerate(kept["selected"])]
        seq, off, sel = kept
        return r, (sel[: (len(off) - 1) // 2] if sel is not None else None)

    def close(self):
        if self.h:
            lib().tk_server_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
