"""ctypes bindings of the CPU oracle (oracle/drtopk_oracle.cpp) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker (or as the reported CPU baseline).  The library itself
never loads it.

All oracle functions take "ordered" uint32 keys; float32 keys are mapped with
`f32_to_ordered` first (sign flip: b ^ (0xFFFFFFFF if b>>31 else 0x80000000)).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

# Distributions of the synthetic generator (orc_generate).
UNIFORM_U32, NORMAL_U32, LOGNORMAL_U32, NORMAL_F32, UNIFORM_F32 = 0, 1, 2, 3, 4

ALGO_RADIX, ALGO_BUCKET, ALGO_BITONIC = 0, 1, 2
FLAG_FILTER, FLAG_BETA, FLAG_SHUFFLE = 0x1, 0x2, 0x4
FLAGS_DEFAULT = FLAG_FILTER | FLAG_BETA | FLAG_SHUFFLE


class Params(ctypes.Structure):
    """Mirror of drtopk_params (include/drtopk.h)."""
    _fields_ = [("alpha", ctypes.c_int32), ("beta", ctypes.c_int32), ("digit_bits", ctypes.c_int32),
                ("num_buckets", ctypes.c_int32), ("flags", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 3)]


class Stats(ctypes.Structure):
    """Mirror of drtopk_stats (include/drtopk.h)."""
    _fields_ = [("algo", ctypes.c_int32), ("alpha", ctypes.c_int32), ("beta", ctypes.c_int32),
                ("path", ctypes.c_int32), ("digit_bits", ctypes.c_int32), ("num_buckets", ctypes.c_int32),
                ("first_passes", ctypes.c_int32), ("second_passes", ctypes.c_int32),
                ("num_subranges", ctypes.c_uint64), ("num_delegates", ctypes.c_uint64),
                ("threshold", ctypes.c_uint32), ("kth_bits", ctypes.c_uint32),
                ("selected_subranges", ctypes.c_uint64), ("lonely_subranges", ctypes.c_uint64),
                ("lonely_elements", ctypes.c_uint64), ("concat_size", ctypes.c_uint64),
                ("count_greater", ctypes.c_uint64), ("ties_taken", ctypes.c_uint64),
                ("phase_ms", ctypes.c_float * 5), ("reserved", ctypes.c_uint32 * 3)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_
                if name not in ("reserved", "phase_ms")}


def params(alpha=-1, beta=2, digit_bits=8, num_buckets=256, flags=FLAGS_DEFAULT) -> Params:
    return Params(alpha, beta, digit_bits, num_buckets, flags, (ctypes.c_uint32 * 3)(0, 0, 0))


_lib = None


def build() -> None:
    """Compile the oracle (g++ -fopenmp) into oracle/build/liboracle.so."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u32p, i64p, u64p = (ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int64),
                            ctypes.POINTER(ctypes.c_uint64))
        L.orc_mt19937.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p]
        L.orc_generate.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, u32p]
        L.orc_f32_to_ordered.argtypes = [u32p, ctypes.c_uint64, u32p]
        L.orc_ordered_to_f32.argtypes = [u32p, ctypes.c_uint64, u32p]
        L.orc_topk_reference.argtypes = [u32p, ctypes.c_uint64, ctypes.c_uint64, u32p, u32p, i64p]
        L.orc_resolve.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(Params),
                                  ctypes.POINTER(ctypes.c_int)]
        L.orc_sample.argtypes = [u32p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, u32p, u32p]
        L.orc_drtopk.argtypes = [ctypes.c_int, u32p, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(Params),
                                 ctypes.c_int, u32p, u32p, i64p, ctypes.POINTER(Stats)]
        L.orc_bitonic_topk64.argtypes = [u64p, ctypes.c_uint64, ctypes.c_uint64, u64p]
        L.orc_radix_kth.argtypes = [u32p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
        L.orc_radix_kth.restype = ctypes.c_uint32
        L.orc_bucket_kth.argtypes = [u32p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
        L.orc_bucket_kth.restype = ctypes.c_uint32
        L.orc_bucket_bounds.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64,
                                        u32p, u32p]
        L.orc_bucket_of.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.orc_bucket_of.restype = ctypes.c_uint64
        L.orc_cpu_baseline.argtypes = [u32p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, u32p, u32p, i64p]
        L.orc_cpu_baseline.restype = ctypes.c_double
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_verify_topk.argtypes = [u32p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, u32p, i64p,
                                      ctypes.c_int, ctypes.c_int]
        L.orc_verify_topk.restype = ctypes.c_int
        L.orc_verify_rows.argtypes = [u32p, ctypes.c_uint64, u64p, u64p, u64p, u64p, ctypes.POINTER(ctypes.c_int),
                                      u32p, u32p, i64p, ctypes.POINTER(ctypes.c_int)]
        L.orc_verify_rows.restype = ctypes.c_uint64
        _lib = L
    return _lib


def _p(a, t=ctypes.c_uint32):
    return a.ctypes.data_as(ctypes.POINTER(t))


def mt19937(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, np.uint32)
    lib().orc_mt19937(seed, count, _p(out))
    return out


def generate(dist: int, seed: int, n: int) -> np.ndarray:
    """Seeded synthetic keys: uint32 array (float32 bit patterns for the float distributions)."""
    out = np.empty(n, np.uint32)
    if lib().orc_generate(dist, seed, n, _p(out)) != 0:
        raise ValueError(f"bad distribution {dist}")
    return out


def f32_to_ordered(bits: np.ndarray) -> np.ndarray:
    bits = np.ascontiguousarray(bits, dtype=np.uint32)
    out = np.empty_like(bits)
    lib().orc_f32_to_ordered(_p(bits), bits.size, _p(out))
    return out


def ordered_to_f32(o: np.ndarray) -> np.ndarray:
    o = np.ascontiguousarray(o, dtype=np.uint32)
    out = np.empty_like(o)
    lib().orc_ordered_to_f32(_p(o), o.size, _p(out))
    return out


def topk_reference(okeys: np.ndarray, k: int):
    """Closed form: nth_element k-th largest, top-k set with the lowest-index tie rule (index order)."""
    okeys = np.ascontiguousarray(okeys, dtype=np.uint32)
    kth = ctypes.c_uint32()
    keys = np.empty(k, np.uint32)
    idx = np.empty(k, np.int64)
    rc = lib().orc_topk_reference(_p(okeys), okeys.size, k, ctypes.byref(kth), _p(keys), _p(idx, ctypes.c_int64))
    if rc != 0:
        raise ValueError("bad k")
    return int(kth.value), keys, idx


def resolve(algo: int, n: int, k: int, p: Params):
    out = (ctypes.c_int * 6)()
    rc = lib().orc_resolve(algo, n, k, ctypes.byref(p), out)
    if rc != 0:
        return None
    return {"alpha": out[0], "beta": out[1], "path": out[2]}


def sample(okeys: np.ndarray, alpha: int, beta: int):
    okeys = np.ascontiguousarray(okeys, dtype=np.uint32)
    L = 1 << alpha
    S = (okeys.size + L - 1) // L
    vals = np.empty(beta * S, np.uint32)
    pos = np.empty(beta * S, np.uint32)
    if lib().orc_sample(_p(okeys), okeys.size, alpha, beta, _p(vals), _p(pos)) != 0:
        raise ValueError("bad sample args")
    return vals, pos


def drtopk(algo: int, okeys: np.ndarray, k: int, p: Params, f32: bool = False):
    """Restated Dr.Top-k: returns (kth, topk_keys, topk_idx, stats dict).  f32: the ordered keys
    come from float32 keys (the beta path then keeps full delegates)."""
    okeys = np.ascontiguousarray(okeys, dtype=np.uint32)
    kth = ctypes.c_uint32()
    keys = np.empty(k, np.uint32)
    idx = np.empty(k, np.int64)
    st = Stats()
    rc = lib().orc_drtopk(algo, _p(okeys), okeys.size, k, ctypes.byref(p), int(f32), ctypes.byref(kth), _p(keys),
                          _p(idx, ctypes.c_int64), ctypes.byref(st))
    if rc == 1:
        raise ValueError("invalid parameters")
    if rc != 0:
        raise RuntimeError(f"oracle internal error {rc}")
    return int(kth.value), keys, idx, st.as_dict()


def bitonic_topk64(a: np.ndarray, k: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    out = np.empty(k, np.uint64)
    if lib().orc_bitonic_topk64(_p(a, ctypes.c_uint64), a.size, k, _p(out, ctypes.c_uint64)) != 0:
        raise ValueError("bad bitonic args")
    return out


def radix_kth(a: np.ndarray, k: int, digit_bits: int = 8, plain: bool = False) -> int:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return int(lib().orc_radix_kth(_p(a), a.size, k, digit_bits, int(plain)))


def bucket_kth(a: np.ndarray, k: int, num_buckets: int = 256) -> int:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return int(lib().orc_bucket_kth(_p(a), a.size, k, num_buckets))


def bucket_bounds(lo: int, hi: int, num_buckets: int, b: int):
    x, y = ctypes.c_uint32(), ctypes.c_uint32()
    lib().orc_bucket_bounds(lo, hi, num_buckets, b, ctypes.byref(x), ctypes.byref(y))
    return int(x.value), int(y.value)


def bucket_of(v: int, lo: int, hi: int, num_buckets: int) -> int:
    return int(lib().orc_bucket_of(v, lo, hi, num_buckets))


def cpu_baseline(okeys: np.ndarray, k: int, threads: int = 0):
    """OpenMP nth_element baseline; returns (seconds, kth, keys, idx)."""
    okeys = np.ascontiguousarray(okeys, dtype=np.uint32)
    kth = ctypes.c_uint32()
    keys = np.empty(k, np.uint32)
    idx = np.empty(k, np.int64)
    t = lib().orc_cpu_baseline(_p(okeys), okeys.size, k, threads, ctypes.byref(kth), _p(keys),
                               _p(idx, ctypes.c_int64))
    return float(t), int(kth.value), keys, idx


def num_threads() -> int:
    return int(lib().orc_num_threads())


def verify_topk(raw: np.ndarray, k: int, kth_raw: int, out_keys=None, out_idx=None, f32: bool = False,
                threads: int = 0) -> int:
    """Exact O(n) check of a finished top-k on raw key bits (float32 bits when f32): 0 when the
    k-th key, the top-k keys and the indices (either may be None) are exactly the defined result;
    else 2 (wrong k-th key), 3 (an index differs) or 4 (a key differs)."""
    raw = np.ascontiguousarray(raw, dtype=np.uint32)
    ok = np.ascontiguousarray(out_keys, dtype=np.uint32) if out_keys is not None else None
    oi = np.ascontiguousarray(out_idx, dtype=np.int64) if out_idx is not None else None
    if (ok is not None and ok.size < k) or (oi is not None and oi.size < k):
        return 1
    return int(lib().orc_verify_topk(_p(raw), raw.size, k, int(kth_raw) & 0xFFFFFFFF,
                                     _p(ok) if ok is not None else None,
                                     _p(oi, ctypes.c_int64) if oi is not None else None, int(f32), threads))


def verify_rows(base: np.ndarray, rows, kth_raw: np.ndarray, out_keys=None, out_idx=None) -> np.ndarray:
    """verify_topk for every row (offset, n, k, out_offset, dtype) of one buffer, rows in parallel;
    returns the per-row codes (0 = exact)."""
    base = np.ascontiguousarray(base, dtype=np.uint32)
    nr = len(rows)
    off = np.array([r[0] for r in rows], np.uint64)
    n = np.array([r[1] for r in rows], np.uint64)
    k = np.array([r[2] for r in rows], np.uint64)
    oo = np.array([r[3] for r in rows], np.uint64)
    f = np.array([r[4] for r in rows], np.int32)
    kth = np.ascontiguousarray(kth_raw, dtype=np.uint32)
    ok = np.ascontiguousarray(out_keys, dtype=np.uint32) if out_keys is not None else None
    oi = np.ascontiguousarray(out_idx, dtype=np.int64) if out_idx is not None else None
    codes = np.zeros(nr, np.int32)
    lib().orc_verify_rows(_p(base), nr, _p(off, ctypes.c_uint64), _p(n, ctypes.c_uint64), _p(k, ctypes.c_uint64),
                          _p(oo, ctypes.c_uint64), _p(f, ctypes.c_int), _p(kth),
                          _p(ok) if ok is not None else None, _p(oi, ctypes.c_int64) if oi is not None else None,
                          _p(codes, ctypes.c_int))
    return codes
