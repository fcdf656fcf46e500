"""Synthetic row mixes for the batched and served forms.

batch_mix:   rows of length log-uniform in [2^lo, 2^hi], key type by a fair coin (uint32
             uniform or float32 N(0,1)), k log-uniform in [1, min(n/8, kmax)]
             (4096 rows, 2^12..2^22, kmax 4096 is the batched workload).
serve_mix:   request lengths log-uniform in [2^lo, 2^hi], k log-uniform in [1, min(n/4, kmax)]
             (2^10..2^26, kmax 2^16 for the served workload), Poisson arrival times.
pack:        16-byte aligned offsets of the rows in one key buffer and output offsets.
"""
from __future__ import annotations

import numpy as np

U32, F32 = 0, 1


def _log_uniform_int(rng, lo_log2: float, hi_log2, size):
    return np.floor(np.exp2(rng.uniform(lo_log2, hi_log2, size))).astype(np.int64)


def batch_mix(seed: int = 7, nrows: int = 4096, lo: int = 12, hi: int = 22, kmax: int = 4096):
    """[(n, k, dtype)] for the batched workload."""
    rng = np.random.default_rng(seed)
    n = np.minimum(_log_uniform_int(rng, lo, hi, nrows), 1 << hi)
    dt = rng.integers(0, 2, nrows)
    kcap = np.maximum(1, np.minimum(n // 8, kmax))
    k = np.minimum(_log_uniform_int(rng, 0.0, np.log2(kcap), nrows), kcap)
    k = np.maximum(k, 1)
    return [(int(a), int(b), int(c)) for a, b, c in zip(n, k, dt)]


def serve_mix(seed: int = 8, nreq: int = 2000, lo: int = 10, hi: int = 26, kmax: int = 1 << 16,
              rate_per_s: float = 2000.0):
    """[(arrival_s, n, k, dtype)] for the served workload (Poisson arrivals at rate_per_s)."""
    rng = np.random.default_rng(seed)
    n = np.minimum(_log_uniform_int(rng, lo, hi, nreq), 1 << hi)
    dt = rng.integers(0, 2, nreq)
    kcap = np.maximum(1, np.minimum(n // 4, kmax))
    k = np.maximum(1, np.minimum(_log_uniform_int(rng, 0.0, np.log2(kcap), nreq), kcap))
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, nreq))
    return [(float(a), int(b), int(c), int(d)) for a, b, c, d in zip(t, n, k, dt)]


def pack(rows):
    """rows [(n, k, dtype)] -> ([(offset, n, k, out_offset, dtype)], keys_len, out_len)."""
    out, off, oo = [], 0, 0
    for n, k, dt in rows:
        out.append((off, n, k, oo, dt))
        off += (n + 3) // 4 * 4
        oo += k
    return out, off, oo
