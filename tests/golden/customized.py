"""The reference's adversarial "customized" key vector, restated (customizedDatasetGenerator.cpp:7-62).

Construction (customizedDatasetGenerator.cpp:25-52): with max = 2^32-1 and 256 buckets, level
j spreads 256 keys at the bucket midpoints of [min_j, max]: diff_j = (max - min_j) / 256 and
key i = min_j + (2i+1) * diff_j / 2; the next level starts at the last of them
(min_{j+1} = key 255, :47-48).  It stops when the range has no room left (diff == 0, min >= max,
:31-32).  Every remaining position holds max (:55-57).  So 1024 keys (four levels) climb ever
closer to max, then about n - 1024 copies of max follow.

Each level packs the next one into the top bucket of the last, which is what drives the bucket
family through one refinement level per level of the vector
(bucket_select/CustomizedDistribution/bucketselect_combined.cuh:1011).

Two departures from the reference's text, both to remove undefined behaviour it has:
- the midpoint uses the bucket index i, as its commented-out line :41 does, not the running
  position `pos`: pos * diff overflows 32 bits from the second level on (:42);
- no key is read before it is written (the reference reads data[-1] at pos 1, :36, and data[pos]
  after a break, :47).

Arithmetic is exact (Python integers); every key fits 32 bits.
"""
from __future__ import annotations

import numpy as np

KEY_MAX = 0xFFFFFFFF
NBUCKET = 256


def levels() -> list[int]:
    """The climbing prefix: 256 keys per level while the range has room."""
    lo, out = 0, []
    while lo < KEY_MAX:
        diff = (KEY_MAX - lo) // NBUCKET
        if diff == 0:
            break
        lvl = [lo + (2 * i + 1) * diff // 2 for i in range(NBUCKET)]
        out += lvl
        lo = lvl[-1]
    return out


def customized(n: int) -> np.ndarray:
    """uint32[n]: the prefix (cut to n) followed by n - len(prefix) copies of 2^32 - 1."""
    pre = np.array(levels(), dtype=np.uint64).astype(np.uint32)
    v = np.full(n, KEY_MAX, dtype=np.uint32)
    m = min(n, pre.size)
    v[:m] = pre[:m]
    return v
