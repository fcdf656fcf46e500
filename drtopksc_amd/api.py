"""ctypes binding of libdrtopk.so (include/drtopk.h).

Device memory and streams come from torch (ROCm build); the selection itself runs in the
HIP kernels of the library.  Mirrors the reference's entry points:
  select(..., algo=ALGO_RADIX)   sample_radix_select / radix_select
                                 (baseline+filter+beta+shuffle/radixselect.cuh:1783, :1473)
  select(..., algo=ALGO_BUCKET)  sample_bucket_select / bucket_select
                                 (bucket_select/bucketselect_combined.cuh:1280, :473)
  select(..., algo=ALGO_BITONIC) bitonic_firstTopk + bitonic (bitonic/bitonic_com.cuh:610, :669)
  sample_delegates(...)          sampleMax / sampleMax_NoReduction_Beta* (radixselect.cuh:883, :967)
  Batch(...)                     many rows of uneven length in one call (no reference counterpart;
                                 the reference's MultiGPU driver loops single vectors,
                                 MultiGPU/main.cu:366-424)
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(HERE, "lib", "libdrtopk.so")

ALGO_RADIX, ALGO_BUCKET, ALGO_BITONIC = 0, 1, 2
U32, F32 = 0, 1
FLAG_FILTER, FLAG_BETA, FLAG_SHUFFLE, FLAG_TIMING = 0x1, 0x2, 0x4, 0x100
FLAGS_DEFAULT = FLAG_FILTER | FLAG_BETA | FLAG_SHUFFLE
ALPHA_AUTO, ALPHA_TUNED = -1, -2  # reference policy; measured table (radix family)

_STATUS = {0: "OK", 1: "INVALID_ARG", 2: "WORKSPACE_TOO_SMALL", 3: "UNSUPPORTED", 4: "HIP", 5: "INTERNAL",
           6: "BUDGET"}


class DrTopKError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"drtopk {_STATUS.get(status, status)}: {msg}")
        self.status = status


class Params(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_int32), ("beta", ctypes.c_int32), ("digit_bits", ctypes.c_int32),
                ("num_buckets", ctypes.c_int32), ("flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 3)]


class Stats(ctypes.Structure):
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
        d = {name: getattr(self, name) for name, _ in self._fields_ if name not in ("reserved", "phase_ms")}
        d["phase_ms"] = list(self.phase_ms)
        return d


class Row(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint64), ("n", ctypes.c_uint64), ("k", ctypes.c_uint64),
                ("out_offset", ctypes.c_uint64), ("dtype", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    """drtopk_tuning: [key type (U32, F32)][ceil log2 n][ceil log2 k] for each field."""
    _fields_ = [("alpha", ((ctypes.c_uint8 * 32) * 32) * 2), ("digit_bits", ((ctypes.c_uint8 * 32) * 32) * 2),
                ("tile_log2", ((ctypes.c_uint8 * 32) * 32) * 2)]


TUNE_DEFAULT = 0xFF


def params(alpha=-1, beta=2, digit_bits=8, num_buckets=256, flags=FLAGS_DEFAULT) -> Params:
    return Params(alpha, beta, digit_bits, num_buckets, flags, (ctypes.c_uint32 * 3)(0, 0, 0))


_lib = None


def lib_path() -> str:
    return _LIB_PATH


def lib():
    """Load libdrtopk.so; raises (no fallback) when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise DrTopKError(4, f"{_LIB_PATH} is missing: run `python -m drtopksc_amd.build`")
        L = ctypes.CDLL(_LIB_PATH)
        vp, u64, i32, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_uint32
        L.drtopk_params_default.argtypes = [ctypes.POINTER(Params)]
        L.drtopk_params_default.restype = None
        L.drtopk_workspace_size.argtypes = [ctypes.c_int, ctypes.c_int, u64, u64, ctypes.POINTER(Params),
                                            ctypes.POINTER(ctypes.c_size_t)]
        L.drtopk_workspace_size.restype = ctypes.c_int
        L.drtopk_select.argtypes = [ctypes.c_int, ctypes.c_int, vp, u64, u64, ctypes.POINTER(Params), vp,
                                    ctypes.c_size_t, vp, vp, vp, ctypes.POINTER(Stats), vp]
        L.drtopk_select.restype = ctypes.c_int
        L.drtopk_select_static.argtypes = [ctypes.c_int, ctypes.c_int, vp, u64, u64, ctypes.POINTER(Params), vp,
                                           vp, vp, ctypes.POINTER(Stats), vp]
        L.drtopk_select_static.restype = ctypes.c_int
        L.drtopk_sample_delegates.argtypes = [ctypes.c_int, vp, u64, i32, i32, u32, vp, vp, vp]
        L.drtopk_sample_delegates.restype = ctypes.c_int
        L.drtopk_batch_plan.argtypes = [ctypes.POINTER(Row), u64, u64, u64, ctypes.POINTER(Tuning), vp,
                                        ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        L.drtopk_batch_plan.restype = ctypes.c_int
        L.drtopk_select_batched.argtypes = [vp, vp, vp, vp, ctypes.c_size_t, vp, vp, vp, vp]
        L.drtopk_select_batched.restype = ctypes.c_int
        L.drtopk_tuning_default.argtypes = [ctypes.POINTER(Tuning)]
        L.drtopk_tuning_default.restype = None
        L.drtopk_workspace_check.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_uint32)]
        L.drtopk_workspace_check.restype = ctypes.c_int
        L.drtopk_testing_set_segment_cap.argtypes = [u64]
        L.drtopk_testing_set_segment_cap.restype = ctypes.c_int
        L.drtopk_file_count.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)]
        L.drtopk_file_count.restype = ctypes.c_int
        L.drtopk_file_read.argtypes = [ctypes.c_char_p, u64, u64, vp, ctypes.c_int, vp]
        L.drtopk_file_read.restype = ctypes.c_int
        L.drtopk_last_error.argtypes = []
        L.drtopk_last_error.restype = ctypes.c_char_p
        L.drtopk_version.argtypes = []
        L.drtopk_version.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise DrTopKError(rc, lib().drtopk_last_error().decode())


def _dtype_code(t) -> int:
    import torch
    if t.dtype == torch.float32:
        return F32
    if t.dtype in (torch.uint32, torch.int32):
        return U32
    raise DrTopKError(1, f"keys must be uint32 or float32, got {t.dtype}")


def workspace_size(n: int, k: int, algo: int = ALGO_RADIX, dtype: int = U32, p: Params | None = None) -> int:
    p = p or params()
    out = ctypes.c_size_t()
    _check(lib().drtopk_workspace_size(algo, dtype, n, k, ctypes.byref(p), ctypes.byref(out)))
    return int(out.value)


def _stream_ptr(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def select(keys, k: int, algo: int = ALGO_RADIX, p: Params | None = None, want_keys: bool = True,
           want_idx: bool = True, stats: bool = False, workspace=None, out=None, stream=None):
    """Top-k of a 1-D device tensor (uint32 / int32 bit patterns, or float32).

    Returns (kth, topk_keys, topk_idx, stats): kth is a 1-element tensor of the key dtype,
    topk_keys / topk_idx (int64) are in ascending index order (None when not requested),
    stats a dict when stats=True (the call then synchronises).
    """
    import torch
    if not keys.is_cuda:
        raise DrTopKError(1, "keys must be a device tensor")
    keys = keys.contiguous()
    p = p or params()
    dt = _dtype_code(keys)
    n = keys.numel()
    ws_bytes = workspace_size(n, k, algo, dt, p)
    if workspace is None or workspace.numel() < ws_bytes:
        workspace = torch.empty(max(ws_bytes, 256), dtype=torch.uint8, device=keys.device)
    if out is None:
        kth = torch.empty(1, dtype=keys.dtype, device=keys.device)
        tk = torch.empty(k, dtype=keys.dtype, device=keys.device) if want_keys else None
        ti = torch.empty(k, dtype=torch.int64, device=keys.device) if want_idx else None
    else:
        kth, tk, ti = out
    st = Stats()
    rc = lib().drtopk_select(algo, dt, ctypes.c_void_p(keys.data_ptr()), n, k, ctypes.byref(p),
                             ctypes.c_void_p(workspace.data_ptr()), workspace.numel(),
                             ctypes.c_void_p(kth.data_ptr()),
                             ctypes.c_void_p(tk.data_ptr()) if tk is not None else None,
                             ctypes.c_void_p(ti.data_ptr()) if ti is not None else None,
                             ctypes.byref(st) if stats else None, _stream_ptr(stream))
    _check(rc)
    return kth, tk, ti, (st.as_dict() if stats else None)


STATIC_SCRATCH_BYTES = 64 << 20  # DRTOPK_STATIC_SCRATCH_BYTES


def select_static(keys, k: int, algo: int = ALGO_RADIX, p: Params | None = None, stats: bool = False, out=None,
                  stream=None):
    """select() through drtopk_select_static: no workspace, the library's per-GPU static scratch
    (problems up to STATIC_SCRATCH_BYTES of workspace).  `out` = (kth, topk_keys or None,
    topk_idx or None), preallocated by the caller; returns the stats dict or None."""
    if not keys.is_cuda:
        raise DrTopKError(1, "keys must be a device tensor")
    if not keys.is_contiguous():
        raise DrTopKError(1, "keys must be contiguous")
    p = p or params()
    kth, tk, ti = out
    st = Stats()
    rc = lib().drtopk_select_static(algo, _dtype_code(keys), ctypes.c_void_p(keys.data_ptr()), keys.numel(), k,
                                    ctypes.byref(p), ctypes.c_void_p(kth.data_ptr()),
                                    ctypes.c_void_p(tk.data_ptr()) if tk is not None else None,
                                    ctypes.c_void_p(ti.data_ptr()) if ti is not None else None,
                                    ctypes.byref(st) if stats else None, _stream_ptr(stream))
    _check(rc)
    return st.as_dict() if stats else None


def sample_delegates(keys, alpha: int, beta: int = 1, shuffle: bool = False, want_pos: bool = True, stream=None):
    """Delegate pass alone: (values [beta*S] in the key dtype, positions [beta*S] uint32 or None)."""
    import torch
    keys = keys.contiguous()
    dt = _dtype_code(keys)
    n = keys.numel()
    S = (n + (1 << alpha) - 1) >> alpha
    vals = torch.empty(beta * S, dtype=keys.dtype, device=keys.device)
    pos = torch.empty(beta * S, dtype=torch.int32, device=keys.device) if want_pos else None
    rc = lib().drtopk_sample_delegates(dt, ctypes.c_void_p(keys.data_ptr()), n, alpha, beta,
                                       FLAG_SHUFFLE if shuffle else 0, ctypes.c_void_p(vals.data_ptr()),
                                       ctypes.c_void_p(pos.data_ptr()) if pos is not None else None,
                                       _stream_ptr(stream))
    _check(rc)
    return vals, pos


def workspace_check(workspace, stream=None) -> int:
    """drtopk_workspace_check: device-side error bits of every call on this workspace since the
    previous check (cleared by this call); raises DrTopKError (INTERNAL) when any is set."""
    bits = ctypes.c_uint32()
    ptr = workspace if isinstance(workspace, int) else workspace.data_ptr()
    _check(lib().drtopk_workspace_check(ctypes.c_void_p(ptr), _stream_ptr(stream), ctypes.byref(bits)))
    return int(bits.value)


def testing_set_segment_cap(cap: int) -> None:
    """Testing only: bound every later concatenation segment to `cap` entries (0: full capacity)."""
    _check(lib().drtopk_testing_set_segment_cap(int(cap)))


def file_count(path: str) -> int:
    """Whole 32-bit values in a raw binary key file."""
    n = ctypes.c_uint64()
    _check(lib().drtopk_file_count(os.fsencode(path), ctypes.byref(n)))
    return int(n.value)


def file_read(path: str, first: int = 0, count: int | None = None, out=None, stream=None):
    """Values [first, first + count) of a raw binary file of 32-bit little-endian integers (the
    reference's real-dataset format) as keys: into `out` (a numpy uint32 array or a device tensor of
    4-byte elements), or a new numpy uint32 array when out is None."""
    import numpy as np
    if count is None:
        count = file_count(path) - first
    if out is None:
        out = np.empty(count, np.uint32)
    if isinstance(out, np.ndarray):
        if out.dtype.itemsize != 4 or out.size < count or not out.flags.c_contiguous:
            raise DrTopKError(1, "out must be a contiguous array of >= count 4-byte elements")
        _check(lib().drtopk_file_read(os.fsencode(path), first, count, ctypes.c_void_p(out.ctypes.data), 0, None))
    else:
        if out.element_size() != 4 or out.numel() < count or not out.is_cuda or not out.is_contiguous():
            raise DrTopKError(1, "out must be a contiguous device tensor of >= count 4-byte elements")
        _check(lib().drtopk_file_read(os.fsencode(path), first, count, ctypes.c_void_p(out.data_ptr()), 1,
                                      _stream_ptr(stream)))
    return out


def tuning_default() -> Tuning:
    t = Tuning()
    lib().drtopk_tuning_default(ctypes.byref(t))
    return t


def tuning_from_table(table: dict) -> Tuning:
    """Tuning from the autotuner's JSON table:
    {"cells": [[dtype, log2n, log2k, alpha, digit_bits, tile_log2], ...]}."""
    t = tuning_default()
    for d, ln, lk, a, nb, tl in table.get("cells", []):
        t.alpha[d][ln][lk] = a
        t.digit_bits[d][ln][lk] = nb
        t.tile_log2[d][ln][lk] = tl
    return t


class Batch:
    """A planned batch of rows (drtopk_batch_plan), resident on the device.

    rows: sequence of (offset, n, k, out_offset, dtype) with dtype U32 or F32; offsets in
    32-bit elements of the shared key buffer (multiples of 4).  run() launches
    drtopk_select_batched on a stream and returns (kth [nrows] int32 bits, topk keys
    [out_len] int32 bits, topk indices [out_len] int64), row r's outputs at its out_offset.
    """

    def __init__(self, rows, keys_len: int, out_len: int, tuning: Tuning | None = None, device="cuda"):
        import numpy as np
        import torch
        self.nrows = len(rows)
        arr = (Row * self.nrows)()
        for i, (off, n, k, oo, dt) in enumerate(rows):
            arr[i] = Row(int(off), int(n), int(k), int(oo), int(dt), 0)
        self._rows = arr
        self.keys_len, self.out_len = int(keys_len), int(out_len)
        tp = ctypes.byref(tuning) if tuning is not None else None
        pb, wb = ctypes.c_size_t(), ctypes.c_size_t()
        _check(lib().drtopk_batch_plan(arr, self.nrows, self.keys_len, self.out_len, tp, None, ctypes.byref(pb),
                                       ctypes.byref(wb)))
        self.h_plan = np.zeros(int(pb.value), dtype=np.uint8)
        _check(lib().drtopk_batch_plan(arr, self.nrows, self.keys_len, self.out_len, tp,
                                       ctypes.c_void_p(self.h_plan.ctypes.data), ctypes.byref(pb), ctypes.byref(wb)))
        self.ws_bytes = int(wb.value)
        self.d_plan = torch.from_numpy(self.h_plan).to(device)
        self.ws = torch.empty(max(self.ws_bytes, 256), dtype=torch.uint8, device=device)

    def alloc_out(self, device="cuda"):
        import torch
        return (torch.empty(self.nrows, dtype=torch.int32, device=device),
                torch.empty(max(self.out_len, 1), dtype=torch.int32, device=device),
                torch.empty(max(self.out_len, 1), dtype=torch.int64, device=device))

    def run(self, keys, out=None, stream=None):
        if keys.numel() < self.keys_len or keys.element_size() != 4 or not keys.is_cuda:
            raise DrTopKError(1, "keys must be a device tensor of >= keys_len 32-bit elements")
        kth, tk, ti = out if out is not None else self.alloc_out(keys.device)
        _check(lib().drtopk_select_batched(ctypes.c_void_p(self.h_plan.ctypes.data),
                                           ctypes.c_void_p(self.d_plan.data_ptr()), ctypes.c_void_p(keys.data_ptr()),
                                           ctypes.c_void_p(self.ws.data_ptr()), self.ws.numel(),
                                           ctypes.c_void_p(kth.data_ptr()), ctypes.c_void_p(tk.data_ptr()),
                                           ctypes.c_void_p(ti.data_ptr()), _stream_ptr(stream)))
        return kth, tk, ti


# ---------------------------------------------------------------- serving
class ServerConfig(ctypes.Structure):
    _fields_ = [("d_store", ctypes.c_void_p), ("store_len", ctypes.c_uint64), ("budget_bytes", ctypes.c_uint64),
                ("max_inflight", ctypes.c_uint32), ("max_batch_rows", ctypes.c_uint32), ("solo_n", ctypes.c_uint64),
                ("solo_k", ctypes.c_uint64), ("tuning", ctypes.POINTER(Tuning)), ("reserved", ctypes.c_uint32 * 4)]


class Query(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint64), ("n", ctypes.c_uint64), ("k", ctypes.c_uint64),
                ("dtype", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("kth_bits", ctypes.c_uint32), ("d_topk_keys", ctypes.c_void_p),
                ("d_topk_idx", ctypes.c_void_p), ("t_submit_ns", ctypes.c_uint64), ("t_launch_ns", ctypes.c_uint64),
                ("t_done_ns", ctypes.c_uint64), ("launch_rows", ctypes.c_uint32), ("solo", ctypes.c_uint32)]


_server_bound = False


def _bind_server():
    global _server_bound
    if _server_bound:
        return
    L = lib()
    vp, u64 = ctypes.c_void_p, ctypes.c_uint64
    L.drtopk_server_config_default.argtypes = [ctypes.POINTER(ServerConfig)]
    L.drtopk_server_config_default.restype = None
    L.drtopk_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    L.drtopk_server_create.restype = ctypes.c_int
    L.drtopk_server_submit.argtypes = [vp, ctypes.POINTER(Query), ctypes.POINTER(u64)]
    L.drtopk_server_submit.restype = ctypes.c_int
    L.drtopk_server_wait.argtypes = [vp, u64, ctypes.c_int64, ctypes.POINTER(Result)]
    L.drtopk_server_wait.restype = ctypes.c_int
    L.drtopk_server_release.argtypes = [vp, u64]
    L.drtopk_server_release.restype = ctypes.c_int
    L.drtopk_server_destroy.argtypes = [vp]
    L.drtopk_server_destroy.restype = ctypes.c_int
    L.drtopk_server_replay.argtypes = [vp, ctypes.POINTER(Query), ctypes.POINTER(u64), u64, ctypes.POINTER(Result),
                                       ctypes.POINTER(u64)]
    L.drtopk_server_replay.restype = ctypes.c_int
    L.drtopk_replay_serial.argtypes = [vp, u64, ctypes.POINTER(Query), ctypes.POINTER(u64), u64,
                                       ctypes.POINTER(Result), ctypes.POINTER(u64)]
    L.drtopk_replay_serial.restype = ctypes.c_int
    _server_bound = True


def _hip():
    return ctypes.CDLL("libamdhip64.so")


def device_to_host(ptr: int, nbytes: int, dtype):
    """Copy nbytes at a device address into a new numpy array (hipMemcpy, device to host)."""
    import numpy as np
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
    if nbytes:
        rc = _hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(nbytes), 2)
        if rc != 0:
            raise DrTopKError(4, f"hipMemcpy failed ({rc})")
    return out


def _queries(qs):
    arr = (Query * len(qs))()
    for i, (off, n, k, dt) in enumerate(qs):
        arr[i] = Query(int(off), int(n), int(k), int(dt), 0)
    return arr


class Server:
    """drtopk_server_*: continuous-batching top-k over a device key store (a 32-bit device tensor)."""

    def __init__(self, store, budget_bytes=0, max_inflight=0, max_batch_rows=0, solo_n=0, solo_k=0, tuning=None):
        _bind_server()
        self.store = store
        c = ServerConfig()
        lib().drtopk_server_config_default(ctypes.byref(c))
        c.d_store = store.data_ptr()
        c.store_len = store.numel()
        c.budget_bytes = budget_bytes
        if max_inflight:
            c.max_inflight = max_inflight
        if max_batch_rows:
            c.max_batch_rows = max_batch_rows
        if solo_n:
            c.solo_n = solo_n
        if solo_k:
            c.solo_k = solo_k
        self._tuning = tuning
        if tuning is not None:
            c.tuning = ctypes.pointer(tuning)
        h = ctypes.c_void_p()
        _check(lib().drtopk_server_create(ctypes.byref(c), ctypes.byref(h)))
        self.h = h

    def submit(self, offset, n, k, dtype=U32) -> int:
        q = Query(int(offset), int(n), int(k), int(dtype), 0)
        t = ctypes.c_uint64()
        _check(lib().drtopk_server_submit(self.h, ctypes.byref(q), ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_ns: int = -1) -> Result:
        r = Result()
        _check(lib().drtopk_server_wait(self.h, ticket, timeout_ns, ctypes.byref(r)))
        return r

    def fetch(self, r: Result, k: int):
        """(kth bits, top-k keys as uint32 bits, int64 indices) of a waited, unreleased request."""
        import numpy as np
        keys = device_to_host(r.d_topk_keys, 4 * k, np.uint32)
        idx = device_to_host(r.d_topk_idx, 8 * k, np.int64)
        return r.kth_bits, keys, idx

    def release(self, ticket: int):
        _check(lib().drtopk_server_release(self.h, ticket))

    def replay(self, queries, arrivals_ns):
        """Replay [(offset, n, k, dtype)] at the given arrival offsets; returns (Result array, t0_ns)."""
        qs = _queries(queries)
        ar = (ctypes.c_uint64 * len(queries))(*[int(a) for a in arrivals_ns])
        res = (Result * len(queries))()
        t0 = ctypes.c_uint64()
        _check(lib().drtopk_server_replay(self.h, qs, ar, len(queries), res, ctypes.byref(t0)))
        return res, t0.value

    def close(self):
        if self.h:
            _check(lib().drtopk_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def replay_serial(store, queries, arrivals_ns):
    """One-at-a-time baseline over the same trace: (Result array, t0_ns)."""
    _bind_server()
    qs = _queries(queries)
    ar = (ctypes.c_uint64 * len(queries))(*[int(a) for a in arrivals_ns])
    res = (Result * len(queries))()
    t0 = ctypes.c_uint64()
    _check(lib().drtopk_replay_serial(ctypes.c_void_p(store.data_ptr()), store.numel(), qs, ar, len(queries), res,
                                      ctypes.byref(t0)))
    return res, t0.value
