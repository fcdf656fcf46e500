"""CPU-side checks of the C ABI: the library loads, exports every declared symbol, and
rejects bad arguments with status codes (validation happens before any HIP call)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "drtopk.h")


@pytest.fixture(scope="module")
def L():
    from drtopksc_amd import api
    if not os.path.exists(api.lib_path()):
        from drtopksc_amd import build
        build.build()
    return api.lib()


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(drtopk_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    for want in ("drtopk_select", "drtopk_workspace_size", "drtopk_sample_delegates", "drtopk_params_default",
                 "drtopk_last_error", "drtopk_version"):
        assert want in names


def test_every_declared_symbol_is_exported(L):
    for name in declared_functions():
        assert hasattr(L, name), name


def test_params_default(L):
    from drtopksc_amd import api
    p = api.Params()
    L.drtopk_params_default(ctypes.byref(p))
    assert (p.alpha, p.beta, p.digit_bits, p.num_buckets, p.flags) == (-1, 2, 8, 256, 7)


def ws(L, algo, n, k, **kw):
    from drtopksc_amd import api
    p = api.params(**kw)
    out = ctypes.c_size_t()
    rc = L.drtopk_workspace_size(algo, 0, n, k, ctypes.byref(p), ctypes.byref(out))
    return rc, out.value


def test_workspace_size_validation(L):
    assert ws(L, 0, 1 << 20, 128)[0] == 0
    assert ws(L, 0, 1 << 20, 0)[0] == 1            # k == 0
    assert ws(L, 0, 1 << 20, (1 << 20) + 1)[0] == 1  # k > n
    assert ws(L, 0, 0, 1)[0] == 1                  # n == 0
    assert ws(L, 0, 1 << 20, 8, digit_bits=3)[0] == 1
    assert ws(L, 1, 1 << 20, 8, num_buckets=2)[0] == 1
    assert ws(L, 0, 1 << 20, 8, beta=5)[0] == 1
    assert ws(L, 0, 1 << 20, 8, alpha=31)[0] == 1
    assert ws(L, 0, 1 << 20, 1 << 19, alpha=8, flags=1)[0] == 1  # 4096 subranges < k
    assert ws(L, 7, 1 << 20, 8)[0] == 1            # bad algo
    assert ws(L, 0, 1 << 31, 8)[0] == 3            # n beyond 2^31-1: unsupported
    msg = L.drtopk_last_error().decode()
    assert "2^31" in msg


def test_alpha_tuned_resolves(L):
    """DRTOPK_ALPHA_TUNED (-2) takes the measured table's alpha / beta where it has a cell and the
    reference policy elsewhere; below -2 is invalid."""
    assert ws(L, 0, 1 << 30, 1 << 16, alpha=-2)[0] == 0
    assert ws(L, 0, 1 << 14, 100, alpha=-2) == ws(L, 0, 1 << 14, 100, alpha=-1)  # no cell below 2^16: the policy
    assert ws(L, 1, 1 << 30, 1 << 16, alpha=-2) == ws(L, 1, 1 << 30, 1 << 16, alpha=-1)  # bucket: the policy
    assert ws(L, 0, 1 << 20, 100, alpha=-3)[0] == 1


def test_workspace_size_grows_with_k(L):
    a = ws(L, 0, 1 << 24, 1)[1]
    b = ws(L, 0, 1 << 24, 1 << 16)[1]
    c = ws(L, 0, 1 << 24, 1 << 22)[1]
    assert 0 < a <= b <= c


def test_select_rejects_bad_pointers(L):
    from drtopksc_amd import api
    p = api.params()
    rc = L.drtopk_select(0, 0, None, 1024, 4, ctypes.byref(p), ctypes.c_void_p(256), 1 << 20, None, None, None,
                         None, None)
    assert rc == 1 and "d_keys" in L.drtopk_last_error().decode()
    rc = L.drtopk_select(0, 0, ctypes.c_void_p(4), 1024, 4, ctypes.byref(p), ctypes.c_void_p(256), 1 << 20, None,
                         None, None, None, None)
    assert rc == 1 and "aligned" in L.drtopk_last_error().decode()
    rc = L.drtopk_select(0, 0, ctypes.c_void_p(256), 1024, 4, ctypes.byref(p), ctypes.c_void_p(256), 16, None,
                         None, None, None, None)
    assert rc == 2  # workspace too small


def test_select_static_validation(L):
    """drtopk_select_static refuses, before any device work, a problem whose workspace exceeds the
    static scratch and the timing flag (it would create events)."""
    from drtopksc_amd import api
    m = re.search(r"#define DRTOPK_STATIC_SCRATCH_BYTES \((\d+)ull << (\d+)\)", open(HEADER).read())
    assert m and int(m.group(1)) << int(m.group(2)) == api.STATIC_SCRATCH_BYTES
    p = api.params()
    need = ws(L, 0, 1 << 26, 1 << 20)[1]
    assert need > api.STATIC_SCRATCH_BYTES
    rc = L.drtopk_select_static(0, 0, ctypes.c_void_p(256), 1 << 26, 1 << 20, ctypes.byref(p), None, None, None,
                                None, None)
    assert rc == 2 and "static scratch" in L.drtopk_last_error().decode()
    assert ws(L, 0, 1 << 22, 1 << 22)[1] <= api.STATIC_SCRATCH_BYTES  # every k fits at n = 2^22
    pt = api.params(flags=api.FLAGS_DEFAULT | api.FLAG_TIMING)
    rc = L.drtopk_select_static(0, 0, ctypes.c_void_p(256), 1 << 20, 10, ctypes.byref(pt), None, None, None,
                                None, None)
    assert rc == 1 and "TIMING" in L.drtopk_last_error().decode()
    rc = L.drtopk_select_static(0, 0, ctypes.c_void_p(4), 1 << 20, 10, ctypes.byref(p), None, None, None, None, None)
    assert rc == 1 and "aligned" in L.drtopk_last_error().decode()


def test_version(L):
    assert b"gfx950" in L.drtopk_version()


def test_bitonic_validation(L):
    """The bitonic family needs a power-of-two k (sampleBitonic.cu:198-202)."""
    assert ws(L, 2, 1 << 20, 1024)[0] == 0
    assert ws(L, 2, 1 << 20, 1024, alpha=0)[0] == 0
    assert ws(L, 2, 1 << 20, 1000)[0] == 1
    assert "power-of-two" in L.drtopk_last_error().decode()
    assert ws(L, 2, 1 << 20, 1 << 16, alpha=8)[0] == 1  # explicit alpha 8 leaves 4096 subranges < k
    assert ws(L, 2, 1 << 20, 1 << 16)[0] == 0           # auto alpha is reduced until S > k


def _rows(api, specs):
    arr = (api.Row * len(specs))()
    for i, (off, n, k, oo, d) in enumerate(specs):
        arr[i] = api.Row(off, n, k, oo, d, 0)
    return arr


def test_batch_plan_validation(L):
    """drtopk_batch_plan is host-only: layout and validation without a GPU."""
    from drtopksc_amd import api
    pb, wb = ctypes.c_size_t(), ctypes.c_size_t()

    def plan(specs, keys_len=1 << 20, out_len=1 << 20, tuning=None):
        rows = _rows(api, specs)
        return L.drtopk_batch_plan(rows, len(specs), keys_len, out_len,
                                   ctypes.byref(tuning) if tuning is not None else None, None, ctypes.byref(pb),
                                   ctypes.byref(wb))

    assert plan([(0, 5000, 10, 0, 0), (5000, 7000, 100, 10, 1)]) == 0 and pb.value > 128 and wb.value >= 256
    assert plan([(2, 5000, 10, 0, 0)]) == 1 and "multiple of 4" in L.drtopk_last_error().decode()
    assert plan([(0, 5000, 0, 0, 0)]) == 1           # k == 0
    assert plan([(0, 5000, 5001, 0, 0)]) == 1        # k > n
    assert plan([(0, 5000, 10, 0, 7)]) == 1          # bad dtype
    assert plan([(0, 5000, 10, 0, 0)], keys_len=4000) == 1   # row outside the key buffer
    assert plan([(0, 5000, 10, 0, 0)], out_len=5) == 1       # outputs outside the output arrays
    t = api.tuning_default()
    t.digit_bits[0][13][4] = 5
    assert plan([(0, 5000, 10, 0, 0)], tuning=t) == 1 and "digit_bits" in L.drtopk_last_error().decode()
    t = api.tuning_default()
    t.tile_log2[1][13][4] = 9
    assert plan([(0, 5000, 10, 0, 0)], tuning=t) == 1


def test_batch_plan_layout(L):
    """The plan blob: header, one 64-byte record per row, one record per delegate-pass tile, the row
    order."""
    import struct

    import numpy as np

    from drtopksc_amd import api
    specs = [(0, 70000, 100, 0, 0), (70000, 4, 4, 100, 1), (70004, 1 << 18, 4096, 104, 0)]
    rows = _rows(api, specs)
    pb, wb = ctypes.c_size_t(), ctypes.c_size_t()
    assert L.drtopk_batch_plan(rows, 3, 1 << 20, 4200, None, None, ctypes.byref(pb), ctypes.byref(wb)) == 0
    buf = np.zeros(pb.value, np.uint8)
    assert L.drtopk_batch_plan(rows, 3, 1 << 20, 4200, None, ctypes.c_void_p(buf.ctypes.data), ctypes.byref(pb),
                               ctypes.byref(wb)) == 0
    hdr = struct.unpack_from("<12Q", buf.tobytes(), 0)
    magic, nrows, nchunks, o_rows, o_crow, o_order = hdr[:6]
    assert nrows == 3 and o_rows == 128
    recs = [struct.unpack_from("<6Q3I4B", buf.tobytes(), o_rows + 64 * i) for i in range(3)]
    tiles = 0
    for (key_off, n, k, out_off, deleg_off, sel_off, S, cb, nc, alpha, f32, nb, tl), (off, nn, kk, oo, d) in zip(
            recs, specs):
        assert (key_off, n, k, out_off, f32) == (off, nn, kk, oo, d)
        if alpha:
            assert S == -(-n // (1 << alpha)) and S >= k and alpha <= tl
            assert cb == tiles and nc == -(-n // (1 << tl))
            tiles += nc
        assert nb in (4, 8)
    assert nchunks == tiles
    # one 16-byte record per delegate-pass tile: first key | alpha << 48 | f32 << 54, then first
    # delegate | (length - 1) << 40
    td = np.frombuffer(buf.tobytes(), np.uint64, 2 * nchunks, o_crow).reshape(nchunks, 2)
    for rec in recs:
        key_off, n, alpha, f32, tl, deleg_off, cb, nc = rec[0], rec[1], rec[9], rec[10], rec[12], rec[4], rec[7], rec[8]
        for c in range(nc):
            start = c << tl
            src, dst = int(td[cb + c, 0]), int(td[cb + c, 1])
            assert src & ((1 << 48) - 1) == key_off + start and (src >> 48) & 63 == alpha and (src >> 54) & 1 == f32
            assert dst & ((1 << 40) - 1) == deleg_off + (start >> alpha)
            assert (dst >> 40) + 1 == min(1 << tl, n - start)
    order = np.frombuffer(buf.tobytes(), np.uint32, 3, o_order)
    assert sorted(order.tolist()) == [0, 1, 2] and order[0] == 2  # largest row first


def test_server_validation(L):
    """drtopk_server_create refuses bad configurations before touching the GPU."""
    from drtopksc_amd import api
    api._bind_server()
    h = ctypes.c_void_p()
    c = api.ServerConfig()
    L.drtopk_server_config_default(ctypes.byref(c))
    assert (c.max_inflight, c.max_batch_rows, c.solo_n, c.solo_k) == (8, 4096, 1 << 25, 4096)
    assert L.drtopk_server_create(ctypes.byref(c), ctypes.byref(h)) == 1  # no key store
    c.d_store = 4
    assert L.drtopk_server_create(ctypes.byref(c), ctypes.byref(h)) == 1  # misaligned
    c.d_store = 256
    c.reserved[0] = 1
    assert L.drtopk_server_create(ctypes.byref(c), ctypes.byref(h)) == 1


def test_workspace_check_validation(L):
    """drtopk_workspace_check validates its pointer before any HIP call."""
    bits = ctypes.c_uint32(7)
    assert L.drtopk_workspace_check(None, None, ctypes.byref(bits)) == 1
    assert bits.value == 0
    assert L.drtopk_workspace_check(ctypes.c_void_p(0x1008), None, None) == 1  # not 256-byte aligned
    assert "aligned" in L.drtopk_last_error().decode()


def test_file_loader_host(L, tmp_path):
    """Raw 32-bit little-endian key files (the reference's real-dataset format): count, whole and
    partial reads into host memory, the same bits as written (negative int32 values become keys
    >= 2^31, as the reference's int -> unsigned store does), and short files rejected."""
    import numpy as np

    from drtopksc_amd import api
    rng = np.random.default_rng(5)
    vals = rng.integers(-2 ** 31, 2 ** 31, 100003, dtype=np.int64).astype("<i4")
    vals[:4] = [-1, 0, 2 ** 31 - 1, -2 ** 31]
    f = tmp_path / "scores.bin"
    f.write_bytes(vals.tobytes() + b"\x01\x02")  # a trailing partial value is not counted
    assert api.file_count(str(f)) == vals.size
    got = api.file_read(str(f))
    assert np.array_equal(got, vals.view(np.uint32))
    assert got[0] == 0xFFFFFFFF and got[3] == 0x80000000
    part = api.file_read(str(f), first=17, count=1000)
    assert np.array_equal(part, vals.view(np.uint32)[17:1017])
    with pytest.raises(api.DrTopKError) as e:
        api.file_read(str(f), first=vals.size - 5, count=6)
    assert e.value.status == 1 and "holds" in str(e.value)
    with pytest.raises(api.DrTopKError):
        api.file_count(str(tmp_path / "missing.bin"))
    assert L.drtopk_file_read(None, 0, 1, None, 0, None) == 1
