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
