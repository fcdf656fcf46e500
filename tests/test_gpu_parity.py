"""GPU parity: the HIP library against the CPU oracle, bit for bit.

Everything goes through the C ABI (libdrtopk.so via drtopksc_amd).  Each case prints its
name before touching the GPU so a failing run names it.  Outputs compared:
k-th largest, top-k keys and int64 indices (ascending index order, lowest index wins
ties), and the intermediate results the reference computes (threshold, selected /
lonely subranges, concatenation size, histogram passes).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dt():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import drtopksc_amd as dt
    dt.lib()
    return dt


def to_dev(v_u32):
    return torch.from_numpy(np.ascontiguousarray(v_u32).view(np.int32)).cuda()


def to_dev_f32(bits_u32):
    return torch.from_numpy(np.ascontiguousarray(bits_u32).view(np.float32)).cuda()


def run_gpu(dt, dev_keys, k, algo, p):
    kth, tk, ti, st = dt.select(dev_keys, k, algo=algo, p=p, stats=True)
    torch.cuda.synchronize()
    return kth.cpu().numpy(), tk.cpu().numpy(), ti.cpu().numpy(), st


STAT_KEYS = ["alpha", "beta", "path", "threshold", "kth_bits", "selected_subranges", "lonely_subranges",
             "lonely_elements", "concat_size", "count_greater", "ties_taken", "first_passes",
             "num_subranges", "num_delegates"]


def compare(orc, dt, name, okeys, dev, k, algo, flags, alpha=-1, beta=2, nb=8, B=256, f32=False):
    print(f"[case] {name} algo={algo} k={k} alpha={alpha} beta={beta} nb={nb} B={B} flags={flags} "
          f"n={okeys.size} f32={f32}", flush=True)
    po = orc.params(alpha=alpha, beta=beta, digit_bits=nb, num_buckets=B, flags=flags)
    # every listed case must be one the reference's parameter rules accept
    assert orc.resolve(algo, okeys.size, k, po) is not None, (name, "parameters rejected", k, alpha, beta)
    K, keys, idx, ost = orc.drtopk(algo, okeys, k, po, f32=f32)
    pg = dt.params(alpha=alpha, beta=beta, digit_bits=nb, num_buckets=B, flags=flags)
    gk, gkeys, gidx, gst = run_gpu(dt, dev, k, algo, pg)
    gk_bits = gk.view(np.uint32)[0]
    gkeys_bits = gkeys.view(np.uint32)
    if f32:
        gk_bits = orc.f32_to_ordered(np.array([gk_bits], np.uint32))[0]
        gkeys_bits = orc.f32_to_ordered(gkeys_bits)
    assert int(gk_bits) == K, (name, "kth", int(gk_bits), K)
    assert np.array_equal(gidx, idx), (name, "idx")
    assert np.array_equal(gkeys_bits, keys), (name, "keys")
    for s in STAT_KEYS:
        assert gst[s] == ost[s], (name, s, gst[s], ost[s])
    if algo == 0 or ost["path"] != 0:
        assert gst["second_passes"] == ost["second_passes"] or (ost["path"] == 0 and okeys.size > (1 << 22)), \
            (name, "second_passes", gst["second_passes"], ost["second_passes"])


# ------------------------------------------------------------------ delegate pass
@pytest.mark.parametrize("alpha", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14, 16])
@pytest.mark.parametrize("beta", [1, 2, 3, 4])
def test_sample_delegates(orc, dt, alpha, beta):
    if beta > (1 << alpha):
        pytest.skip("beta > subrange size")
    for n in [(1 << 18), (1 << 18) + 5, 1000]:
        if (1 << alpha) > n:
            continue
        for dist in (0, 2, 3):
            v = orc.generate(dist, 100 + alpha * 4 + beta, n)
            ok = orc.f32_to_ordered(v) if dist == 3 else v
            want_vals, want_pos = orc.sample(ok, alpha, beta)
            dev = to_dev_f32(v) if dist == 3 else to_dev(v)
            for shuffle in (False, True):
                print(f"[case] sample alpha={alpha} beta={beta} n={n} dist={dist} shuffle={shuffle}", flush=True)
                vals, pos = dt.sample_delegates(dev, alpha, beta, shuffle=shuffle)
                torch.cuda.synchronize()
                gv = vals.cpu().numpy().view(np.uint32)
                if dist == 3:
                    gv = orc.f32_to_ordered(gv)
                assert np.array_equal(gv, want_vals)
                assert np.array_equal(pos.cpu().numpy().view(np.uint32), want_pos)
                if beta == 1:  # values only: the lane-merge form without positions
                    vals, _ = dt.sample_delegates(dev, alpha, beta, shuffle=shuffle, want_pos=False)
                    torch.cuda.synchronize()
                    gv = vals.cpu().numpy().view(np.uint32)
                    if dist == 3:
                        gv = orc.f32_to_ordered(gv)
                    assert np.array_equal(gv, want_vals), "values without positions"


# ------------------------------------------------------------------ full pipeline
def datasets(orc):
    rng = np.random.default_rng(7)
    yield "uniform", orc.generate(0, 2, 1 << 20), False
    yield "normal", orc.generate(1, 3, 1 << 20), False
    yield "lognormal", orc.generate(2, 4, 1 << 20), False
    yield "ties8", rng.integers(0, 8, 300001, dtype=np.uint64).astype(np.uint32), False
    yield "allequal", np.full(70000, 123456, np.uint32), False
    yield "ragged", orc.generate(0, 9, 1000003), False
    yield "f32normal", orc.generate(3, 5, 1 << 20), True
    yield "f32uniform", orc.generate(4, 6, (1 << 19) + 77), True


@pytest.mark.parametrize("algo", [0, 1])
def test_pipeline_default_config(orc, dt, algo):
    """The reference's final configuration (FILTER|BETA|SHUFFLE, 8-bit digits, beta 2, auto alpha)."""
    for name, v, f32 in datasets(orc):
        ok = orc.f32_to_ordered(v) if f32 else v
        dev = to_dev_f32(v) if f32 else to_dev(v)
        n = v.size
        for k in [1, 2, 7, 128, 1000, 1 << 12, 1 << 15, n // 16, n // 3, n]:
            compare(orc, dt, name, ok, dev, k, algo, orc.FLAGS_DEFAULT, f32=f32)


@pytest.mark.parametrize("want", ["kth", "keys", "idx"])
def test_output_subsets(orc, dt, want):
    """k-th key only, top-k keys only, indices only: the delegate pass keeps no positions when
    neither output is asked for, and the beta path's compact delegate form (beta 2, FILTER,
    8-bit digits) then takes its position-free store; the k-th key and statistics must not move."""
    for name, v, f32 in [("uniform", orc.generate(0, 21, 1 << 20), False),
                         ("ties8", np.random.default_rng(5).integers(0, 8, 300001, dtype=np.uint64).astype(np.uint32), False),
                         ("f32normal", orc.generate(3, 22, (1 << 20) + 3), True)]:
        ok = orc.f32_to_ordered(v) if f32 else v
        dev = to_dev_f32(v) if f32 else to_dev(v)
        for k in [1000, 5000, 20000, 70000]:
            if k > v.size:
                continue
            print(f"[case] outputs={want} {name} k={k}", flush=True)
            po = orc.params(flags=orc.FLAGS_DEFAULT)
            K, keys, idx, ost = orc.drtopk(0, ok, k, po, f32=f32)
            kth, tk, ti, gst = dt.select(dev, k, p=dt.params(), want_keys=(want == "keys"),
                                         want_idx=(want == "idx"), stats=True)
            gk = kth.cpu().numpy().view(np.uint32)[:1]
            if f32:
                gk = orc.f32_to_ordered(gk)
            assert int(gk[0]) == K, (name, k, want, "kth")
            if want == "keys":
                kb = tk.cpu().numpy().view(np.uint32)
                assert np.array_equal(orc.f32_to_ordered(kb) if f32 else kb, keys), (name, k, "keys")
            if want == "idx":
                assert np.array_equal(ti.cpu().numpy(), idx), (name, k, "idx")
            for s in STAT_KEYS:
                assert gst[s] == ost[s], (name, k, want, s, gst[s], ost[s])


def test_f32_special_values(orc, dt):
    """float32 keys with -0.0 / +0.0 ties, +-inf, NaNs of both signs and several payloads, and
    denormals beside N(0,1) values: every family orders them by the sign-flipped bits (a NaN with
    the sign bit clear above +inf, one with it set below -inf), as the oracle does."""
    rng = np.random.default_rng(123)
    n = (1 << 20) + 3
    v = rng.standard_normal(n).astype(np.float32).view(np.uint32).copy()
    kind = rng.integers(0, 100, n)
    v[kind < 15] = 0x80000000                                   # -0.0
    v[(kind >= 15) & (kind < 30)] = 0                           # +0.0
    v[(kind >= 30) & (kind < 33)] = 0x7F800000                  # +inf
    v[(kind >= 33) & (kind < 36)] = 0xFF800000                  # -inf
    v[kind == 36] = 0x7FC00000                                  # quiet NaN
    v[kind == 37] = 0xFFC00001                                  # negative NaN
    v[kind == 38] = 0x7F800001                                  # signalling NaN payload
    den = kind == 39
    v[den] = rng.integers(1, 0x007FFFFF, int(den.sum()), dtype=np.uint64).astype(np.uint32) | \
        (rng.integers(0, 2, int(den.sum()), dtype=np.uint64).astype(np.uint32) << 31)
    ok = orc.f32_to_ordered(v)
    dev = to_dev_f32(v)
    for algo in (0, 1):
        for k in (1, 7, 1000, 1 << 15, n // 3, n):
            compare(orc, dt, "f32special", ok, dev, k, algo, orc.FLAGS_DEFAULT, f32=True)
    for k in (128, 4096):
        compare(orc, dt, "f32special", ok, dev, k, 2, orc.FLAGS_DEFAULT, alpha=6, f32=True)


@pytest.mark.parametrize("flags", [0, 1, 2, 3, 5, 6, 7])
@pytest.mark.parametrize("nb", [4, 8, 16])
def test_pipeline_radix_options(orc, dt, flags, nb):
    """Option surface: filter / beta / shuffle subsets and digit widths, explicit alphas."""
    rng = np.random.default_rng(flags * 10 + nb)
    v = orc.generate(0, 11, 1 << 18)
    t = rng.integers(0, 50, 1 << 17, dtype=np.uint64).astype(np.uint32)
    for name, arr in (("uniform", v), ("ties50", t)):
        dev = to_dev(arr)
        for k, alpha, beta in [(1, -1, 2), (64, 6, 2), (1000, 4, 2), (5000, 3, 3), (3000, 2, 4), (17, 9, 2),
                               (arr.size // 4, 1, 2), (300, 5, 3)]:
            compare(orc, dt, name, arr, dev, k, 0, flags, alpha=alpha, beta=beta, nb=nb)


@pytest.mark.parametrize("B", [4, 16, 256, 4096])
def test_pipeline_bucket_options(orc, dt, B):
    v = orc.generate(1, 12, 1 << 18)
    dev = to_dev(v)
    for flags in (7, 5, 1):
        for k, alpha in [(1, -1), (100, 7), (2000, 4), (9000, 3), (50000, 2)]:
            compare(orc, dt, "normal", v, dev, k, 1, flags, alpha=alpha, B=B)


def test_plain_paths(orc, dt):
    """alpha == 0: radix_select / bucket_select over V (radixselect.cuh:1473, bucketselect_combined.cuh:473)."""
    for name, v, f32 in datasets(orc):
        ok = orc.f32_to_ordered(v) if f32 else v
        dev = to_dev_f32(v) if f32 else to_dev(v)
        for algo in (0, 1):
            for k in (1, 33, v.size // 2, v.size):
                compare(orc, dt, name, ok, dev, k, algo, orc.FLAGS_DEFAULT, alpha=0, f32=f32)


def test_kth_only_and_keys_only(orc, dt):
    v = orc.generate(2, 4, 1 << 20)
    dev = to_dev(v)
    for algo in (0, 1):
        for k in (1, 999, 1 << 16):
            K, keys, idx = orc.topk_reference(v, k)
            kth, tk, ti, _ = dt.select(dev, k, algo=algo, want_keys=False, want_idx=False)
            assert int(kth.cpu().numpy().view(np.uint32)[0]) == K
            kth, tk, ti, _ = dt.select(dev, k, algo=algo, want_keys=True, want_idx=False)
            assert np.array_equal(tk.cpu().numpy().view(np.uint32), keys)


def test_repeat_calls_share_workspace(orc, dt):
    """A workspace is reusable: every call re-initialises it on the device."""
    v = orc.generate(0, 21, 1 << 20)
    dev = to_dev(v)
    ws = torch.empty(dt.workspace_size(v.size, 1 << 16), dtype=torch.uint8, device="cuda")
    for k in (1 << 16, 5, 1 << 16, 12345):
        K, keys, idx = orc.topk_reference(v, k)
        kth, tk, ti, _ = dt.select(dev, k, workspace=ws if k <= (1 << 16) else None)
        torch.cuda.synchronize()
        assert int(kth.cpu().numpy().view(np.uint32)[0]) == K
        assert np.array_equal(ti.cpu().numpy(), idx)


def test_full_size_uniform(orc, dt):
    """|V| = 2^30 uint32 uniform (seed 2): k-th against the OpenMP nth_element oracle, plus properties."""
    n = 1 << 30
    v = orc.generate(0, 2, n)
    dev = to_dev(v)
    for k in (1, 1 << 10, 1 << 20, 1 << 22, 1 << 24):
        print(f"[case] full-size uniform k={k}", flush=True)
        kth, tk, ti, st = dt.select(dev, k, stats=True)
        torch.cuda.synchronize()
        _, K, keys, idx = orc.cpu_baseline(v, k, threads=0)
        assert int(kth.cpu().numpy().view(np.uint32)[0]) == K
        assert np.array_equal(ti.cpu().numpy(), idx)
        assert np.array_equal(tk.cpu().numpy().view(np.uint32), keys)
    del dev
    torch.cuda.empty_cache()


@pytest.mark.parametrize("name", ["uniform", "ties8", "lognormal", "f32normal", "ragged"])
def test_pipeline_bitonic(orc, dt, name):
    """Bitonic family (bitonic_firstTopk + bitonic, bitonic_com.cuh:610-728): k a power of two;
    the 64 KB LDS tile network (k <= 2048), the 128 KB one (k = 4096, 8192) and the global network
    (k > 8192)."""
    data = {n: (v, f) for n, v, f in datasets(orc)}
    v, f32 = data[name]
    ok = orc.f32_to_ordered(v) if f32 else v
    dev = to_dev_f32(v) if f32 else to_dev(v)
    for k, alpha in [(1, -1), (2, 10), (16, -1), (128, 12), (1024, 8), (2048, 7), (4096, 5), (8192, 6),
                     (1 << 14, 5), (1 << 15, 4), (64, 0), (8192, 0)]:
        # the bitonic family needs more subranges than k (sampleBitonic.cu:217-221): ties8 holds
        # 300001 keys, too few for 2^alpha-key subranges at (128, 12), (8192, 6), (16384, 5) and
        # (32768, 4)
        if alpha > 0 and -(-v.size >> alpha) <= k:
            continue
        compare(orc, dt, name, ok, dev, k, 2, orc.FLAGS_DEFAULT, alpha=alpha, f32=f32)


@pytest.mark.parametrize("dist,seed", [(3, 5), (4, 6)])
def test_f32_2p28_bitonic_and_radix(orc, dt, dist, seed):
    """|V| = 2^28 float32 keys, N(0,1) and U(-1,1): bitonic top-k (alpha from 16 down to 5 as k
    grows) and radix select on the sign-flipped bits give the oracle's exact top-k."""
    n = 1 << 28
    v = orc.generate(dist, seed, n)
    ok = orc.f32_to_ordered(v)
    dev = to_dev_f32(v)
    for k in (1 << 4, 1 << 10, 1 << 16, 1 << 20):
        _, K, keys, idx = orc.cpu_baseline(ok, k, threads=0)
        for algo in (2, 0):
            print(f"[case] f32 2^28 dist={dist} k={k} algo={algo}", flush=True)
            kth, tk, ti, st = dt.select(dev, k, algo=algo, stats=True)
            torch.cuda.synchronize()
            assert int(orc.f32_to_ordered(kth.cpu().numpy().view(np.uint32))[0]) == K, (k, algo)
            assert np.array_equal(ti.cpu().numpy(), idx), (k, algo, "idx")
            assert np.array_equal(orc.f32_to_ordered(tk.cpu().numpy().view(np.uint32)), keys), (k, algo, "keys")
            if algo == 2:
                assert 5 <= st["alpha"] <= 16
    del dev
    torch.cuda.empty_cache()


def test_customized_vector(orc, dt):
    """The reference's adversarial vector (customizedDatasetGenerator.cpp, restated in
    tests/golden/customized.py) at 2^20 keys: radix, bucket (plain and Dr.Top-k) and bitonic
    against the oracle, on both sides of the boundary between the climbing prefix and the tail of
    2^32 - 1 copies; the plain bucket path refines through four levels there."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import customized as cz
    n = 1 << 20
    v = cz.customized(n)
    dev = to_dev(v)
    T = n - 1024
    for k in [1, 1024, T, T + 1, T + 300, n]:
        for algo in (0, 1):
            for alpha in (-1, 0):
                compare(orc, dt, "customized", v, dev, k, algo, orc.FLAGS_DEFAULT, alpha=alpha)
    for k in (1, 1 << 10, 1 << 14):
        compare(orc, dt, "customized", v, dev, k, 2, orc.FLAGS_DEFAULT)
