"""GPU test of drtopk_select_static, the entry point that allocates nothing: its scratch is the
library's file-scope device array.  Two calls in a row (the second on a smaller input, so stale
scratch contents from the first are present) must equal drtopk_select with a caller workspace,
bit for bit, on the k-th key, top-k keys and indices, for each family and both paths."""
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


def to_dev(v):
    return torch.from_numpy(np.ascontiguousarray(v).view(np.int32)).cuda()


def outputs(dev, k):
    return (torch.empty(1, dtype=dev.dtype, device="cuda"), torch.empty(k, dtype=dev.dtype, device="cuda"),
            torch.empty(k, dtype=torch.int64, device="cuda"))


# (n, k, algo, alpha): default path (auto alpha 8), beta path (auto alpha 4), the bucket and
# bitonic families, the plain path (alpha 0); each pair runs the larger input first, then the
# smaller one
CASES = [
    ((1 << 22, 1000, 0, -1), (300_001, 77, 0, -1)),
    ((1 << 22, 1 << 17, 0, -1), ((1 << 20) + 12, 1 << 15, 0, -1)),
    ((1 << 22, 5000, 1, -1), (123_457, 40, 1, -1)),
    ((1 << 21, 1024, 2, -1), (1 << 18, 64, 2, -1)),
    ((3_000_000, 1 << 16, 0, -1), (65_536, 1 << 12, 0, -1)),
    ((1 << 21, 3000, 0, 0), (100_003, 3, 0, 0)),
]


@pytest.mark.parametrize("pair", CASES, ids=lambda c: f"{c[0][0]}x{c[0][1]}-a{c[0][2]}")
def test_static_equals_workspace_call(orc, dt, pair):
    for seed, (n, k, algo, alpha) in enumerate(pair, start=61):
        v = orc.generate(0 if seed % 2 else 1, seed, n)
        dev = to_dev(v)
        p = dt.params(alpha=alpha)
        print(f"[case] static n={n} k={k} algo={algo} alpha={alpha}", flush=True)
        kth, tk, ti, st = dt.select(dev, k, algo=algo, p=p, stats=True)
        o = outputs(dev, k)
        st2 = dt.select_static(dev, k, algo=algo, p=p, stats=True, out=o)
        torch.cuda.synchronize()
        assert torch.equal(o[0], kth), (n, k, algo)
        assert torch.equal(o[1], tk), (n, k, algo)
        assert torch.equal(o[2], ti), (n, k, algo)
        for f in ("threshold", "kth_bits", "concat_size", "count_greater", "ties_taken", "selected_subranges"):
            assert st2[f] == st[f], (f, n, k, algo)
        # without stats: asynchronous, same outputs
        o3 = outputs(dev, k)
        dt.select_static(dev, k, algo=algo, p=p, out=o3)
        torch.cuda.synchronize()
        assert all(torch.equal(a, b) for a, b in zip(o3, o)), (n, k, algo)
        K, _, _ = orc.topk_reference(v, k)
        assert int(o[0].cpu().numpy().view(np.uint32)[0]) == K


def test_static_f32_and_kth_only(orc, dt):
    n, k = 1 << 20, 300
    v = orc.generate(3, 71, n)
    dev = torch.from_numpy(v.view(np.float32)).cuda()
    kth, _, _, _ = dt.select(dev, k, want_keys=False, want_idx=False)
    o = (torch.empty(1, dtype=torch.float32, device="cuda"), None, None)
    print("[case] static f32 kth only", flush=True)
    dt.select_static(dev, k, out=o)
    torch.cuda.synchronize()
    assert torch.equal(o[0], kth)


def test_static_refuses_oversize(dt):
    n, k = 1 << 26, 1 << 20
    assert dt.workspace_size(n, k) > dt.STATIC_SCRATCH_BYTES
    dev = torch.empty(n, dtype=torch.int32, device="cuda")
    o = outputs(dev, k)
    with pytest.raises(dt.DrTopKError) as e:
        dt.select_static(dev, k, out=o)
    assert e.value.status == 2 and "static scratch" in str(e.value)
    with pytest.raises(dt.DrTopKError) as e:
        dt.select_static(dev[:1 << 20], 10, p=dt.params(flags=dt.FLAGS_DEFAULT | dt.FLAG_TIMING), stats=True,
                         out=outputs(dev, 10))
    assert e.value.status == 1
