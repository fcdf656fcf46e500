"""GPU parity of the batched form (drtopk_batch_plan + drtopk_select_batched) through the C ABI.

Every row's k-th largest, top-k keys and indices must equal the oracle's (std::nth_element +
lowest-index extraction on that row alone), and equal what drtopk_select returns for the same
row, whatever alpha / digit width / tile the tuning table picks."""
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


def make_rows(orc, specs, seed0=1000):
    """specs [(n, k, dtype, kind)] -> packed rows, host key buffer (u32 bits), per-row ordered keys."""
    from drtopksc_amd.workloads import pack
    rows, keys_len, out_len = pack([(n, k, d) for n, k, d, _ in specs])
    buf = np.zeros(max(keys_len, 4), np.uint32)
    ordered = []
    rng = np.random.default_rng(seed0)
    for i, ((off, n, k, oo, d), (_, _, _, kind)) in enumerate(zip(rows, specs)):
        if kind == "rand":
            v = orc.generate(3 if d == 1 else 0, seed0 + i, n)
        elif kind == "ties":
            v = rng.integers(0, 8, n, dtype=np.uint64).astype(np.uint32)
            if d == 1:
                v = np.array([-1.0, 0.0, -0.0, 2.5, 1e-30, -7.0, 3.0, 2.5], np.float32)[v].view(np.uint32)
        elif kind == "equal":
            v = np.full(n, 0x3F800000 if d == 1 else 77, np.uint32)
        else:
            raise ValueError(kind)
        buf[off:off + n] = v
        ordered.append(orc.f32_to_ordered(v) if d == 1 else v)
    return rows, buf, keys_len, out_len, ordered


def check(orc, rows, ordered, kth, tk, ti, label):
    kth = kth.cpu().numpy().view(np.uint32)
    tk = tk.cpu().numpy().view(np.uint32)
    ti = ti.cpu().numpy()
    for r, (off, n, k, oo, d) in enumerate(rows):
        K, keys, idx = orc.topk_reference(ordered[r], k)
        gk = kth[r:r + 1]
        gkeys = tk[oo:oo + k]
        if d == 1:
            gk = orc.f32_to_ordered(gk)
            gkeys = orc.f32_to_ordered(gkeys)
        assert int(gk[0]) == K, (label, r, n, k, d, "kth")
        assert np.array_equal(ti[oo:oo + k], idx), (label, r, n, k, d, "idx")
        assert np.array_equal(gkeys, keys), (label, r, n, k, d, "keys")


def run_batch(dt, rows, buf, keys_len, out_len, tuning=None):
    b = dt.Batch(rows, keys_len, out_len, tuning=tuning)
    dev = torch.from_numpy(buf.view(np.int32)).cuda()
    kth, tk, ti = b.run(dev)
    torch.cuda.synchronize()
    return b, dev, kth, tk, ti


def test_batch_mix(orc, dt):
    """A scaled-down batched workload: lengths log-uniform 2^12..2^18, coin-flip key type."""
    from drtopksc_amd.workloads import batch_mix
    specs = [(n, k, d, "rand") for n, k, d in batch_mix(seed=11, nrows=160, lo=12, hi=18)]
    rows, buf, kl, ol, ordered = make_rows(orc, specs)
    print(f"[case] batch mix rows={len(rows)} keys={kl}", flush=True)
    b, dev, kth, tk, ti = run_batch(dt, rows, buf, kl, ol)
    check(orc, rows, ordered, kth, tk, ti, "mix")
    # the same batch again (the row queue and workspace are reset by the call itself)
    kth2, tk2, ti2 = b.run(dev)
    torch.cuda.synchronize()
    assert torch.equal(kth, kth2) and torch.equal(tk, tk2) and torch.equal(ti, ti2)


def test_batch_edges(orc, dt):
    specs = [(1, 1, 0, "rand"), (5, 5, 1, "rand"), (7, 3, 0, "rand"), (4096, 4096, 0, "rand"),
             (4099, 1, 1, "rand"), (70001, 1000, 0, "equal"), (70001, 1000, 1, "equal"),
             (300001, 4096, 0, "ties"), (300001, 777, 1, "ties"), (1 << 20, 1 << 16, 0, "rand"),
             (1 << 20, 20000, 1, "rand"), (100003, 25000, 0, "rand"), (2 ** 21 + 3, 4096, 1, "rand"),
             # k above the row phase's LDS concatenation cache (6144 keys): the concatenation goes
             # through the row's workspace region, here with ties and equal keys filling it
             (300001, 9000, 0, "ties"), (300003, 7000, 1, "ties"), (70001, 8000, 1, "equal")]
    rows, buf, kl, ol, ordered = make_rows(orc, specs)
    print(f"[case] batch edges rows={len(rows)}", flush=True)
    _, _, kth, tk, ti = run_batch(dt, rows, buf, kl, ol)
    check(orc, rows, ordered, kth, tk, ti, "edges")


@pytest.mark.parametrize("alpha,nb,tile", [(0, 8, 13), (1, 8, 13), (2, 4, 10), (3, 8, 11), (4, 8, 13), (5, 8, 15),
                                           (6, 8, 14), (7, 4, 13), (8, 8, 15), (9, 4, 13), (10, 8, 13), (11, 8, 15),
                                           (12, 8, 12), (13, 8, 13), (14, 8, 15), (15, 8, 15)])
def test_batch_tuning_overrides(orc, dt, alpha, nb, tile):
    """Every alpha / digit width / tile the planner accepts gives the same rows."""
    specs = [(n, k, d, kind) for n, k, d, kind in [(1 << 16, 100, 0, "rand"), (50000, 3000, 1, "rand"),
                                                   (200003, 64, 0, "ties"), (4096, 500, 1, "rand"),
                                                   (1 << 18, 1, 0, "rand"), (33333, 2048, 1, "equal"),
                                                   (1 << 18, 7000, 0, "rand")]]
    rows, buf, kl, ol, ordered = make_rows(orc, specs, seed0=2000 + alpha)
    t = dt.tuning_default()
    for d in range(2):
        for ln in range(32):
            for lk in range(32):
                t.alpha[d][ln][lk] = alpha
                t.digit_bits[d][ln][lk] = nb
                t.tile_log2[d][ln][lk] = tile
    print(f"[case] batch tuning alpha={alpha} nb={nb} tile={tile}", flush=True)
    _, _, kth, tk, ti = run_batch(dt, rows, buf, kl, ol, tuning=t)
    check(orc, rows, ordered, kth, tk, ti, f"tune a{alpha} nb{nb} t{tile}")


def test_batch_equals_single_calls(orc, dt):
    """Batched outputs are identical to looping drtopk_select over the rows."""
    from drtopksc_amd.workloads import batch_mix
    specs = [(n, k, d, "rand") for n, k, d in batch_mix(seed=12, nrows=48, lo=12, hi=20)]
    rows, buf, kl, ol, ordered = make_rows(orc, specs, seed0=3000)
    print(f"[case] batch vs single rows={len(rows)}", flush=True)
    _, dev, kth, tk, ti = run_batch(dt, rows, buf, kl, ol)
    for r, (off, n, k, oo, d) in enumerate(rows):
        x = dev[off:off + n]
        if d == 1:
            x = x.view(torch.float32)
        sk, skeys, sidx, _ = dt.select(x, k)
        assert torch.equal(sk.view(torch.int32).cpu(), kth[r:r + 1].cpu())
        assert torch.equal(skeys.view(torch.int32).cpu(), tk[oo:oo + k].cpu())
        assert torch.equal(sidx.cpu(), ti[oo:oo + k].cpu())


def test_batch_graph_replay(orc, dt):
    """The batched call captured in a HIP graph: every replay equals the eager result."""
    from drtopksc_amd.workloads import batch_mix
    specs = [(n, k, d, "rand") for n, k, d in batch_mix(seed=13, nrows=64, lo=12, hi=17)]
    rows, buf, kl, ol, ordered = make_rows(orc, specs, seed0=4000)
    print(f"[case] batch graph rows={len(rows)}", flush=True)
    b, dev, kth, tk, ti = run_batch(dt, rows, buf, kl, ol)
    out = b.alloc_out()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        b.run(dev, out=out, stream=s)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        b.run(dev, out=out, stream=torch.cuda.current_stream())
    for rep in range(3):
        for o in out:
            o.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out[0], kth) and torch.equal(out[1], tk) and torch.equal(out[2], ti), rep


@pytest.mark.timeout(900)
def test_batch_full_workload(orc, dt):
    """The whole batched workload: 4096 rows, lengths log-uniform 2^12..2^22 (2.47 G keys), coin-flip
    uint32 uniform / float32 N(0,1), k log-uniform in [1, min(n/8, 4096)] (the bench's rows and keys).
    Every row is checked exactly (k-th key, keys, indices) by the O(n) checker; 256 rows, the 16
    longest among them, also against nth_element."""
    from drtopksc_amd.workloads import batch_mix, pack
    mix = batch_mix(seed=7, nrows=4096)
    rows, keys_len, out_len = pack(mix)
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    buf = torch.randint(-2 ** 31, 2 ** 31 - 1, (keys_len,), dtype=torch.int32, device="cuda", generator=g)
    for off, n, k, oo, d in rows:
        if d == 1:
            buf[off:off + n] = torch.randn(n, device="cuda", generator=g).view(torch.int32)
    print(f"[case] batch full workload rows={len(rows)} keys={keys_len}", flush=True)
    b = dt.Batch(rows, keys_len, out_len)
    kth, tk, ti = b.run(buf)
    torch.cuda.synchronize()
    host = buf.cpu().numpy().view(np.uint32)
    kth_h = kth.cpu().numpy().view(np.uint32)
    tk_h = tk.cpu().numpy().view(np.uint32)
    ti_h = ti.cpu().numpy()
    codes = orc.verify_rows(host, rows, kth_h, tk_h, ti_h)
    bad = np.nonzero(codes)[0]
    assert bad.size == 0, [(int(r), rows[r], int(codes[r])) for r in bad[:10]]
    rng = np.random.default_rng(77)
    longest = np.argsort([-r[1] for r in rows], kind="stable")[:16]
    sample = np.unique(np.concatenate([longest, rng.choice(len(rows), 240, replace=False)]))
    for r in sample:
        off, n, k, oo, d = rows[r]
        raw = host[off:off + n]
        o = orc.f32_to_ordered(raw) if d == 1 else raw
        K, keys, idx = orc.topk_reference(o, k)
        gk = orc.f32_to_ordered(kth_h[r:r + 1]) if d == 1 else kth_h[r:r + 1]
        gkeys = orc.f32_to_ordered(tk_h[oo:oo + k]) if d == 1 else tk_h[oo:oo + k]
        assert int(gk[0]) == K and np.array_equal(ti_h[oo:oo + k], idx) and np.array_equal(gkeys, keys), (r, n, k, d)
    del buf
    torch.cuda.empty_cache()
