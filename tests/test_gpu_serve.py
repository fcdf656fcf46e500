"""GPU tests of the served form (drtopk_server_*): every request's result in the mix equals the
oracle's and drtopk_select's for the same request alone; admission by the arena budget."""
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


def build_store(orc, specs, seed0=5000):
    """specs [(n, k, dtype)] -> (device store, queries [(offset, n, k, dtype)], ordered host rows)."""
    from drtopksc_amd.workloads import pack
    rows, keys_len, _ = pack(specs)
    buf = np.zeros(max(keys_len, 4), np.uint32)
    ordered = []
    for i, (off, n, k, oo, d) in enumerate(rows):
        v = orc.generate(3 if d == 1 else 0, seed0 + i, n)
        buf[off:off + n] = v
        ordered.append(orc.f32_to_ordered(v) if d == 1 else v)
    store = torch.from_numpy(buf.view(np.int32)).cuda()
    return store, [(off, n, k, d) for off, n, k, oo, d in rows], ordered


def check_one(orc, dt, srv, store, q, ordered, r, label):
    off, n, k, d = q
    kth, keys, idx = srv.fetch(r, k)
    K, rkeys, ridx = orc.topk_reference(ordered, k)
    kk = np.array([kth], np.uint32)
    if d == 1:
        kk = orc.f32_to_ordered(kk)
        keys = orc.f32_to_ordered(keys)
    assert int(kk[0]) == K, (label, q, "kth")
    assert np.array_equal(idx, ridx), (label, q, "idx")
    assert np.array_equal(keys, rkeys), (label, q, "keys")
    # the same request alone through the single-vector entry point
    x = store[off:off + n]
    sk, skeys, sidx, _ = dt.select(x.view(torch.float32) if d == 1 else x, k)
    assert int(sk.view(torch.int32).cpu().numpy().view(np.uint32)[0]) == kth, (label, q, "alone kth")
    assert np.array_equal(sidx.cpu().numpy(), idx), (label, q, "alone idx")


def test_mix_equals_alone(orc, dt):
    from drtopksc_amd.workloads import serve_mix
    specs = [(n, k, d) for _, n, k, d in serve_mix(seed=21, nreq=120, lo=10, hi=19, kmax=1 << 13)]
    specs += [(1 << 20, 1 << 15, 0), (3 << 18, 100, 1), (1 << 19, 5000, 0)]
    store, qs, ordered = build_store(orc, specs)
    print(f"[case] serve mix requests={len(qs)}", flush=True)
    with dt.Server(store, budget_bytes=1 << 30, solo_n=1 << 18, solo_k=4096) as srv:
        tickets = [srv.submit(*q) for q in qs]
        results = [srv.wait(t) for t in tickets]
        rows = [r.launch_rows for r in results]
        assert any(r.solo for r in results) and any(r.launch_rows > 1 for r in results), rows
        for q, o, r, t in zip(qs, ordered, results, tickets):
            check_one(orc, dt, srv, store, q, o, r, "mix")
            assert r.t_submit_ns <= r.t_launch_ns <= r.t_done_ns
            srv.release(t)


def test_budget_admission(orc, dt):
    specs = [(50000 + 997 * i, 1 + 31 * i, i % 2) for i in range(40)]
    store, qs, ordered = build_store(orc, specs, seed0=6000)
    print("[case] serve small budget", flush=True)
    # 4 MiB holds one launch's scratch and outputs at a time: requests queue but all finish
    with dt.Server(store, budget_bytes=4 << 20, max_batch_rows=8) as srv:
        tickets = [srv.submit(*q) for q in qs]
        for q, o, t in zip(qs, ordered, tickets):
            r = srv.wait(t)
            check_one(orc, dt, srv, store, q, o, r, "budget")
            srv.release(t)
        with pytest.raises(dt.DrTopKError) as e:
            srv.submit(0, store.numel() - 4, store.numel() // 2, 0)
        assert e.value.status == 6  # DRTOPK_ERR_BUDGET


def test_budget_large_request_in_order(orc, dt):
    """An arena just above one large request's need, small requests queued before and after it,
    every request waited for and released in submission order.  The rows of a batch share one
    output block, freed once all of them are released: a batch that took small requests from
    behind the large one without leaving it room would keep it from ever fitting.  A finite wait
    turns such a stall into a failure instead of a hang."""
    small_a = [(40000 + 1009 * i, 1 + 7 * i, i % 2) for i in range(8)]
    large = [(1 << 20, 3000, 0)]
    small_b = [(41000 + 1013 * i, 2 + 5 * i, (i + 1) % 2) for i in range(8)]
    store, qs, ordered = build_store(orc, small_a + large + small_b, seed0=9000)

    def a256(x):
        return (x + 255) // 256 * 256
    # the server's single-vector calls use the reference configuration with the measured table
    pt = dt.params(alpha=dt.ALPHA_TUNED)
    need = a256(dt.workspace_size(1 << 20, 3000, dt.ALGO_RADIX, 0, pt)) + a256(a256(4) + a256(4 * 3000) + 8 * 3000)
    budget = need + (64 << 10)
    print(f"[case] serve large request in order budget={budget}", flush=True)
    with dt.Server(store, budget_bytes=budget, solo_n=1 << 18, max_batch_rows=16) as srv:
        tickets = [srv.submit(*q) for q in qs]
        for q, o, t in zip(qs, ordered, tickets):
            r = srv.wait(t, timeout_ns=60 * 10 ** 9)
            check_one(orc, dt, srv, store, q, o, r, "large-in-order")
            srv.release(t)


def test_batch_takes_small_requests_behind_large(orc, dt):
    """With room in the arena a batch takes the small requests queued behind a large one.  One
    stream; 24 single-vector requests of 2^26 keys keep it busy (milliseconds) while the replay
    queues 8 small requests, a large one and 8 more small ones (microseconds), so when the stream
    frees up the queue holds all 17: the 16 small requests go in one launch, then the large one
    runs alone."""
    small_a = [(30000 + 1009 * i, 1 + 7 * i, i % 2) for i in range(8)]
    large = [(1 << 20, 3000, 0)]
    small_b = [(31000 + 1013 * i, 2 + 5 * i, (i + 1) % 2) for i in range(8)]
    busy = [(1 << 26, 1000, 0)]
    store, qs, ordered = build_store(orc, small_a + large + small_b + busy, seed0=9500)
    nbusy = 24
    seq = [qs[-1]] * nbusy + qs[:-1]
    print("[case] serve batch behind large request", flush=True)
    with dt.Server(store, budget_bytes=1 << 32, max_inflight=1, solo_n=1 << 18, max_batch_rows=64) as srv:
        res, _ = srv.replay(seq, [0] * len(seq))
    rs = list(res)[nbusy:]
    assert all(r.status == 0 for r in res)
    for i, (q, o, r) in enumerate(zip(qs[:-1], ordered[:-1], rs)):
        K, _, _ = orc.topk_reference(o, q[2])
        kk = np.array([r.kth_bits], np.uint32)
        if q[3] == 1:
            kk = orc.f32_to_ordered(kk)
        assert int(kk[0]) == K, (i, q)
        if i == 8:
            assert r.solo == 1 and r.launch_rows == 1, (i, r.solo, r.launch_rows)
        else:
            assert r.solo == 0 and r.launch_rows == 16, (i, r.solo, r.launch_rows)
    assert rs[8].t_launch_ns >= rs[9].t_launch_ns == rs[0].t_launch_ns


def test_replay_matches_serial(orc, dt):
    from drtopksc_amd.workloads import serve_mix
    mix = serve_mix(seed=22, nreq=80, lo=10, hi=20, kmax=1 << 12, rate_per_s=20000.0)
    specs = [(n, k, d) for _, n, k, d in mix]
    store, qs, ordered = build_store(orc, specs, seed0=7000)
    arr = [int(t * 1e9) for t, _, _, _ in mix]
    print(f"[case] serve replay requests={len(qs)}", flush=True)
    with dt.Server(store, budget_bytes=1 << 30) as srv:
        res, t0 = srv.replay(qs, arr)
    ser, t1 = dt.replay_serial(store, qs, arr)
    for i, (q, o) in enumerate(zip(qs, ordered)):
        K, _, _ = orc.topk_reference(o, q[2])
        a = np.array([res[i].kth_bits, ser[i].kth_bits], np.uint32)
        if q[3] == 1:
            a = orc.f32_to_ordered(a)
        assert res[i].status == 0 and ser[i].status == 0
        assert int(a[0]) == K and int(a[1]) == K, (i, q)


@pytest.mark.timeout(900)
def test_serve_full_trace(orc, dt):
    """The whole served workload: 2000 requests, n log-uniform 2^10..2^26 (10.7 G keys in one device
    store), coin-flip uint32 uniform / float32 N(0,1), k log-uniform in [1, min(n/4, 2^16)], default
    server (64 GiB arena).  All requests are queued at once, then waited for in order; every result
    (k-th key, keys, indices) is checked exactly by the O(n) checker, 256 of them (the 16 longest
    among them) also against nth_element.  The same trace replayed with Poisson arrivals at 20 K/s
    returns the same k-th keys."""
    from drtopksc_amd.workloads import pack, serve_mix
    mix = serve_mix(seed=8, nreq=2000, rate_per_s=20000.0)
    rows, keys_len, _ = pack([(n, k, d) for _, n, k, d in mix])
    g = torch.Generator(device="cuda")
    g.manual_seed(8)
    store = torch.randint(-2 ** 31, 2 ** 31 - 1, (keys_len,), dtype=torch.int32, device="cuda", generator=g)
    for off, n, k, oo, d in rows:
        if d == 1:
            store[off:off + n] = torch.randn(n, device="cuda", generator=g).view(torch.int32)
    torch.cuda.synchronize()
    qs = [(off, n, k, d) for off, n, k, oo, d in rows]
    out_len = sum(q[2] for q in qs)
    kth = np.zeros(len(qs), np.uint32)
    okeys = np.zeros(out_len, np.uint32)
    oidx = np.zeros(out_len, np.int64)
    print(f"[case] serve full trace requests={len(qs)} keys={keys_len}", flush=True)
    with dt.Server(store) as srv:
        tickets = [srv.submit(*q) for q in qs]
        for i, (t, (off, n, k, oo, d)) in enumerate(zip(tickets, rows)):
            r = srv.wait(t)
            kb, keys, idx = srv.fetch(r, k)
            kth[i] = kb
            okeys[oo:oo + k] = keys
            oidx[oo:oo + k] = idx
            srv.release(t)
        arr = [int(t * 1e9) for t, _, _, _ in mix]
        res, _ = srv.replay(qs, arr)
        assert all(res[i].status == 0 and res[i].kth_bits == kth[i] for i in range(len(qs)))
    host = store.cpu().numpy().view(np.uint32)
    codes = orc.verify_rows(host, rows, kth, okeys, oidx)
    bad = np.nonzero(codes)[0]
    assert bad.size == 0, [(int(r), rows[r], int(codes[r])) for r in bad[:10]]
    rng = np.random.default_rng(88)
    longest = np.argsort([-r[1] for r in rows], kind="stable")[:16]
    sample = np.unique(np.concatenate([longest, rng.choice(len(rows), 240, replace=False)]))
    for r in sample:
        off, n, k, oo, d = rows[r]
        raw = host[off:off + n]
        o = orc.f32_to_ordered(raw) if d == 1 else raw
        K, keys, idx = orc.topk_reference(o, k)
        gk = orc.f32_to_ordered(kth[r:r + 1]) if d == 1 else kth[r:r + 1]
        gkeys = orc.f32_to_ordered(okeys[oo:oo + k]) if d == 1 else okeys[oo:oo + k]
        assert int(gk[0]) == K and np.array_equal(oidx[oo:oo + k], idx) and np.array_equal(gkeys, keys), (r, n, k, d)
    del store, host
    torch.cuda.empty_cache()
