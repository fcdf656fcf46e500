"""GPU tests of the error and I/O surface: device-side checks reported without stats
(drtopk_workspace_check), their path through the served form, the raw binary key loader
(drtopk_file_read) feeding a top-k, and the server under a budget that holds one large request."""
import os
import subprocess

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
    yield dt
    dt.testing_set_segment_cap(0)


def to_dev(v):
    return torch.from_numpy(np.ascontiguousarray(v).view(np.int32)).cuda()


@pytest.mark.parametrize("k", [1000, 1 << 17])  # default path (alpha 8) and beta path (alpha 4)
def test_device_check_reported_without_stats(orc, dt, k):
    """A segment bound forced below the concatenation's needs: the call itself returns OK (it does
    not synchronise), drtopk_workspace_check names the capacity check, a call with stats raises, and
    once the bound is lifted the same workspace gives the oracle's result and a clean check."""
    n = 1 << 20
    v = orc.generate(0, 31, n)
    dev = to_dev(v)
    ws = torch.empty(dt.workspace_size(n, k), dtype=torch.uint8, device="cuda")
    print(f"[case] device check k={k}", flush=True)
    dt.select(dev, k, workspace=ws)
    assert dt.workspace_check(ws) == 0
    dt.testing_set_segment_cap(1)
    try:
        dt.select(dev, k, workspace=ws)
        dt.select(dev, k, workspace=ws, want_keys=False, want_idx=False)
        with pytest.raises(dt.DrTopKError) as e:
            dt.workspace_check(ws)
        assert e.value.status == 5 and "bound" in str(e.value)
        assert dt.workspace_check(ws) == 0  # read and cleared
        with pytest.raises(dt.DrTopKError) as e:
            dt.select(dev, k, workspace=ws, stats=True)
        assert e.value.status == 5
    finally:
        dt.testing_set_segment_cap(0)
    with pytest.raises(dt.DrTopKError):
        dt.workspace_check(ws)  # the stats call's failure is still pending in the sticky bits
    kth, tk, ti, _ = dt.select(dev, k, workspace=ws)
    torch.cuda.synchronize()
    assert dt.workspace_check(ws) == 0
    assert orc.verify_topk(v, k, int(kth.cpu().numpy().view(np.uint32)[0]), tk.cpu().numpy().view(np.uint32),
                           ti.cpu().numpy()) == 0


def test_batched_workspace_check(orc, dt):
    from drtopksc_amd.workloads import batch_mix, pack
    rows, kl, ol = pack(batch_mix(seed=41, nrows=64, lo=12, hi=16))
    buf = to_dev(orc.generate(0, 41, kl))
    b = dt.Batch(rows, kl, ol)
    print("[case] batched workspace check", flush=True)
    b.run(buf)
    b.run(buf)
    assert dt.workspace_check(b.ws) == 0


def test_server_reports_device_error(orc, dt):
    """A single-vector request whose device-side check fails comes back with DRTOPK_ERR_INTERNAL;
    the server keeps serving afterwards."""
    n, k = 1 << 20, 5000
    store = to_dev(orc.generate(0, 32, n))
    print("[case] server device error", flush=True)
    with dt.Server(store, budget_bytes=1 << 30, solo_n=1 << 16) as srv:
        dt.testing_set_segment_cap(1)
        try:
            t = srv.submit(0, n, k, 0)
            with pytest.raises(dt.DrTopKError) as e:
                srv.wait(t, timeout_ns=60_000_000_000)
            assert e.value.status == 5
            srv.release(t)
        finally:
            dt.testing_set_segment_cap(0)
        t = srv.submit(0, n, k, 0)
        r = srv.wait(t, timeout_ns=60_000_000_000)
        K, _, _ = orc.topk_reference(store.cpu().numpy().view(np.uint32), k)
        assert r.status == 0 and r.kth_bits == K
        srv.release(t)


def test_serve_budget_holds_one_large_request(orc, dt):
    """Budget just above one large request's scratch and outputs, small requests queued before and
    after it, every request waited for and released in submission order: all finish (small
    requests are never batched past a queued large one, so released outputs always make room)."""
    from drtopksc_amd.workloads import pack
    big_n, big_k = 1 << 22, 1 << 14
    specs = [(20000 + 101 * i, 50 + i, i % 2) for i in range(6)] + [(big_n, big_k, 0)] + \
            [(30000 + 77 * i, 20 + i, i % 2) for i in range(6)]
    rows, kl, _ = pack(specs)
    store = to_dev(orc.generate(0, 33, kl))
    need = dt.workspace_size(big_n, big_k) + 12 * big_k + 4096
    print(f"[case] serve tight budget need={need}", flush=True)
    with dt.Server(store, budget_bytes=need + (1 << 20), solo_n=1 << 20, max_batch_rows=4) as srv:
        tickets = [srv.submit(off, n, k, d) for off, n, k, oo, d in rows]
        host = store.cpu().numpy().view(np.uint32)
        for t, (off, n, k, oo, d) in zip(tickets, rows):
            r = srv.wait(t, timeout_ns=120_000_000_000)
            kb, keys, idx = srv.fetch(r, k)
            assert orc.verify_topk(host[off:off + n], k, kb, keys, idx, f32=bool(d)) == 0, (n, k, d)
            srv.release(t)


def test_file_keys_to_device(orc, dt, tmp_path):
    """Keys read from a raw 32-bit little-endian file straight into device memory (pinned staging
    in pieces) equal the file, and their top-k equals the oracle's."""
    n = (1 << 24) + 12345  # more than one 32 MiB staging piece
    raw = orc.generate(2, 34, n)  # log-normal: many ties at 2^32 - 1
    f = tmp_path / "keys.bin"
    raw.astype("<u4").tofile(f)
    dev = torch.empty(n, dtype=torch.int32, device="cuda")
    print("[case] file to device", flush=True)
    dt.file_read(str(f), out=dev)
    torch.cuda.synchronize()
    assert np.array_equal(dev.cpu().numpy().view(np.uint32), raw)
    for k in (1, 4096, 1 << 20):
        kth, tk, ti, _ = dt.select(dev, k)
        K, keys, idx = orc.topk_reference(raw, k)
        assert int(kth.cpu().numpy().view(np.uint32)[0]) == K
        assert np.array_equal(ti.cpu().numpy(), idx) and np.array_equal(tk.cpu().numpy().view(np.uint32), keys)


def test_cli_reads_file(orc, dt, tmp_path):
    """The command-line driver's --file flag (the reference's score-file argument) with --check."""
    n = 1 << 20
    raw = orc.generate(0, 35, n).view(np.int32)
    f = tmp_path / "scores.bin"
    raw.astype("<i4").tofile(f)
    exe = os.path.join(os.path.dirname(dt.lib_path()), "drtopk_topk")
    print("[case] cli --file", flush=True)
    r = subprocess.run([exe, "20", "1000", "8", "2", "--file", str(f), "--check"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Success!" in r.stdout and f"read {n} keys" in r.stdout
