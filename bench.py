"""Benchmark: Dr.Top-k on one MI355X at |V| = 2^30 (uint32 uniform, seed 2).

One step = the reference's headline sweep (baseline+filter+beta+shuffle/run.bash:14-22):
top-k of the same device-resident 2^30-key vector for every k = 2^0 .. 2^24, in the
reference's final configuration (FILTER|BETA|SHUFFLE, 8-bit digits, beta 2, auto alpha),
each call returning the k-th key, the top-k keys and their int64 indices.
value = keys examined per second over the sweep (25 * 2^30 / step time), in Gelem/s.

Printed as one JSON line; also carries the per-k times, the HBM roofline of the
dominant kernel (the delegate pass) and the CPU baseline timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "top-k Gelem/s + %HBM peak at |V|=2^30; batched req/s + p50/p99 latency"


def algorithmic_bytes(st: dict, n: int, k: int) -> int:
    """4n (the delegate pass reads V once) + second-pass traffic: delegate write, first-top-k
    passes over delegates, reads of the selected subranges, concatenation write, second-top-k
    passes over it, and k keys + int64 indices out."""
    D = st["num_delegates"]
    C = st["concat_size"]
    L = 1 << st["alpha"]
    p1 = max(0, st["first_passes"] - 1) + 1     # unfused digit passes + the range-selection pass
    p2 = max(0, st["second_passes"] - 1) + 1    # unfused digit passes + the extraction pass
    if st["path"] == 2:
        sel_read = st["selected_subranges"] * L * 4
    else:
        sel_read = st["selected_subranges"] * L * 4
    return 4 * n + 4 * D + 4 * D * p1 + sel_read + 8 * C + 4 * C * p2 + 12 * k


def batched_leg(dt, torch, steps: int, warmup: int, nrows: int = 4096, graph_loop: bool = True):
    """Batched workload: nrows rows, lengths log-uniform 2^12..2^22, coin-flip uint32 uniform /
    float32 N(0,1), k log-uniform in [1, min(n/8, 4096)], top-k keys + int64 indices per row.
    One batched call (two launches) against a loop of single-vector calls over the same rows,
    eagerly and captured in a HIP graph; outputs must be identical."""
    from drtopksc_amd.workloads import batch_mix, pack
    mix = batch_mix(seed=7, nrows=nrows)
    rows, keys_len, out_len = pack(mix)
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    buf = torch.randint(-2 ** 31, 2 ** 31 - 1, (keys_len,), dtype=torch.int32, device="cuda", generator=g)
    for off, n, k, oo, d in rows:
        if d == 1:
            buf[off:off + n] = torch.randn(n, device="cuda", generator=g).view(torch.int32)
    b = dt.Batch(rows, keys_len, out_len)
    out = b.alloc_out()
    stream = torch.cuda.current_stream()
    for _ in range(max(1, warmup)):
        b.run(buf, out=out, stream=stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        b.run(buf, out=out, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps

    # loop of single-vector calls over the same rows (shared workspace, preallocated outputs)
    views = []
    for off, n, k, oo, d in rows:
        x = buf[off:off + n]
        views.append((x.view(torch.float32) if d == 1 else x, k, oo))
    wsb = max(dt.workspace_size(n, k, dt.ALGO_RADIX, d) for _, n, k, _, d in rows)
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    lk = torch.empty(len(rows), dtype=torch.int32, device="cuda")
    lkeys = torch.empty(max(out_len, 1), dtype=torch.int32, device="cuda")
    lidx = torch.empty(max(out_len, 1), dtype=torch.int64, device="cuda")

    def loop(st):
        for r, (x, k, oo) in enumerate(views):
            o = (lk[r:r + 1].view(x.dtype), lkeys[oo:oo + k].view(x.dtype), lidx[oo:oo + k])
            dt.select(x, k, workspace=ws, out=o, stream=st)

    loop(stream)
    torch.cuda.synchronize()
    identical = bool(torch.equal(lk, out[0]) and torch.equal(lkeys[:out_len], out[1][:out_len])
                     and torch.equal(lidx[:out_len], out[2][:out_len]))
    t0 = time.perf_counter()
    e0.record(stream)
    loop(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    loop_ms = e0.elapsed_time(e1)
    loop_wall_ms = (time.perf_counter() - t0) * 1e3
    loop_graph_ms = None
    if graph_loop:
        s = torch.cuda.Stream()
        s.wait_stream(stream)
        gr = torch.cuda.CUDAGraph()
        with torch.cuda.graph(gr, stream=s):
            loop(torch.cuda.current_stream())
        gr.replay()
        torch.cuda.synchronize()
        e0.record(stream)
        gr.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        loop_graph_ms = e0.elapsed_time(e1)
        del gr
    total = sum(n for _, n, _, _, _ in rows)
    best_loop = min(x for x in (loop_ms, loop_graph_ms) if x is not None)
    res = {"workload": f"{len(rows)} rows, n log-uniform 2^12..2^22, uint32 uniform / float32 N(0,1) coin flip, "
                       "k log-uniform [1, min(n/8, 4096)], keys + int64 indices out",
           "rows": len(rows), "keys": total, "ms": round(ms, 4), "req_s": round(len(rows) / (ms * 1e-3), 1),
           "gelem_s": round(total / (ms * 1e-3) / 1e9, 1),
           "hbm_frac_4n": round(4 * total / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
           "loop_ms": round(loop_ms, 3), "loop_wall_ms": round(loop_wall_ms, 3),
           "loop_graph_ms": round(loop_graph_ms, 3) if loop_graph_ms is not None else None,
           "speedup_vs_loop": round(best_loop / ms, 2), "outputs_identical": identical}
    del buf, ws
    torch.cuda.empty_cache()
    return res


def served_leg(dt, torch, nreq: int = 2000, rate: float = 20000.0):
    """Served workload: nreq requests, n log-uniform 2^10..2^26, coin-flip uint32 uniform /
    float32 N(0,1), k log-uniform in [1, min(n/4, 2^16)], Poisson arrivals at `rate`, keys resident
    in one device store, 64 GiB scratch arena.  The same trace is served (a) by the continuous-
    batching server and (b) one request at a time by single-vector calls; both also saturated
    (every request arriving at t = 0)."""
    from drtopksc_amd.workloads import pack, serve_mix
    mix = serve_mix(seed=8, nreq=nreq, rate_per_s=rate)
    rows, keys_len, _ = pack([(n, k, d) for _, n, k, d in mix])
    g = torch.Generator(device="cuda")
    g.manual_seed(8)
    store = torch.randint(-2 ** 31, 2 ** 31 - 1, (keys_len,), dtype=torch.int32, device="cuda", generator=g)
    for off, n, k, oo, d in rows:
        if d == 1:
            store[off:off + n] = torch.randn(n, device="cuda", generator=g).view(torch.int32)
    torch.cuda.synchronize()
    qs = [(off, n, k, d) for off, n, k, oo, d in rows]
    arr = [int(t * 1e9) for t, _, _, _ in mix]
    sat = [0] * len(qs)
    total = sum(n for _, n, _, _ in qs)

    def summ(res, t0):
        lat = np.array([(r.t_done_ns - r.t_submit_ns) / 1e3 for r in res])
        span = (max(r.t_done_ns for r in res) - t0) / 1e9
        assert all(r.status == 0 for r in res)
        return {"req_s": round(len(res) / span, 1), "gelem_s": round(total / span / 1e9, 2),
                "p50_us": round(float(np.percentile(lat, 50)), 1), "p99_us": round(float(np.percentile(lat, 99)), 1)}

    with dt.Server(store) as srv:
        srv.replay(qs[:64], [0] * 64)  # warm-up: code objects, pinned staging
        r_rate, t0 = srv.replay(qs, arr)
        served_rate = summ(r_rate, t0)
        r_sat, t1 = srv.replay(qs, sat)
        served_sat = summ(r_sat, t1)
    s_rate, t2 = dt.replay_serial(store, qs, arr)
    serial_rate = summ(s_rate, t2)
    s_sat, t3 = dt.replay_serial(store, qs, sat)
    serial_sat = summ(s_sat, t3)
    identical = all(r_rate[i].kth_bits == s_rate[i].kth_bits == r_sat[i].kth_bits for i in range(len(qs)))
    del store
    torch.cuda.empty_cache()
    return {"workload": f"{len(qs)} requests, n log-uniform 2^10..2^26, uint32 uniform / float32 N(0,1) coin flip, "
                        f"k log-uniform [1, min(n/4, 2^16)], Poisson arrivals at {rate:.0f}/s, device-resident keys, "
                        "64 GiB scratch arena",
            "requests": len(qs), "keys": total, "rate_per_s": rate,
            "served_at_rate": served_rate, "served_saturated": served_sat,
            "serial_at_rate": serial_rate, "serial_saturated": serial_sat,
            "speedup_vs_serial": round(served_sat["req_s"] / serial_sat["req_s"], 2),
            "kth_identical": identical}


def side_legs_only(args):
    """Run only the batched and/or served legs (for profiling); the headline value is null."""
    import torch

    import drtopksc_amd as dt
    out = {"metric": METRIC, "value": None, "unit": "Gelem/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "note": "sweep leg skipped (--legs)"}
    if not args.no_batched:
        out["batched"] = batched_leg(dt, torch, args.steps, args.warmup, nrows=args.batch_rows)
    if not args.no_served:
        out["served"] = served_leg(dt, torch, nreq=args.served_requests)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=30)
    ap.add_argument("--kmax-log2", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-rows", type=int, default=4096)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--legs", default="sweep,batched,served",
                    help="comma list of sweep, batched, served (profiling runs pick one)")
    ap.add_argument("--served-requests", type=int, default=2000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU exchange is "
              "design-only in this build", file=sys.stderr)
        sys.exit(2)
    legs = set(args.legs.split(","))
    args.no_batched = args.no_batched or "batched" not in legs
    args.no_served = args.no_served or "served" not in legs
    if "sweep" not in legs:
        return side_legs_only(args)

    import torch

    import drtopksc_amd as dt
    from oracle import oracle

    n = 1 << args.log2n
    ks = [1 << j for j in range(args.kmax_log2 + 1)]
    t0 = time.time()
    v = oracle.generate(oracle.UNIFORM_U32, 2, n)
    print(f"[bench] generated {n} keys in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dev = torch.from_numpy(v.view(np.int32)).cuda()
    p = dt.params()
    ws_bytes = max(dt.workspace_size(n, k, dt.ALGO_RADIX, 0, p) for k in ks)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    kmax = max(ks)
    kth = torch.empty(1, dtype=dev.dtype, device="cuda")
    okeys = torch.empty(kmax, dtype=dev.dtype, device="cuda")
    oidx = torch.empty(kmax, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()

    def call(k, stats=False, timing=False):
        pp = dt.params(flags=dt.FLAGS_DEFAULT | (dt.FLAG_TIMING if timing else 0))
        return dt.select(dev, k, p=pp, stats=stats, workspace=ws, out=(kth, okeys[:k], oidx[:k]), stream=stream)

    def sweep():
        for k in ks:
            call(k)

    # per-k stats + phase times (outside the timed region) and a correctness spot check
    per_k = []
    for k in ks:
        _, _, _, st = call(k, stats=True, timing=True)
        for _ in range(2):
            _, _, _, st = call(k, stats=True, timing=True)
        per_k.append(st)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        sweep()
    torch.cuda.synchronize()
    # per-k wall times with HIP events on the launch stream
    k_ms = []
    for k in ks:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record(stream)
        for _ in range(reps):
            call(k)
        e1.record(stream)
        e1.synchronize()
        k_ms.append(e0.elapsed_time(e1) / reps)

    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_start = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        sweep()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_step = e0.elapsed_time(e1) / args.steps
    value = len(ks) * n / (ms_step * 1e-3) / 1e9

    # roofline of the dominant kernel: the delegate pass (phase 0), k = 2^10 (alpha 11)
    rows = []
    for k, st, ms in zip(ks, per_k, k_ms):
        B = algorithmic_bytes(st, n, k)
        rows.append({"k": k, "alpha": st["alpha"], "path": st["path"], "ms": round(ms, 4),
                     "gelem_s": round(n / (ms * 1e-3) / 1e9, 1),
                     "hbm_frac": round(B / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                     "hbm_frac_4n": round(4 * n / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                     "concat": st["concat_size"], "phase_ms": [round(x, 4) for x in st["phase_ms"]]})
    samp = [st["phase_ms"][0] for st in per_k]
    ref_i = ks.index(1 << 10) if (1 << 10) in ks else 0
    st_ref = per_k[ref_i]
    samp_bytes = 4 * n + 4 * st_ref["num_delegates"]
    samp_ms = samp[ref_i]
    achieved = samp_bytes / (samp_ms * 1e-3) / 1e9
    # HBM bytes per launch of this kernel from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on the
    # same workload, FETCH_SIZE scaled by 2.000 and WRITE_SIZE by 1.000 (the factors a 4 GiB
    # device-to-device copy reads back): profiles/delegate_pass_2p30_k1024.md
    traffic = (2 * 2_097_192 + 4_399) * 1024 if args.log2n == 30 else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "delegate pass (k_sample_waves, alpha=11)", "bytes_per_launch": samp_bytes,
                "avg_ms": round(samp_ms, 4)}

    cpu = None
    if not args.no_cpu:
        kc = 1 << 10
        threads = min(16, os.cpu_count() or 1)
        reps = 3  # about 1.5 s of wall time on 16 threads: ~20 s of CPU work
        runs = []
        for _ in range(reps):
            t_one, K, _, _ = oracle.cpu_baseline(v, kc, threads=threads)
            runs.append(t_one)
        t_cpu = sorted(runs)[reps // 2]
        kth_gpu = None
        call(kc)
        torch.cuda.synchronize()
        kth_gpu = int(kth.cpu().numpy().view(np.uint32)[0])
        assert kth_gpu == K, (kth_gpu, K)
        cpu = {"value": round(n / t_cpu / 1e9, 3), "unit": "Gelem/s", "cores": threads, "kind": "port",
               "sample": f"full 2^{args.log2n} uniform vector, k=2^10 (nth_element per chunk + merge + sort + "
                         f"index pass), median of {reps} runs of {t_cpu:.2f}s on {threads} threads"}

    batched = None
    if not args.no_batched:
        del dev, ws
        torch.cuda.empty_cache()
        batched = batched_leg(dt, torch, args.steps, args.warmup, nrows=args.batch_rows)
        print(f"[bench] batched {batched}", file=sys.stderr, flush=True)
    served = None
    if not args.no_served:
        if args.no_batched:
            del dev, ws
            torch.cuda.empty_cache()
        served = served_leg(dt, torch, nreq=args.served_requests)
        print(f"[bench] served {served}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Gelem/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"single-vector top-k sweep, |V|=2^{args.log2n} uint32 uniform (mt19937 seed 2), "
                               f"k=2^0..2^{args.kmax_log2}, keys+int64 indices out",
                   "n": n, "k_sweep": [ks[0], ks[-1]], "algo": "radix", "flags": "FILTER|BETA|SHUFFLE",
                   "digit_bits": 8, "beta": 2, "alpha": "auto (reference policy)"},
        "roofline": roofline, "cpu_baseline": cpu, "batched": batched, "served": served,
        "min_hbm_frac": min(r["hbm_frac"] for r in rows),
        "per_k": rows, "wall_s": round(wall, 3),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
