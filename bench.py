"""Benchmark: Dr.Top-k on one MI355X at |V| = 2^30 (uint32 uniform, seed 2).

One step = the reference's headline sweep (baseline+filter+beta+shuffle/run.bash:14-22):
top-k of the same device-resident 2^30-key vector for every k = 2^0 .. 2^24, in the
reference's final configuration (FILTER|BETA|SHUFFLE, 8-bit digits, beta 2, auto alpha),
each call returning the k-th key, the top-k keys and their int64 indices.
value = keys examined per second over the sweep (25 * 2^30 / step time), in Gelem/s.

Printed as one JSON line; besides the headline it carries the per-k times with two %HBM
figures, the HBM roofline of the dominant kernel (the delegate pass), the other single-vector
workloads (bucket family, normal and log-normal keys, float32 2^28 with the bitonic family,
the plain radix select, the option ablation, the bitonic reference points), the batched and
served forms, and the OpenMP CPU baseline timed on this host for each key distribution.
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

# Published reference times (V100S), set beside the matching lines below.
REF_BUCKET_UNIFORM_MS = {  # bucket_select/Uniform_Unsigned__N_30_SOKBucket.csv:1-125, mean of 5 runs per k
    1: 7.77, 2: 7.71, 4: 7.65, 8: 7.18, 16: 6.95, 32: 6.98, 64: 7.11, 128: 6.88, 256: 7.29, 512: 7.24,
    1024: 7.09, 2048: 7.44, 4096: 7.28, 8192: 7.31, 16384: 7.71, 32768: 7.79, 65536: 8.72, 131072: 8.77,
    262144: 10.55, 524288: 10.53, 1048576: 14.18, 2097152: 16.31, 4194304: 26.00, 8388608: 27.02,
    16777216: 31.41}
REF_RADIX_UNIFORM_MS = {  # baseline+filter+beta+shuffle/optimum_FICO, "Total Time" lines 831..2532
    1: 7.16, 2: 7.18, 4: 6.56, 8: 6.59, 16: 6.41, 32: 6.43, 64: 6.33, 128: 6.33, 256: 6.30, 512: 6.45,
    1024: 6.41, 2048: 6.49, 4096: 6.49, 8192: 7.32, 16384: 6.97, 32768: 6.96, 65536: 7.66, 131072: 7.80,
    262144: 9.07, 524288: 9.70, 1048576: 12.70, 2097152: 14.05, 4194304: 18.65, 8388608: 20.53,
    16777216: 33.69}
REF_PLAIN_RADIX_2P26_MS = 4.5   # baseline/onlyRadixSelect/U_OnlyRadiXSelect_2^30.csv:1-26 (flat in k)
REF_BITONIC_2P29_K128_MS = 6.66  # bitonic/testV100.csv:1 (alpha 12)
REF_BITONIC_LARGEK_MS = {4096: 16.35, 8192: 23.80, 16384: 63.11}  # largerK/testV100.csv:48,:50,:52 (alpha 10)
REF_ABLATION_2P29_K2P24_MS = {  # baseline{,+filter,+filter+shuffle}/U_OptimumFICO_RADIX.csv:25
    "baseline": 53.29, "filter": 36.81, "filter+shuffle": 28.46}


def bytes_build(st: dict, n: int, k: int) -> int:
    """This build's algorithmic bytes for one top-k: what each pass must move at least.
    - 4n: the delegate pass reads V once;
    - 4D: the delegate write;
    - 4D per first-top-k pass that reads the delegates (digit passes after the first, whose
      histogram rides on the delegate pass, plus the range-selection / concatenation pass);
    - 4 * 2^alpha per selected subrange (its keys are read);
    - 8C: the concatenation writes each entry's key and its 32-bit index;
    - 4C per second-top-k pass (digit passes after the first, plus the extraction);
    - 12k: k keys and k int64 indices out."""
    D, C, L = st["num_delegates"], st["concat_size"], 1 << st["alpha"]
    p1 = max(0, st["first_passes"] - 1) + 1
    p2 = max(0, st["second_passes"] - 1) + 1
    return 4 * n + 4 * D + 4 * D * p1 + 4 * st["selected_subranges"] * L + 8 * C + 4 * C * p2 + 12 * k


def bytes_survey(st: dict, n: int, k: int) -> int:
    """The byte formula as the hot-path scope states it: 4n + 4D + 4D*p1 (p1 = digit passes
    + 1 selection pass) + 4*S_sel*2^alpha + 4C + 4C*p2 (p2 = second-top-k passes) + 12k."""
    D, C, L = st["num_delegates"], st["concat_size"], 1 << st["alpha"]
    p1 = st["first_passes"] + 1
    p2 = st["second_passes"]
    return 4 * n + 4 * D + 4 * D * p1 + 4 * st["selected_subranges"] * L + 4 * C + 4 * C * p2 + 12 * k


def checks_clean(dt, ws) -> bool:
    """drtopk_workspace_check over every call made on the workspace: True when no device-side
    check failed."""
    try:
        return dt.workspace_check(ws) == 0
    except dt.DrTopKError:
        return False


def frac(nbytes: float, ms: float) -> float:
    return round(nbytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)


class Runner:
    """Single-vector calls on one stream with one shared workspace and output buffers."""

    def __init__(self, dt, torch, kmax: int):
        self.dt, self.torch = dt, torch
        self.stream = torch.cuda.current_stream()
        self.ws = torch.empty(256, dtype=torch.uint8, device="cuda")
        self.kmax = kmax
        self.okeys = torch.empty(kmax, dtype=torch.int32, device="cuda")
        self.oidx = torch.empty(kmax, dtype=torch.int64, device="cuda")
        self.kth = torch.empty(1, dtype=torch.int32, device="cuda")

    def reserve(self, nbytes: int):
        if self.ws.numel() < nbytes:
            del self.ws
            self.torch.cuda.empty_cache()
            self.ws = self.torch.empty(nbytes, dtype=self.torch.uint8, device="cuda")

    def call(self, dev, k, algo=0, p=None, stats=False, timing=False):
        dt = self.dt
        p = p or dt.params()
        if timing:
            p = dt.params(p.alpha, p.beta, p.digit_bits, p.num_buckets, p.flags | dt.FLAG_TIMING)
        o = (self.kth.view(dev.dtype), self.okeys[:k].view(dev.dtype), self.oidx[:k])
        return self.dt.select(dev, k, algo=algo, p=p, stats=stats, workspace=self.ws, out=o, stream=self.stream)

    def time_ms(self, fn, reps: int = 5) -> float:
        torch = self.torch
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        e0.record(self.stream)
        for _ in range(reps):
            fn()
        e1.record(self.stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    def sweep(self, dev, ks, algo=0, p=None, reps=5, ref=None):
        """Per-k stats (one synchronised call) and event-timed ms (reps calls), with both %HBM
        figures and the conservative 4n one."""
        dt = self.dt
        n = dev.numel()
        p = p or dt.params()
        self.reserve(max(dt.workspace_size(n, k, algo, dt.F32 if dev.dtype == self.torch.float32 else dt.U32, p)
                         for k in ks))
        rows = []
        for k in ks:
            _, _, _, st = self.call(dev, k, algo, p, stats=True)
            ms = self.time_ms(lambda: self.call(dev, k, algo, p), reps)
            r = {"k": k, "alpha": st["alpha"], "path": st["path"], "ms": round(ms, 4),
                 "gelem_s": round(n / (ms * 1e-3) / 1e9, 1), "hbm_frac": frac(bytes_build(st, n, k), ms),
                 "hbm_frac_survey": frac(bytes_survey(st, n, k), ms), "hbm_frac_4n": frac(4 * n, ms),
                 "concat": st["concat_size"]}
            if ref and k in ref:
                r["ref_v100s_ms"] = ref[k]
            rows.append(r)
        return rows

    def kth_u32(self) -> int:
        return int(self.kth.cpu().numpy().view(np.uint32)[0])


def cpu_figures(oracle, okeys: np.ndarray, k: int, P: int, reps: int = 5, one_thread_n: int = 1 << 27):
    """OpenMP nth_element baseline on P threads over the whole vector and on 1 thread over its
    first one_thread_n keys: each the median of `reps` runs after one warm-up run."""
    def med(a, threads):
        oracle.cpu_baseline(a, k, threads=threads)  # warm-up
        ts = sorted(oracle.cpu_baseline(a, k, threads=threads)[0] for _ in range(reps))
        return ts[reps // 2]
    t_all = med(okeys, P)
    sub = okeys[:min(one_thread_n, okeys.size)]
    t_one = med(sub, 1)
    _, K, keys, idx = oracle.cpu_baseline(okeys, k, threads=P)
    return {"threads": P, "gelem_s": round(okeys.size / t_all / 1e9, 3), "s": round(t_all, 4),
            "one_thread_gelem_s": round(sub.size / t_one / 1e9, 3), "one_thread_s": round(t_one, 4),
            "one_thread_sample": f"first {sub.size} keys", "runs": reps, "k": k}, (K, keys, idx)


def spot_check(R: Runner, dev, k, algo, want, f32=False):
    """k-th key, keys and indices of one call against the CPU baseline's exact result."""
    K, keys, idx = want
    R.call(dev, k, algo)
    R.torch.cuda.synchronize()
    got_k = R.kth_u32()
    if f32:  # compare in the ordered domain the CPU baseline works in
        b = got_k
        got_k = b ^ (0xFFFFFFFF if b >> 31 else 0x80000000)
        gk = R.okeys[:k].cpu().numpy().view(np.uint32)
        gk = gk ^ np.where(gk >> 31, np.uint32(0xFFFFFFFF), np.uint32(0x80000000))
    else:
        gk = R.okeys[:k].cpu().numpy().view(np.uint32)
    gi = R.oidx[:k].cpu().numpy()
    return bool(got_k == K and np.array_equal(gk, keys) and np.array_equal(gi, idx))


def configs_leg(dt, torch, oracle, R: Runner, v_uni: np.ndarray, dev_uni, P: int, ks, cpu: bool):
    """Every other single-vector workload, outside the timed region."""
    out = {}
    t0 = time.time()
    kfew = [1, 1 << 10, 1 << 16, 1 << 20, 1 << 22, 1 << 24]
    # the headline sweep with the measured (alpha, beta) table instead of the reference policy
    rows = R.sweep(dev_uni, ks, p=dt.params(alpha=dt.ALPHA_TUNED))
    tot = sum(r["ms"] for r in rows)
    out["uniform_2p30_tuned"] = {"alpha": "measured table (DRTOPK_ALPHA_TUNED)", "sweep_ms": round(tot, 4),
                                 "gelem_s": round(len(ks) * dev_uni.numel() / (tot * 1e-3) / 1e9, 1), "rows": rows}
    # bucket family on the headline vector
    out["uniform_2p30_bucket"] = {"algo": "bucket", "rows": R.sweep(dev_uni, ks, algo=dt.ALGO_BUCKET,
                                                                    ref=REF_BUCKET_UNIFORM_MS)}
    # option ablation at 2^29 (the reference's ablation size), k = 2^24 and others
    n29 = 1 << 29
    d29 = dev_uni[:n29]
    abl = {}
    for name, fl in (("baseline", 0), ("filter", dt.FLAG_FILTER), ("filter+beta", dt.FLAG_FILTER | dt.FLAG_BETA),
                     ("filter+shuffle", dt.FLAG_FILTER | dt.FLAG_SHUFFLE), ("filter+beta+shuffle", dt.FLAGS_DEFAULT)):
        rows = R.sweep(d29, [1 << 10, 1 << 20, 1 << 22, 1 << 24], p=dt.params(flags=fl), reps=3)
        abl[name] = {"rows": rows}
        if name in REF_ABLATION_2P29_K2P24_MS:
            abl[name]["ref_v100s_ms_k2p24"] = REF_ABLATION_2P29_K2P24_MS[name]
    out["ablation_2p29"] = abl
    # plain radix select (alpha 0) at 2^26
    d26 = dev_uni[:1 << 26]
    out["plain_radix_2p26"] = {"alpha": 0, "ref_v100s_ms": REF_PLAIN_RADIX_2P26_MS,
                               "rows": R.sweep(d26, [1, 1 << 4, 1 << 10, 1 << 16, 1 << 20, 1 << 24],
                                               p=dt.params(alpha=0))}
    # bitonic reference points: 2^29 k = 128 at alpha 12; 2^30 k = 2^12..2^14 at alpha 10
    bt = R.sweep(d29, [128], algo=dt.ALGO_BITONIC, p=dt.params(alpha=12))
    bt[0]["ref_v100s_ms"] = REF_BITONIC_2P29_K128_MS
    bl = R.sweep(dev_uni, [4096, 8192, 16384], algo=dt.ALGO_BITONIC, p=dt.params(alpha=10),
                 ref=REF_BITONIC_LARGEK_MS)
    out["bitonic_uniform"] = {"2p29_k128_alpha12": bt, "2p30_alpha10": bl}
    print(f"[bench] uniform configs done ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    # other key distributions at 2^30: radix and bucket sweeps, CPU baseline, exact spot check
    for name, dist, seed in (("normal_2p30", oracle.NORMAL_U32, 3), ("lognormal_2p30", oracle.LOGNORMAL_U32, 4)):
        t1 = time.time()
        v = oracle.generate(dist, seed, 1 << 30)
        dev = torch.from_numpy(v.view(np.int32)).cuda()
        ent = {"dist": name, "seed": seed,
               "radix": R.sweep(dev, ks), "bucket": R.sweep(dev, kfew, algo=dt.ALGO_BUCKET)}
        if cpu:
            ent["cpu"], want = cpu_figures(oracle, v, 1 << 10, P)
            ent["spot_check_k1024_exact"] = spot_check(R, dev, 1 << 10, 0, want)
        out[name] = ent
        del dev, v
        torch.cuda.empty_cache()
        print(f"[bench] {name} done ({time.time() - t1:.1f}s)", file=sys.stderr, flush=True)
    # float32 2^28: N(0,1) and U(-1,1); radix and bitonic at k = 2^4, 2^10, 2^16, 2^20
    kf = [1 << 4, 1 << 10, 1 << 16, 1 << 20]
    for name, dist, seed in (("f32_normal_2p28", oracle.NORMAL_F32, 5), ("f32_uniform_2p28", oracle.UNIFORM_F32, 6)):
        t1 = time.time()
        v = oracle.generate(dist, seed, 1 << 28)
        dev = torch.from_numpy(v.view(np.float32)).cuda()
        ent = {"dist": name, "seed": seed, "radix": R.sweep(dev, kf), "bitonic": R.sweep(dev, kf, algo=dt.ALGO_BITONIC)}
        if cpu:
            ov = oracle.f32_to_ordered(v)
            ent["cpu"], want = cpu_figures(oracle, ov, 1 << 10, P)
            ent["spot_check_k1024_exact"] = spot_check(R, dev, 1 << 10, 0, want, f32=True)
            del ov
        out[name] = ent
        del dev, v
        torch.cuda.empty_cache()
        print(f"[bench] {name} done ({time.time() - t1:.1f}s)", file=sys.stderr, flush=True)
    return out


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
    device_ok = checks_clean(dt, b.ws)

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
           "speedup_vs_loop": round(best_loop / ms, 2), "outputs_identical": identical,
           "device_checks_clean": device_ok}
    del buf, ws
    torch.cuda.empty_cache()
    return res


def served_leg(dt, torch, nreq: int = 2000, rate: float = 20000.0):
    """Served workload: nreq requests, n log-uniform 2^10..2^26, coin-flip uint32 uniform /
    float32 N(0,1), k log-uniform in [1, min(n/4, 2^16)], Poisson arrivals at `rate`, keys resident
    in one device store, 64 GiB scratch arena.  The same trace is served (a) by the continuous-
    batching server and (b) one request at a time by single-vector calls; both also saturated
    (every request arriving at t = 0), three times each, the median run reported."""
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

    def median_of(runs):  # the run with the median throughput, and every run's throughput
        runs = sorted(runs, key=lambda x: x["req_s"])
        return dict(runs[len(runs) // 2], runs_req_s=[x["req_s"] for x in runs])

    with dt.Server(store) as srv:
        srv.replay(qs[:64], [0] * 64)  # warm-up: code objects, pinned staging
        r_rate, t0 = srv.replay(qs, arr)
        served_rate = summ(r_rate, t0)
        sat_runs = []
        for _ in range(3):
            r_sat, t1 = srv.replay(qs, sat)
            sat_runs.append(summ(r_sat, t1))
        served_sat = median_of(sat_runs)
    # launches in flight (worker threads, one stream each): this host gives a process 4 hardware
    # queues (GPU_MAX_HW_QUEUES), so streams beyond 4 share queues
    inflight = {}
    for m in (2, 4, 8, 12):
        with dt.Server(store, max_inflight=m) as srv:
            srv.replay(qs[:64], [0] * 64)
            r_m, tm = srv.replay(qs, sat)
            inflight[m] = summ(r_m, tm)
    s_rate, t2 = dt.replay_serial(store, qs, arr)
    serial_rate = summ(s_rate, t2)
    ser_runs = []
    for _ in range(3):
        s_sat, t3 = dt.replay_serial(store, qs, sat)
        ser_runs.append(summ(s_sat, t3))
    serial_sat = median_of(ser_runs)
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
            "saturated_by_max_inflight": inflight,
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


def host_info(oracle):
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"nproc": os.cpu_count(), "affinity_cpus": aff, "omp_max_threads": oracle.num_threads(),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


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
    ap.add_argument("--legs", default="sweep,configs,batched,served",
                    help="comma list of sweep, configs, batched, served (profiling runs pick one)")
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

    P = oracle.num_threads()  # OpenMP threads of this host share (OMP_NUM_THREADS), fixed before any run
    host = host_info(oracle)
    n = 1 << args.log2n
    ks = [1 << j for j in range(args.kmax_log2 + 1)]
    t0 = time.time()
    v = oracle.generate(oracle.UNIFORM_U32, 2, n)
    print(f"[bench] generated {n} keys in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dev = torch.from_numpy(v.view(np.int32)).cuda()
    R = Runner(dt, torch, max(ks))
    p = dt.params()
    R.reserve(max(dt.workspace_size(n, k, dt.ALGO_RADIX, 0, p) for k in ks))
    stream = R.stream

    def sweep():
        for k in ks:
            R.call(dev, k)

    # per-k stats + phase times (outside the timed region)
    per_k = []
    for k in ks:
        for _ in range(3):
            _, _, _, st = R.call(dev, k, stats=True, timing=True)
        per_k.append(st)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        sweep()
    torch.cuda.synchronize()
    # per-k times with HIP events on the launch stream
    k_ms = [R.time_ms(lambda: R.call(dev, k), 5) for k in ks]

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
    device_ok = checks_clean(dt, R.ws)  # device-side checks of every call above

    rows = []
    for k, st, ms in zip(ks, per_k, k_ms):
        r = {"k": k, "alpha": st["alpha"], "path": st["path"], "ms": round(ms, 4),
             "gelem_s": round(n / (ms * 1e-3) / 1e9, 1),
             "hbm_frac": frac(bytes_build(st, n, k), ms), "hbm_frac_survey": frac(bytes_survey(st, n, k), ms),
             "hbm_frac_4n": frac(4 * n, ms), "concat": st["concat_size"],
             "phase_ms": [round(x, 4) for x in st["phase_ms"]]}
        if args.log2n == 30 and k in REF_RADIX_UNIFORM_MS:
            r["ref_v100s_ms"] = REF_RADIX_UNIFORM_MS[k]
        rows.append(r)
    # roofline of the dominant kernel: the delegate pass (phase 0), k = 2^10 (alpha 11)
    ref_i = ks.index(1 << 10) if (1 << 10) in ks else 0
    st_ref = per_k[ref_i]
    samp_bytes = 4 * n + 4 * st_ref["num_delegates"]
    # its average duration over 8 launches (HIP events around it on the call's stream)
    samp = [R.call(dev, ks[ref_i], stats=True, timing=True)[3]["phase_ms"][0] for _ in range(8)]
    samp_ms = sum(samp) / len(samp)
    achieved = samp_bytes / (samp_ms * 1e-3) / 1e9
    # HBM bytes per launch of this kernel from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on the
    # same workload, FETCH_SIZE scaled by 2.000 and WRITE_SIZE by 1.000 (the factors a 4 GiB
    # device-to-device copy reads back): profiles/delegate_pass_2p30_k1024.md
    traffic = (2 * 2_097_212 + 2_251) * 1024 if args.log2n == 30 else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "delegate pass (k_sample_waves, alpha=11)", "bytes_per_launch": samp_bytes,
                "avg_ms": round(samp_ms, 4)}

    cpu = None
    cpu_uniform = None
    if not args.no_cpu:
        kc = 1 << 10
        cpu_uniform, want = cpu_figures(oracle, v, kc, P)
        assert spot_check(R, dev, kc, 0, want), "GPU result differs from the CPU baseline at k=2^10"
        cpu = {"value": cpu_uniform["gelem_s"], "unit": "Gelem/s", "cores": P, "kind": "port",
               "sample": f"full 2^{args.log2n} uniform vector, k=2^10 (nth_element per chunk + merge + sort + "
                         f"index pass), median of {cpu_uniform['runs']} runs after a warm-up on {P} OpenMP threads "
                         f"({cpu_uniform['s']}s each); 1 thread over the first 2^27 keys: "
                         f"{cpu_uniform['one_thread_gelem_s']} Gelem/s"}

    configs = None
    if "configs" in legs:
        configs = configs_leg(dt, torch, oracle, R, v, dev, P, ks, cpu=not args.no_cpu)
        if cpu_uniform is not None:
            configs["uniform_2p30_cpu"] = cpu_uniform
    del dev, R
    torch.cuda.empty_cache()

    batched = None
    if not args.no_batched:
        batched = batched_leg(dt, torch, args.steps, args.warmup, nrows=args.batch_rows)
        print(f"[bench] batched {batched}", file=sys.stderr, flush=True)
    served = None
    if not args.no_served:
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
        "roofline": roofline, "cpu_baseline": cpu, "host": host,
        "min_hbm_frac": min(r["hbm_frac"] for r in rows),
        "min_hbm_frac_survey": min(r["hbm_frac_survey"] for r in rows),
        "device_checks_clean": device_ok,
        "per_k": rows, "configs": configs, "batched": batched, "served": served, "wall_s": round(wall, 3),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
