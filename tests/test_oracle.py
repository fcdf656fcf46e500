"""Pins the CPU oracle (oracle/drtopk_oracle.cpp) before anything is checked against it.

The reference ships no golden vectors (its only check is the disabled `Enabletest`
assert, baseline+filter+beta+shuffle/main.cu:281-296: sort a host copy and compare
vec1[n-k]).  So every oracle function is pinned here by something it does not share
code with: the C++ standard's mt19937 known answer, numpy sorts (the reference's own
check), brute force on tiny inputs, closed forms of the threshold rules, and
distribution moments for the generators.
"""
import math

import numpy as np
import pytest

ALGOS = [0, 1, 2]  # radix, bucket, bitonic


def np_topk(v, k):
    """numpy brute force: k-th largest (ascending rank n-k) and lowest-index top-k set."""
    K = int(np.sort(v, kind="stable")[v.size - k])
    gt = np.nonzero(v > K)[0]
    eq = np.nonzero(v == K)[0][: k - gt.size]
    idx = np.sort(np.concatenate([gt, eq]))
    return K, v[idx], idx


# ---------------------------------------------------------------- generators
def test_mt19937_known_answer(orc):
    # [rand.predef]: the 10000th consecutive invocation of a default-constructed mt19937
    # (seed 5489) produces 4123659995.
    assert int(orc.mt19937(5489, 10000)[-1]) == 4123659995


def test_uniform_is_raw_stream(orc):
    assert np.array_equal(orc.generate(orc.UNIFORM_U32, 7, 1000), orc.mt19937(7, 1000))


def test_normal_lognormal_moments(orc):
    n = 1 << 20
    x = orc.generate(orc.NORMAL_U32, 3, n).astype(np.float64)
    assert abs(x.mean() - 2.0**31) < 4 * 2.0**28 / math.sqrt(n)
    assert abs(x.std() / 2.0**28 - 1.0) < 0.01
    y = orc.generate(orc.LOGNORMAL_U32, 4, n)
    sat = float(np.mean(y == 0xFFFFFFFF))
    # P(Z > (ln(2^32) - 20)/2) for Z ~ N(0,1)
    z = (32 * math.log(2.0) - 20.0) / 2.0
    p = 0.5 * math.erfc(z / math.sqrt(2.0))
    assert abs(sat - p) < 5 * math.sqrt(p * (1 - p) / n)
    # median of lognormal(20, 2) is e^20
    assert abs(math.log(float(np.median(y))) - 20.0) < 0.01


def test_float_generators(orc):
    n = 1 << 20
    z = orc.generate(orc.NORMAL_F32, 5, n).view(np.float32).astype(np.float64)
    assert abs(z.mean()) < 4 / math.sqrt(n)
    assert abs(z.std() - 1.0) < 0.01
    u = orc.generate(orc.UNIFORM_F32, 6, n).view(np.float32)
    assert u.min() >= -1.0 and u.max() < 1.0
    assert abs(float(u.astype(np.float64).mean())) < 4 * 0.58 / math.sqrt(n)
    # values lie on the 2^-23 grid
    g = (u.astype(np.float64) + 1.0) * 2.0**23
    assert np.all(g == np.floor(g))


def test_generators_are_seeded(orc):
    for d in range(5):
        a, b = orc.generate(d, 11, 4097), orc.generate(d, 11, 4097)
        c = orc.generate(d, 12, 4097)
        assert np.array_equal(a, b) and not np.array_equal(a, c)


# ---------------------------------------------------------------- float ordering
def test_f32_ordering_is_monotone(orc):
    rng = np.random.default_rng(0)
    f = np.concatenate([rng.standard_normal(5000).astype(np.float32),
                        np.array([0.0, -0.0, np.inf, -np.inf, 1e-45, -1e-45, 3.4e38, -3.4e38], np.float32)])
    o = orc.f32_to_ordered(f.view(np.uint32))
    order_o = np.argsort(o, kind="stable")
    fs = f[order_o]
    assert np.all(fs[1:] >= fs[:-1])
    # -0 < +0 strictly in the key order
    neg0, pos0 = orc.f32_to_ordered(np.array([0x80000000, 0], np.uint32))
    assert neg0 < pos0
    assert np.array_equal(orc.ordered_to_f32(o), f.view(np.uint32))


# ---------------------------------------------------------------- closed-form oracle
@pytest.mark.parametrize("n,k", [(1, 1), (17, 1), (17, 17), (1000, 10), (4096, 4096), (5000, 2500)])
def test_topk_reference_vs_numpy(orc, n, k):
    rng = np.random.default_rng(n * 31 + k)
    for hi in (4, 1 << 32):  # heavy ties and distinct
        v = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
        K, keys, idx = orc.topk_reference(v, k)
        K2, keys2, idx2 = np_topk(v, k)
        assert K == K2
        assert np.array_equal(idx, idx2) and np.array_equal(keys, keys2)


# ---------------------------------------------------------------- delegates
def np_sample(v, alpha, beta):
    """Brute-force top-beta per subrange: order by (value desc, index asc); pads are (0, 0xFFFFFFFF)."""
    L = 1 << alpha
    S = (v.size + L - 1) // L
    A = np.full(S * L, -1, np.int64)
    A[: v.size] = v
    A = A.reshape(S, L)
    I = np.arange(S * L, dtype=np.int64).reshape(S, L)
    order = np.lexsort((I, -A), axis=-1)[:, :beta]
    vals = np.take_along_axis(A, order, axis=1)
    pos = np.take_along_axis(I, order, axis=1)
    pos = np.where(vals < 0, 0xFFFFFFFF, pos)
    vals = np.where(vals < 0, 0, vals)
    return vals.reshape(-1).astype(np.uint32), pos.reshape(-1).astype(np.uint32)


@pytest.mark.parametrize("alpha,beta", [(1, 1), (2, 2), (3, 3), (4, 4), (5, 2), (6, 1), (7, 1), (9, 1)])
def test_sample_vs_bruteforce(orc, alpha, beta):
    rng = np.random.default_rng(alpha * 10 + beta)
    for n in (1 << 12, (1 << 12) + 3):
        for hi in (5, 1 << 32):
            v = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
            vals, pos = orc.sample(v, alpha, beta)
            v2, p2 = np_sample(v, alpha, beta)
            assert np.array_equal(vals, v2) and np.array_equal(pos, p2)


# ---------------------------------------------------------------- building blocks
@pytest.mark.parametrize("nb", [1, 2, 4, 8, 16])
def test_radix_kth(orc, nb):
    rng = np.random.default_rng(nb)
    for hi in (3, 1 << 32):
        v = rng.integers(0, hi, 3000, dtype=np.uint64).astype(np.uint32)
        for k in (1, 2, 100, 2999, 3000):
            want = int(np.sort(v)[v.size - k])
            assert orc.radix_kth(v, k, nb, plain=False) == want
            assert orc.radix_kth(v, k, nb, plain=True) == want


def test_radix_plain_phase1(orc):
    # n > 2^22 exercises Phase I compaction (radixselect.cuh:1497-1535)
    v = orc.generate(orc.UNIFORM_U32, 21, (1 << 22) + 1000)
    s = np.sort(v)
    for k in (1, 1 << 10, 1 << 21):
        assert orc.radix_kth(v, k, 8, plain=True) == int(s[v.size - k])


@pytest.mark.parametrize("B", [2, 3, 16, 256, 4096])
def test_bucket_map_is_exact(orc, B):
    rng = np.random.default_rng(B)
    for lo, hi in [(0, 0xFFFFFFFF), (5, 9), (100, 100 + B - 2), (7, 7 + 3 * B + 1), (0, 1)]:
        if hi <= lo:  # the map is defined for lo < hi; lo == hi ends the refinement first
            continue
        for v in np.concatenate([rng.integers(lo, hi + 1, 50), [lo, hi]]):
            v = int(v)
            b = orc.bucket_of(v, lo, hi, B)
            assert b == ((v - lo) * (B - 1)) // (hi - lo)  # exact integer value of the double map
            blo, bhi = orc.bucket_bounds(lo, hi, B, b)
            assert blo <= v <= bhi
        # bucket ranges tile [lo, hi] in order, last bucket is {hi}
        prev = lo - 1
        for b in range(B):
            blo, bhi = orc.bucket_bounds(lo, hi, B, b)
            if blo <= bhi:
                assert blo == prev + 1
                prev = bhi
        assert prev == hi and orc.bucket_bounds(lo, hi, B, B - 1) == (hi, hi)


@pytest.mark.parametrize("B", [4, 16, 256])
def test_bucket_kth(orc, B):
    rng = np.random.default_rng(B + 1)
    for hi in (2, 1000, 1 << 32):
        v = rng.integers(0, hi, 2000, dtype=np.uint64).astype(np.uint32)
        for k in (1, 7, 1000, 2000):
            assert orc.bucket_kth(v, k, B) == int(np.sort(v)[v.size - k])


@pytest.mark.parametrize("n,k", [(8, 1), (8, 8), (64, 4), (1000, 16), (4096, 256), (3000, 1024)])
def test_bitonic_network_is_partial_sort(orc, n, k):
    rng = np.random.default_rng(n + k)
    a = rng.integers(1, 1 << 63, n, dtype=np.uint64)
    out = orc.bitonic_topk64(a, k)
    assert np.array_equal(out, np.sort(a)[-k:])


# ---------------------------------------------------------------- full restatement
def cases():
    rng = np.random.default_rng(1234)
    yield "uniform", rng.integers(0, 1 << 32, 1 << 14, dtype=np.uint64).astype(np.uint32)
    yield "ties4", rng.integers(0, 4, 1 << 14, dtype=np.uint64).astype(np.uint32)
    yield "allequal", np.full(5000, 77, np.uint32)
    yield "sorted", np.arange(1 << 13, dtype=np.uint32)
    yield "reverse", np.arange(1 << 13, dtype=np.uint32)[::-1].copy()
    yield "ragged", rng.integers(0, 1 << 32, 10007, dtype=np.uint64).astype(np.uint32)
    yield "lognormal", None  # filled from the generator below
    yield "smallrange", rng.integers(1000, 1040, 9000, dtype=np.uint64).astype(np.uint32)


def load_case(orc, name, v):
    if name == "lognormal":
        return orc.generate(orc.LOGNORMAL_U32, 4, 1 << 14)
    return v


def param_grid():
    F, B, S = 1, 2, 4
    for algo in ALGOS:
        for flags in ([F | B | S, F | S, F | B, B | S, 0] if algo != 2 else [F | B | S]):
            for nb in ([8, 4, 16] if algo == 0 else [8]):
                for beta in ([2, 3, 4] if (flags & B) and algo != 2 else [2]):
                    yield algo, flags, nb, beta


@pytest.mark.parametrize("algo,flags,nb,beta", list(param_grid()))
def test_drtopk_matches_closed_form(orc, algo, flags, nb, beta):
    rng = np.random.default_rng(algo * 1000 + flags * 100 + nb * 10 + beta)
    loaded = [(name, load_case(orc, name, v)) for name, v in cases()]
    tried = 0
    for name, v in loaded:
        n = v.size
        ks = [1, 2, 16, 128, 1024, n // 8]
        if algo != 2:
            ks += [5, 1000, n // 3, n - 1, n]
        for k in rng.choice(ks, size=4, replace=False):
            k = int(k)
            for alpha in rng.choice([-1, 0, 2, 4, 5, 6, 8], size=3, replace=False):
                alpha = int(alpha)
                if k < 1 or k > n or (algo == 2 and k & (k - 1)):
                    continue
                p = orc.params(alpha=alpha, beta=beta, digit_bits=nb, flags=flags)
                if orc.resolve(algo, n, k, p) is None:
                    continue
                tried += 1
                K, keys, idx, st = orc.drtopk(algo, v, k, p)
                K2, keys2, idx2 = np_topk(v, k)
                assert K == K2, (name, algo, flags, nb, beta, k, alpha)
                assert np.array_equal(idx, idx2), (name, algo, flags, k, alpha)
                assert np.array_equal(keys, keys2)
                check_invariants(orc, algo, v, k, st, flags)
    assert tried > 20


def check_invariants(orc, algo, v, k, st, flags):
    K = st["kth_bits"]
    assert st["count_greater"] == int(np.sum(v > K)) and st["count_greater"] < k
    assert st["ties_taken"] == k - st["count_greater"] >= 1
    if st["path"] == 0:
        assert st["concat_size"] == v.size
        return
    alpha, beta = st["alpha"], st["beta"]
    L = 1 << alpha
    S = (v.size + L - 1) // L
    assert st["num_subranges"] == S and st["num_delegates"] == beta * S
    thr = st["threshold"]
    assert thr <= K  # the first top-k never overshoots the k-th largest key
    vals, pos = np_sample(v, alpha, beta)
    d_sorted = np.sort(vals.reshape(S, beta)[:, 0] if beta == 1 else vals)
    real = vals if beta == 1 else vals  # padded slots are 0 and never exceed the k-th delegate
    dk = int(np.sort(real)[real.size - k])
    if st["path"] == 1:
        assert st["selected_subranges"] == k
        if algo == 1 or algo == 2:
            assert thr == dk  # bucket and bitonic: exact k-th delegate
        else:
            # radix: prefix of the k-th delegate at the first digit where it is alone in its bucket
            nb = st["digit_bits"]
            want, passes = None, 0
            for d in range(32 // nb - 1, -1, -1):
                passes += 1
                sh = d * nb
                pf = (dk >> sh) << sh
                if int(np.sum((d_sorted.astype(np.uint64) >> sh) == (pf >> sh))) == 1:
                    want = pf
                    break
            if want is None:
                want = dk
            assert thr == want and st["first_passes"] == passes
        if flags & 1 and algo != 2:
            # filtered concatenation = keys >= threshold inside the selected subranges
            assert st["concat_size"] <= k * L
    else:
        nb = st["digit_bits"]
        if algo == 0:
            sh = 3 * nb  # digits 32/nb-1 .. 3 are resolved, the lower three digits stay 0
            want = (dk >> sh) << sh if 32 // nb - 1 >= 3 else 0
            # with full delegates (not uint32 keys with 8-bit digits, beta 2 and FILTER), a dense
            # threshold (more than 4k delegates at or above it) takes one digit more
            compact = nb == 8 and beta == 2 and (flags & 1)
            if not compact and 32 // nb - 1 >= 3 and int(np.sum(real.astype(np.uint64) >= want)) > 4 * k:
                want = (dk >> (2 * nb)) << (2 * nb)
            assert thr == want
        # every key >= threshold reaches the concatenation (filter) or its whole subrange does
        n_ge = int(np.sum(v >= thr))
        assert st["concat_size"] >= n_ge
        if flags & 1:
            assert st["concat_size"] == n_ge + int(np.sum(vals_lonely_below(v, vals, alpha, beta, thr)))


def vals_lonely_below(v, vals, alpha, beta, thr):
    """Delegates below the threshold that the lonely rule still appends (radixselect.cuh:496-507)."""
    L = 1 << alpha
    S = (v.size + L - 1) // L
    d = vals.reshape(S, beta).astype(np.int64)
    m = np.minimum(beta, v.size - np.arange(S) * L)
    real = np.arange(beta)[None, :] < m[:, None]
    below = real & (d < thr)
    lonely = (d[:, 0] >= thr) & below.any(axis=1)
    return below[lonely].sum(axis=1)


def test_invalid_parameters(orc):
    v = np.arange(1024, dtype=np.uint32)
    with pytest.raises(ValueError):
        orc.drtopk(0, v, 0, orc.params())
    with pytest.raises(ValueError):
        orc.drtopk(0, v, 2000, orc.params())
    with pytest.raises(ValueError):
        orc.drtopk(0, v, 4, orc.params(digit_bits=3))
    with pytest.raises(ValueError):
        orc.drtopk(2, v, 3, orc.params())  # bitonic needs a power-of-two k
    with pytest.raises(ValueError):
        orc.drtopk(0, v, 512, orc.params(alpha=4, flags=1))  # 64 subranges < k on the default path


def test_alpha_policy_table(orc):
    # Observed alpha at |V|=2^30 for k=2^0..2^24 (baseline+filter+beta+shuffle/optimum_FICO:768-2532)
    want = [16, 16, 15, 15, 14, 14, 13, 13, 12, 12, 11, 11, 10, 10, 9, 9, 8, 8, 7, 7, 6, 6, 5, 5, 4]
    got = [orc.resolve(0, 1 << 30, 1 << j, orc.params())["alpha"] for j in range(25)]
    assert got == want
    paths = [orc.resolve(0, 1 << 30, 1 << j, orc.params())["path"] for j in range(25)]
    assert paths == [1] * 22 + [2] * 3
    # bucket: alpha++ when alpha < 5 (bucket_select/main.cu:66-67)
    assert orc.resolve(1, 1 << 30, 1 << 24, orc.params())["alpha"] == 5


def test_cpu_baseline_matches(orc):
    v = orc.generate(orc.LOGNORMAL_U32, 4, 1 << 16)
    for k in (1, 100, 5000):
        t, K, keys, idx = orc.cpu_baseline(v, k, threads=4)
        K2, keys2, idx2 = np_topk(v, k)
        assert t >= 0 and K == K2 and np.array_equal(idx, idx2) and np.array_equal(keys, keys2)


# ---------------------------------------------------------------- exact checker
@pytest.mark.parametrize("n,k", [(1, 1), (10, 3), (1000, 1), (1000, 999), (4099, 64), (65537, 4096)])
@pytest.mark.parametrize("kind", ["u32", "ties", "f32"])
def test_verify_topk_accepts_numpy_result(orc, n, k, kind):
    """The checker accepts the numpy brute-force top-k (on every thread count) and rejects each
    planted mistake: a k-th key one off, two indices swapped, a later tie instead of the first
    untaken one, a changed key."""
    rng = np.random.default_rng(n + k)
    if kind == "u32":
        raw = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        ordk = raw
    elif kind == "ties":
        raw = rng.integers(0, 5, n, dtype=np.uint64).astype(np.uint32)
        ordk = raw
    else:
        raw = np.array([-2.0, -0.0, 0.0, 1.5, -1e-20, 3.0, 1.5, np.inf], np.float32)[
            rng.integers(0, 8, n)].view(np.uint32)
        ordk = orc.f32_to_ordered(raw)
    K, keys_o, idx = np_topk(ordk, k)
    kth_raw = int(raw[idx[np.nonzero(keys_o == K)[0][0]]]) if np.any(keys_o == K) else None
    keys_raw = raw[idx]
    f32 = kind == "f32"
    for threads in (1, 3, 8):
        assert orc.verify_topk(raw, k, kth_raw, keys_raw, idx, f32=f32, threads=threads) == 0
    assert orc.verify_topk(raw, k, kth_raw, None, None, f32=f32) == 0
    # planted mistakes (in the checker's input, on the CPU)
    # K + 1 (ordered) has #(v >= K + 1) = #(v > K) < k keys at or above it: never the k-th largest
    if K < 0xFFFFFFFF:
        bad_k = orc.ordered_to_f32(np.array([K + 1], np.uint32))[0] if f32 else K + 1
        assert orc.verify_topk(raw, k, int(bad_k), None, None, f32=f32) == 2
    # K - 1 has #(v > K - 1) = #(v >= K) >= k keys above it: never the k-th largest either
    if K > 0:
        bad_k = orc.ordered_to_f32(np.array([K - 1], np.uint32))[0] if f32 else K - 1
        assert orc.verify_topk(raw, k, int(bad_k), None, None, f32=f32) == 2
    if k >= 2:
        j = idx.copy()
        j[0], j[1] = j[1], j[0]
        assert orc.verify_topk(raw, k, kth_raw, keys_raw, j, f32=f32) == 3
        kk = keys_raw.copy()
        kk[-1] ^= 1
        assert orc.verify_topk(raw, k, kth_raw, kk, idx, f32=f32) == 4
    ties_all = np.nonzero(ordk == K)[0]
    taken = np.intersect1d(ties_all, idx)
    if taken.size and taken.size < ties_all.size:
        j = idx.copy()
        pos = np.nonzero(j == taken[-1])[0][0]
        later = ties_all[ties_all > taken[-1]][0]
        j[pos] = later
        j.sort()
        assert orc.verify_topk(raw, k, kth_raw, None, j, f32=f32) == 3


def test_verify_rows(orc):
    """The row-parallel checker: every row of a packed buffer exact; one wrong row is named."""
    rng = np.random.default_rng(3)
    rows, parts, off, oo = [], [], 0, 0
    kth, ok, oi = [], [], []
    for r in range(40):
        n = int(rng.integers(1, 5000))
        k = int(rng.integers(1, n + 1))
        d = r % 2
        raw = (rng.standard_normal(n).astype(np.float32).view(np.uint32) if d else
               rng.integers(0, 50, n, dtype=np.uint64).astype(np.uint32))
        o = orc.f32_to_ordered(raw) if d else raw
        K, _, idx = np_topk(o, k)
        kth.append(int(raw[idx[np.nonzero(o[idx] == K)[0][0]]]))
        ok.append(raw[idx])
        oi.append(idx)
        pad = (-n) % 4
        parts.append(np.concatenate([raw, np.zeros(pad, np.uint32)]))
        rows.append((off, n, k, oo, d))
        off += n + pad
        oo += k
    base = np.concatenate(parts)
    okeys, oidx = np.concatenate(ok), np.concatenate(oi)
    kth = np.array(kth, np.uint32)
    assert not np.any(orc.verify_rows(base, rows, kth, okeys, oidx))
    bad = oidx.copy()
    r7 = rows[7]
    if r7[2] >= 2:
        bad[r7[3]], bad[r7[3] + 1] = bad[r7[3] + 1], bad[r7[3]]
        codes = orc.verify_rows(base, rows, kth, okeys, bad)
        assert list(np.nonzero(codes)[0]) == [7]


# ------------------------------------------------------------------ the reference's adversarial vector
def test_customized_vector_closed_form(orc):
    """customizedDatasetGenerator.cpp's vector (tests/golden/customized.py): 1024 keys climbing
    through four ever finer levels, then copies of 2^32 - 1.  The oracle's k-th key, top-k keys
    and indices equal the sort-based closed form on both sides of the boundary, and the plain
    bucket select refines through one level per level of the vector (four) just past the tail."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import customized as cz
    n = 1 << 16
    v = cz.customized(n)
    assert v.dtype == np.uint32 and (v[1024:] == 0xFFFFFFFF).all() and v[0] == 8388607
    T = n - 1024
    s = np.sort(v)
    for k in [1, T - 1, T, T + 1, T + 10, T + 256, T + 257, T + 700, n]:
        K = int(s[n - k])
        gt = np.nonzero(v > K)[0]
        eq = np.nonzero(v == K)[0][:k - gt.size]
        want_idx = np.sort(np.concatenate([gt, eq]))
        for algo in (0, 1):
            for alpha in (-1, 0):
                kk, keys, idx, st = orc.drtopk(algo, v, k, orc.params(alpha=alpha))
                assert kk == K, (k, algo, alpha)
                assert np.array_equal(idx, want_idx) and np.array_equal(keys, v[want_idx]), (k, algo, alpha)
    _, _, _, st = orc.drtopk(1, v, T + 1, orc.params(alpha=0))
    assert st["second_passes"] == 4
    _, _, _, st = orc.drtopk(1, v, T, orc.params(alpha=0))
    assert st["second_passes"] == 1


def test_dense_beta_threshold_takes_one_more_digit(orc):
    """Float keys U(-1,1): the top byte of their ordered bits is the sign and seven exponent bits,
    so the reference's one-digit beta-path threshold leaves far more than 4k delegates above it.
    With full delegates (float keys) the threshold takes the second digit of the k-th largest
    delegate; the compact uint32 form keeps the reference's one digit.  Outputs are identical."""
    n, k = 1 << 20, 1 << 13
    v = orc.f32_to_ordered(orc.generate(4, 77, n))
    p = orc.params()
    K2, keys2, idx2 = np_topk(v, k)
    vals, _ = np_sample(v, 5, 2)  # alpha 5 = (20 + 3 - 13) // 2, beta 2
    dk = int(np.sort(vals)[vals.size - k])
    assert int(np.sum(vals >= ((dk >> 24) << 24))) > 4 * k  # the one-byte threshold is dense here
    for f32, want, passes in ((True, (dk >> 16) << 16, 2), (False, (dk >> 24) << 24, 1)):
        K, keys, idx, st = orc.drtopk(0, v, k, p, f32=f32)
        assert st["alpha"] == 5 and st["path"] == 2
        assert st["threshold"] == want and st["first_passes"] == passes, (f32, hex(st["threshold"]))
        assert K == K2 and np.array_equal(idx, idx2) and np.array_equal(keys, keys2)
