// drtopk_oracle.cpp -- CPU ORACLE FOR TESTS ONLY.
//
// This file is test infrastructure.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker (or as the
// reported CPU baseline); the library never calls it and has no CPU fallback.
//
// Contents
//   1. Seeded synthetic generators (std::mt19937 + Box-Muller, project-owned transform).
//   2. The closed-form oracle: std::nth_element for the k-th largest, plus the
//      top-k set under the lowest-index-wins tie rule.
//   3. A CPU restatement of Dr.Top-k as the reference computes it -- delegate
//      sampling, first top-k over delegates, range selection, concatenation,
//      second top-k -- for the radix, bucket and bitonic families, each step
//      citing the reference file:line it restates.  Every function works on
//      "ordered" uint32 keys (float32 keys are mapped by sign flip first).
//   4. The OpenMP CPU baseline timed beside the GPU (nth_element per chunk,
//      nth_element over the candidates, sort of the top k).
//   5. An exact O(n) checker of a finished top-k (k-th key, keys, indices) for
//      full-size inputs and many rows, where nth_element per case is too slow.
//
// Deviations from the reference, all forced by exactness or determinism:
//   - Prefix masks are computed explicitly: a prefix of length 0 is 0.  The
//     reference writes (v>>32)<<32 (radixselect.cuh:367,:397,:1432), which is
//     0 on NVIDIA (shift clamps) but v on AMD (shift count mod 32).
//   - Tied delegates are filled lowest subrange id first (the reference's
//     FillWithKElementsID, radixselect.cuh:603-618, is a racy first-come fill).
//   - The bucket map is integer: bucket(v) = floor((v-lo)*(B-1)/(hi-lo)), the
//     exact value of the reference's double expression (bucketselect_combined.cuh:180).
//     The beta-path threshold is the lower bound of the level-1 k-bucket, which
//     is always <= the k-th delegate (the reference takes an arbitrary key of
//     that bucket, bucketselect_combined.cuh:1075-1079, which is not).
//   - The concatenation is built in subrange order (the reference appends with
//     global atomics, so its order is not reproducible); the multiset is the same.
//   - Parity status: every function here is pinned by tests/test_oracle.py
//     (closed form, brute force, invariants); none is left unpinned.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/drtopk.h"

typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

// ---------------------------------------------------------------------------
// 1. Generators
// ---------------------------------------------------------------------------
// dist: 0 uniform u32 (raw mt19937 output; equal to libstdc++'s
//       uniform_int_distribution<uint32_t>(0, 2^32-1), baseline+filter+beta+shuffle/main.cu:101)
//       1 normal(2^31, 2^28) clamped to [0, 2^32-1], rounded half up
//       2 lognormal(mu=20, sigma=2) clamped, rounded half up
//       3 float32 N(0,1)   (output: float bit patterns)
//       4 float32 U(-1,1)  (24-bit grid: (a>>8)*2^-23 - 1)
// Doubles in [0,1) use two draws (a>>5, b>>6) -> (a*2^26+b)*2^-53 (53-bit).
// Box-Muller: r = sqrt(-2 ln(1-u1)), z0 = r cos(2 pi u2), z1 = r sin(2 pi u2).

static inline double res53(std::mt19937& g) {
    u32 a = g() >> 5, b = g() >> 6;
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

static inline u32 clamp_round_u32(double x) {
    if (!(x > 0.0)) return 0u;              // also maps NaN to 0
    if (x >= 4294967295.0) return 0xFFFFFFFFu;
    double r = std::floor(x + 0.5);
    if (r >= 4294967295.0) return 0xFFFFFFFFu;
    return (u32)r;
}

extern "C" void orc_mt19937(u32 seed, u64 count, u32* out) {
    std::mt19937 g(seed);
    for (u64 i = 0; i < count; ++i) out[i] = g();
}

extern "C" int orc_generate(int dist, u32 seed, u64 n, u32* out) {
    std::mt19937 g(seed);
    const double two_pi = 6.283185307179586476925286766559;
    if (dist == 0) {
        for (u64 i = 0; i < n; ++i) out[i] = g();
        return 0;
    }
    if (dist == 4) {
        for (u64 i = 0; i < n; ++i) {
            float f = (float)((double)(g() >> 8) * (1.0 / 8388608.0) - 1.0);
            std::memcpy(&out[i], &f, 4);
        }
        return 0;
    }
    if (dist < 1 || dist > 3) return -1;
    // Draw the uniforms sequentially (the stream order fixes the data), transform in parallel.
    u64 pairs = (n + 1) / 2;
    std::vector<double> u(2 * pairs);
    for (u64 p = 0; p < pairs; ++p) { u[2 * p] = res53(g); u[2 * p + 1] = res53(g); }
#pragma omp parallel for schedule(static)
    for (i64 p = 0; p < (i64)pairs; ++p) {
        double r = std::sqrt(-2.0 * std::log(1.0 - u[2 * p]));
        double th = two_pi * u[2 * p + 1];
        double z[2] = {r * std::cos(th), r * std::sin(th)};
        for (int j = 0; j < 2; ++j) {
            u64 i = 2 * (u64)p + j;
            if (i >= n) break;
            if (dist == 1) out[i] = clamp_round_u32(2147483648.0 + 268435456.0 * z[j]);
            else if (dist == 2) out[i] = clamp_round_u32(std::exp(20.0 + 2.0 * z[j]));
            else { float f = (float)z[j]; std::memcpy(&out[i], &f, 4); }
        }
    }
    return 0;
}

// float32 bits <-> ordered uint32 (monotone: a<b as floats (total order) iff ord(a)<ord(b))
static inline u32 f2o(u32 b) { return b ^ ((b >> 31) ? 0xFFFFFFFFu : 0x80000000u); }
static inline u32 o2f(u32 o) { return o ^ ((o >> 31) ? 0x80000000u : 0xFFFFFFFFu); }
extern "C" void orc_f32_to_ordered(const u32* in, u64 n, u32* out) { for (u64 i = 0; i < n; ++i) out[i] = f2o(in[i]); }
extern "C" void orc_ordered_to_f32(const u32* in, u64 n, u32* out) { for (u64 i = 0; i < n; ++i) out[i] = o2f(in[i]); }

// ---------------------------------------------------------------------------
// 2. Closed-form oracle
// ---------------------------------------------------------------------------
// kth = ascending rank n-k (the reference's own check: sort, vec1[n-k]==TopKElement,
// baseline+filter+beta+shuffle/main.cu:281-296).  Top-k set = {i : v[i] > kth} plus the
// lowest r = k - #gt indices with v[i] == kth; written in ascending index order.
static void extract_from(const u32* keys, const u32* idx, u64 m, u64 k, u32 kth, u32* out_keys,
                         i64* out_idx, u64* cnt_gt_out) {
    u64 gt = 0;
    for (u64 i = 0; i < m; ++i) gt += keys[i] > kth;
    u64 r = k - gt, ties = 0, w = 0;
    for (u64 i = 0; i < m; ++i) {
        bool take = keys[i] > kth || (keys[i] == kth && ties++ < r);
        if (take) {
            if (out_keys) out_keys[w] = keys[i];
            if (out_idx) out_idx[w] = idx ? (i64)idx[i] : (i64)i;
            ++w;
        }
    }
    if (cnt_gt_out) *cnt_gt_out = gt;
}

extern "C" int orc_topk_reference(const u32* keys, u64 n, u64 k, u32* kth, u32* out_keys, i64* out_idx) {
    if (k == 0 || k > n) return -1;
    std::vector<u32> tmp(keys, keys + n);
    std::nth_element(tmp.begin(), tmp.begin() + (n - k), tmp.end());
    u32 K = tmp[n - k];
    if (kth) *kth = K;
    extract_from(keys, nullptr, n, k, K, out_keys, out_idx, nullptr);
    return 0;
}

// ---------------------------------------------------------------------------
// 3. Restatement of Dr.Top-k
// ---------------------------------------------------------------------------
static inline u32 prefix_of(u32 v, int shleft) { return shleft >= 32 ? 0u : (v >> shleft) << shleft; }

static int floor_log2(u64 x) { int l = -1; while (x) { x >>= 1; ++l; } return l; }
static int ceil_log2(u64 x) { int l = floor_log2(x); return ((1ull << l) == x) ? l : l + 1; }

struct Resolved {
    int algo, alpha, beta, path, nbits, nbuckets;
    unsigned flags;
    u64 S, D;  // subranges, delegates (beta*S)
};

// Number of real delegates (a short last subrange holds fewer than beta keys).
static u64 real_delegates(u64 n, int alpha, int beta) {
    u64 L = 1ull << alpha, S = (n + L - 1) / L, last = n - (S - 1) * L;
    return (u64)beta * (S - 1) + std::min<u64>((u64)beta, last);
}

// Parameter policy.  Radix: alpha = floor(0.5*(log2|V| - log2 k + 3)), beta path iff alpha<=5
// (baseline+filter+beta+shuffle/main.cu:83-89).  Bucket: same formula, alpha++ if alpha<5,
// beta path iff alpha<=5 (bucket_select/main.cu:66-67, :86-90).  Bitonic: alpha is a user input
// (bitonic/sampleBitonic.cu:203); auto uses the radix formula clamped to [5,16] (README.md:52:
// "about 16 to 5 ... for increasing k").  For k not a power of two log2 k is rounded up, which
// equals the floor of the reference's real-valued expression.
extern "C" int orc_resolve(int algo, u64 n, u64 k, const drtopk_params* p, int* out /*[6]*/) {
    if (!p || k == 0 || k > n || n == 0) return 1;
    if (p->digit_bits != 1 && p->digit_bits != 2 && p->digit_bits != 4 && p->digit_bits != 8 &&
        p->digit_bits != 16) return 1;
    if (p->num_buckets < 4 || p->num_buckets > 4096) return 1;  // B=2 narrows by one value per level
    if (p->beta < 1 || p->beta > 4) return 1;
    int Ln = floor_log2(n);
    int alpha = p->alpha;
    bool autoa = alpha < 0;
    if (autoa) {
        alpha = (Ln + 3 - ceil_log2(k)) / 2;
        if (alpha < 0) alpha = 0;
        if (algo == DRTOPK_ALGO_BUCKET && alpha < 5) alpha++;
        if (algo == DRTOPK_ALGO_BITONIC) alpha = std::min(16, std::max(5, alpha));
        if (alpha > Ln) alpha = Ln;
    }
    if (alpha > 30 || alpha > Ln) return 1;
    int beta = 1, path = alpha == 0 ? 0 : 1;
    auto want_beta = [&](int a) {
        return algo != DRTOPK_ALGO_BITONIC && (p->flags & DRTOPK_FLAG_BETA) && a <= 5 && p->beta >= 2;
    };
    auto valid = [&](int a) -> bool {
        if (a == 0) return true;
        u64 L = 1ull << a, S = (n + L - 1) / L;
        if (algo == DRTOPK_ALGO_BITONIC) return S > k;  // sampleBitonic.cu:286 (NSubranges>k)
        if (want_beta(a)) return p->beta <= (int)L && real_delegates(n, a, p->beta) >= k;
        return S >= k;
    };
    if (!valid(alpha)) {
        if (!autoa) return 1;
        while (alpha > 0 && !valid(alpha)) --alpha;
        path = alpha == 0 ? 0 : 1;
    }
    if (alpha > 0 && want_beta(alpha)) { beta = p->beta; path = 2; }
    if (algo == DRTOPK_ALGO_BITONIC && (k & (k - 1))) return 1;  // sampleBitonic.cu:198-202
    out[0] = alpha; out[1] = beta; out[2] = path;
    out[3] = p->digit_bits; out[4] = p->num_buckets; out[5] = (int)p->flags;
    return 0;
}

// --- delegates: sampleMax (radixselect.cuh:883-918), sampleMax_NoReduction_Beta2/3 (:967-1117) ---
// Top-beta multiset of each subrange, largest first; among equal values the lowest index first.
// Missing slots of a short subrange: value 0, position 0xFFFFFFFF.
extern "C" int orc_sample(const u32* keys, u64 n, int alpha, int beta, u32* vals, u32* pos) {
    if (alpha < 0 || alpha > 30 || beta < 1 || beta > 4 || n == 0) return 1;
    u64 L = 1ull << alpha, S = (n + L - 1) / L;
    for (u64 s = 0; s < S; ++s) {
        u32 bv[4] = {0, 0, 0, 0};
        u32 bp[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
        int have = 0;
        u64 e = std::min(n, (s + 1) * L);
        for (u64 i = s * L; i < e; ++i) {
            u32 v = keys[i];
            // insert (v, i) keeping order value desc, index asc; later index loses ties
            int j = have < beta ? have : beta;
            if (have == beta && !(v > bv[beta - 1])) continue;
            if (have < beta) ++have;
            j = have - 1;
            while (j > 0 && v > bv[j - 1]) { bv[j] = bv[j - 1]; bp[j] = bp[j - 1]; --j; }
            bv[j] = v; bp[j] = (u32)i;
        }
        for (int j = 0; j < beta; ++j) {
            vals[s * beta + j] = bv[j];
            if (pos) pos[s * beta + j] = bp[j];
        }
    }
    return 0;
}

// --- radix digit step: CumulateCount_inplace (radixselect.cuh:192-220) + host loop (:1427-1445) ---
struct RadixState { u32 flag; u64 below; int digit; u64 kcount; };

static void radix_digit_step(const u32* a, u64 m, u64 k, int nb, RadixState& st) {
    const u64 B = 1ull << nb;
    const int shleft = (st.digit + 1) * nb, shright = st.digit * nb;
    std::vector<u64> cnt(B, 0);
    for (u64 i = 0; i < m; ++i)
        if (prefix_of(a[i], shleft) == st.flag) cnt[(a[i] >> shright) & (B - 1)]++;
    u64 target = m - k + 1, sum = st.below, prev = st.below;
    u64 kd = 0;
    for (u64 i = 0; i < B; ++i) {
        sum += cnt[i];
        if (sum >= target && (i == 0 ? true : prev < target)) { kd = i; break; }
        prev = sum;
    }
    st.flag |= (u32)(kd << shright);
    st.below = prev;  // cumulative count below the k-digit (Belowcount itself when kd == 0)
    st.kcount = cnt[kd];
    st.digit--;
}

// radix_select_inplace (radixselect.cuh:1411-1470): every digit, k-th largest = flag.
static u32 radix_inplace(const u32* a, u64 m, u64 k, int nb, int* passes) {
    RadixState st{0, 0, 32 / nb - 1, 0};
    int p = 0;
    while (st.digit >= 0) { radix_digit_step(a, m, k, nb, st); ++p; }
    if (passes) *passes += p;
    return st.flag;
}

// radix_select (radixselect.cuh:1473-1543): Phase I compacts while m > 2^22, then Phase II.
static u32 radix_plain(const u32* a0, u64 m, u64 k, int nb, int* passes) {
    const u64 limit = 1ull << 22;
    if (m <= limit) return radix_inplace(a0, m, k, nb, passes);
    const u64 B = 1ull << nb;
    std::vector<u32> cur(a0, a0 + m), nxt;
    u32 flag = 0;
    int digit = 32 / nb - 1;
    while (m > limit && digit >= 0) {
        int shift = digit * nb;
        std::vector<u64> cnt(B, 0);
        for (u64 i = 0; i < m; ++i) cnt[(cur[i] >> shift) & (B - 1)]++;  // CalculateOccurence :622-649
        ++*passes;
        u64 target = m - k + 1, sum = 0, kd = 0, cumk = 0;
        for (u64 i = 0; i < B; ++i) { sum += cnt[i]; if (sum >= target) { kd = i; cumk = sum; break; } }
        flag |= (u32)(kd << shift);
        nxt.clear();
        for (u64 i = 0; i < m; ++i) if (prefix_of(cur[i], shift) == flag) nxt.push_back(cur[i]);  // :290-303
        k = k - (m - cumk);
        m = cnt[kd];
        cur.swap(nxt);
        if (m == 1) return cur[0];
        digit--;
    }
    RadixState st{flag, 0, digit, 0};
    while (st.digit >= 0) { radix_digit_step(cur.data(), m, k, nb, st); ++*passes; }
    return st.flag;
}

// --- bucket select, integer map (bucketselect_combined.cuh:167-216, :350-470) ---
static inline u64 bucket_of(u32 v, u32 lo, u32 hi, u64 B) { return ((u64)(v - lo) * (B - 1)) / (u64)(hi - lo); }
static inline u32 bucket_lo(u64 b, u32 lo, u32 hi, u64 B) {
    u64 W = hi - lo; return lo + (u32)((b * W + (B - 2)) / (B - 1));
}
static inline u32 bucket_hi(u64 b, u32 lo, u32 hi, u64 B) { return b + 1 == B ? hi : bucket_lo(b + 1, lo, hi, B) - 1; }

struct BucketLevel { u64 kb, below, kcount; };
// FindKBucket (bucketselect_combined.cuh:200-216): first bucket whose cumulative count exceeds a.
static BucketLevel bucket_level(const u32* x, u64 m, u32 lo, u32 hi, u64 a, u64 B) {
    std::vector<u64> cnt(B, 0);
    for (u64 i = 0; i < m; ++i) if (x[i] >= lo && x[i] <= hi) cnt[bucket_of(x[i], lo, hi, B)]++;
    u64 sum = 0;
    for (u64 b = 0; b < B; ++b) { if (sum + cnt[b] > a) return {b, sum, cnt[b]}; sum += cnt[b]; }
    return {B - 1, sum - cnt[B - 1], cnt[B - 1]};
}
// Exact element of ascending rank a among x[0..m): refine until the k-bucket holds one key or one value.
static u32 bucket_kth(const u32* x, u64 m, u64 a, u64 B, int* passes) {
    u32 lo = *std::min_element(x, x + m), hi = *std::max_element(x, x + m);
    while (lo != hi) {
        BucketLevel L = bucket_level(x, m, lo, hi, a, B);
        ++*passes;
        u32 nlo = bucket_lo(L.kb, lo, hi, B), nhi = bucket_hi(L.kb, lo, hi, B);
        a -= L.below;
        if (L.kcount == 1) {
            for (u64 i = 0; i < m; ++i) if (x[i] >= nlo && x[i] <= nhi) return x[i];  // GetKValue :271-298
        }
        lo = nlo; hi = nhi;
    }
    return lo;
}

// --- bitonic top-k network (bitonic/bitonic_com.cuh:101-228, :231-386, :388-510) ---
// Composite keys make every comparison strict: larger key wins, then lower index.
static void bitonic_cmpx(std::vector<u64>& A, u64 n, u64 len, u64 inc) {
    u64 dir = len << 1;
    for (u64 t = 0; t < n / 2; ++t) {
        u64 low = t & (inc - 1), i = (t << 1) - low;
        bool up = (dir & i) == 0;
        if (up == (A[i] > A[i + inc])) std::swap(A[i], A[i + inc]);
    }
}
// Local sort into alternating k-runs (:271-295), then merge = elementwise max of run pairs
// (:299-321) and rebuild (:332-356) until k elements remain; result sorted ascending.
static void bitonic_topk(std::vector<u64>& A, u64 k) {
    u64 n = 1; while (n < A.size()) n <<= 1;
    if (n < k) n = k;
    A.resize(n, 0ull);
    for (u64 len = 1; len < k; len <<= 1)
        for (u64 inc = len; inc > 0; inc >>= 1) bitonic_cmpx(A, n, len, inc);
    while (n > k) {
        std::vector<u64> B(n / 2);
        for (u64 t = 0; t < n / 2; ++t) { u64 low = t & (k - 1), i = (t << 1) - low; B[t] = std::max(A[i], A[i + k]); }
        A.swap(B); n >>= 1;
        for (u64 inc = k >> 1; inc > 0; inc >>= 1) bitonic_cmpx(A, n, k >> 1, inc);
    }
    A.resize(k);
}

// --- the pipeline ---
// sample_radix_select (radixselect.cuh:1783-1894), sample_bucket_select
// (bucketselect_combined.cuh:1280-1358), bitonic driver (sampleBitonic.cu:283-308).
// f32: the ordered keys come from float32 keys (only the storage of the beta path's delegates,
// and so the dense-threshold rule below, depends on it).
extern "C" int orc_drtopk(int algo, const u32* keys, u64 n, u64 k, const drtopk_params* p, int f32, u32* kth_out,
                          u32* out_keys, i64* out_idx, drtopk_stats* st) {
    int r[6];
    if (orc_resolve(algo, n, k, p, r)) return 1;
    const int alpha = r[0], beta = r[1], path = r[2], nb = r[3];
    const u64 B = (u64)r[4];
    const unsigned flags = (unsigned)r[5];
    drtopk_stats s;
    std::memset(&s, 0, sizeof(s));
    s.algo = algo; s.alpha = alpha; s.beta = beta; s.path = path; s.digit_bits = nb; s.num_buckets = (int)B;

    std::vector<u32> ck, ci;  // concatenation: keys and their indices, in subrange order
    u32 K = 0;
    if (path == 0) {
        // No delegates: radix_select / bucket_select / plain bitonic over all keys.
        s.num_subranges = 0; s.num_delegates = 0;
        ck.assign(keys, keys + n);
        ci.resize(n);
        for (u64 i = 0; i < n; ++i) ci[i] = (u32)i;
        s.concat_size = n;
        int passes = 0;
        if (algo == DRTOPK_ALGO_RADIX) K = radix_plain(keys, n, k, nb, &passes);
        else if (algo == DRTOPK_ALGO_BUCKET) K = bucket_kth(keys, n, n - k, B, &passes);
        else {
            std::vector<u64> A(n);
            for (u64 i = 0; i < n; ++i) A[i] = ((u64)keys[i] << 32) | (0xFFFFFFFFu - (u32)i);
            bitonic_topk(A, k);
            K = (u32)(A[0] >> 32);
        }
        s.second_passes = passes;
    } else {
        const u64 L = 1ull << alpha, S = (n + L - 1) / L, D = (u64)beta * S;
        s.num_subranges = S; s.num_delegates = D;
        std::vector<u32> dv(D), dp(D);
        orc_sample(keys, n, alpha, beta, dv.data(), dp.data());
        const u64 Dreal = real_delegates(n, alpha, beta);
        auto sub_len = [&](u64 sidx) { return std::min(n, (sidx + 1) * L) - sidx * L; };
        auto gather = [&](u64 sidx, u32 thr, bool filt) {
            for (u64 i = sidx * L, e = sidx * L + sub_len(sidx); i < e; ++i)
                if (!filt || keys[i] >= thr) { ck.push_back(keys[i]); ci.push_back((u32)i); }
        };
        const bool filt = (flags & DRTOPK_FLAG_FILTER) != 0;
        u32 thr = 0;
        int fpasses = 0;
        if (path == 1) {
            // ---- default path: one max per subrange ----
            int shleft = 0;  // prefix length used by range selection
            if (algo == DRTOPK_ALGO_RADIX) {
                // radix_select_inplace_firsttopk (radixselect.cuh:1667-1724): digits MSD->LSD, stop once
                // the k-digit bucket holds a single delegate.
                RadixState rs{0, 0, 32 / nb - 1, 0};
                while (rs.digit >= 0) {
                    radix_digit_step(dv.data(), S, k, nb, rs); ++fpasses;
                    if (rs.kcount == 1) break;
                }
                thr = rs.flag;
                shleft = (rs.digit + 1) * nb;
            } else if (algo == DRTOPK_ALGO_BUCKET) {
                // bucket_select_inplaceFirstTopk (bucketselect_combined.cuh:1132-1276): exact k-th delegate.
                thr = bucket_kth(dv.data(), S, S - k, B, &fpasses);
                shleft = 0;
            } else {
                // bitonic_firstTopk (bitonic_com.cuh:610-666): network over (max, subrange id).
                std::vector<u64> A(S);
                for (u64 i = 0; i < S; ++i) A[i] = ((u64)dv[i] << 32) | (0xFFFFFFFFu - (u32)i);
                bitonic_topk(A, k);
                std::vector<u64> ids(k);
                for (u64 j = 0; j < k; ++j) ids[j] = 0xFFFFFFFFu - (u32)(A[j] & 0xFFFFFFFFu);
                std::sort(ids.begin(), ids.end());
                thr = (u32)(A[0] >> 32);
                for (u64 j = 0; j < k; ++j) gather(ids[j], 0, false);  // CreateConcatenateRange :72-98 (unfiltered)
                s.selected_subranges = k;
                s.threshold = thr;
            }
            if (algo != DRTOPK_ALGO_BITONIC) {
                // FindTopKRanges (radixselect.cuh:413-427) + FillWithKElementsID (:603-618): every delegate
                // whose prefix exceeds the flag, then lowest-id delegates with prefix == flag up to k.
                u64 gt = 0;
                for (u64 i = 0; i < S; ++i) gt += prefix_of(dv[i], shleft) > thr;
                u64 fill = k - gt, taken = 0;
                for (u64 i = 0; i < S; ++i) {
                    u32 pf = prefix_of(dv[i], shleft);
                    bool sel = pf > thr || (pf == thr && taken++ < fill);
                    if (sel) { gather(i, thr, filt); s.selected_subranges++; }  // CreateConcatenateRange_optimized :1575-1607
                }
                s.threshold = thr;
            }
        } else {
            // ---- beta path: top-beta delegates per subrange ----
            if (algo == DRTOPK_ALGO_RADIX) {
                // radix_select_inplace_firsttopk_beta (radixselect.cuh:1728-1780): digits while CurrentDigit>=3.
                RadixState rs{0, 0, 32 / nb - 1, 0};
                while (rs.digit >= 3) { radix_digit_step(dv.data(), D, k, nb, rs); ++fpasses; }
                // Build deviation: one digit more when the threshold is dense, i.e. more than 4k
                // delegates lie at or above it (float keys, whose top byte is the sign and seven
                // exponent bits; skewed keys).  Any prefix of the k-th largest delegate is a valid
                // threshold, so the outputs are unchanged; the concatenation shrinks.  Not with the
                // build's compact delegates (uint32 keys, 8-bit digits, beta 2, FILTER), which
                // keep only the second delegate's top byte and so the reference's one digit.
                const bool compact = !f32 && nb == 8 && beta == 2 && (flags & DRTOPK_FLAG_FILTER);
                if (!compact && fpasses > 0 && rs.digit == 2 && D - rs.below > 4 * k) {
                    radix_digit_step(dv.data(), D, k, nb, rs);
                    ++fpasses;
                }
                thr = rs.flag;
            } else {
                // bucket_select_inplace_betaFirstTopk (bucketselect_combined.cuh:973-1095): min/max, level 0,
                // one refinement level; threshold = lower bound of the level-1 k-bucket.
                u32 lo = *std::min_element(dv.begin(), dv.end()), hi = *std::max_element(dv.begin(), dv.end());
                u64 a = D - k;
                if (lo == hi) thr = hi;
                else {
                    BucketLevel L0 = bucket_level(dv.data(), D, lo, hi, a, B); ++fpasses;
                    u32 nlo = bucket_lo(L0.kb, lo, hi, B), nhi = bucket_hi(L0.kb, lo, hi, B);
                    a -= L0.below;
                    if (nlo == nhi) thr = nlo;
                    else {
                        BucketLevel L1 = bucket_level(dv.data(), D, nlo, nhi, a, B); ++fpasses;
                        thr = bucket_lo(L1.kb, nlo, nhi, B);
                    }
                }
            }
            s.threshold = thr;
            // FindTopKRanges_beta (radixselect.cuh:472-600): all beta delegates >= flag -> whole subrange,
            // else if the maximum >= flag -> its delegates go to the concatenation ("lonely").
            for (u64 i = 0; i < S; ++i) {
                u64 m = std::min<u64>((u64)beta, sub_len(i));
                if (dv[i * beta] < thr) continue;
                bool all = true;
                for (u64 j = 0; j < m; ++j) all = all && dv[i * beta + j] >= thr;
                if (all) { gather(i, thr, filt); s.selected_subranges++; }  // CreateConcatenateRange_Beta_optimized :1610-1639
                else {
                    s.lonely_subranges++; s.lonely_elements += m;
                    // lonely delegates, placed by position inside the subrange
                    u32 pv[4], pp[4];
                    for (u64 j = 0; j < m; ++j) { pv[j] = dv[i * beta + j]; pp[j] = dp[i * beta + j]; }
                    for (u64 x = 0; x < m; ++x)
                        for (u64 y = x + 1; y < m; ++y)
                            if (pp[y] < pp[x]) { std::swap(pp[x], pp[y]); std::swap(pv[x], pv[y]); }
                    for (u64 j = 0; j < m; ++j) { ck.push_back(pv[j]); ci.push_back(pp[j]); }
                }
            }
        }
        (void)Dreal;
        s.first_passes = fpasses;
        s.concat_size = ck.size();
        // ---- second top-k over the concatenation ----
        int passes = 0;
        const u64 C = ck.size();
        // cannot happen: kth(V) >= kth(delegates) >= flag, and every key >= flag is in the
        // concatenation (a selected subrange or a lonely delegate); reported as an internal error
        if (C < k) return 2;
        if (algo == DRTOPK_ALGO_RADIX) K = radix_inplace(ck.data(), C, k, nb, &passes);   // :1891
        else if (algo == DRTOPK_ALGO_BUCKET) K = bucket_kth(ck.data(), C, C - k, B, &passes);  // :1355
        else {
            std::vector<u64> A(C);
            for (u64 i = 0; i < C; ++i) A[i] = ((u64)ck[i] << 32) | (0xFFFFFFFFu - ci[i]);
            bitonic_topk(A, k);  // bitonic (bitonic_com.cuh:669-728)
            K = (u32)(A[0] >> 32);
        }
        s.second_passes = passes;
    }
    // ---- top-k set: findTopk + FillTopk (MultiGPU/radixselect.cuh:1355-1383), lowest index first ----
    u64 gt = 0;
    extract_from(ck.data(), ci.data(), ck.size(), k, K, out_keys, out_idx, &gt);
    s.kth_bits = K;
    s.count_greater = gt;
    s.ties_taken = k - gt;
    if (kth_out) *kth_out = K;
    if (st) *st = s;
    return 0;
}

// Bitonic network alone on composite keys, for pinning it against a sort (tests only).
extern "C" int orc_bitonic_topk64(const u64* in, u64 n, u64 k, u64* out) {
    if (k == 0 || (k & (k - 1)) || k > n) return 1;
    std::vector<u64> A(in, in + n);
    bitonic_topk(A, k);
    std::copy(A.begin(), A.end(), out);
    return 0;
}

// Radix and bucket k-th alone (tests pin them against nth_element).
extern "C" u32 orc_radix_kth(const u32* a, u64 m, u64 k, int nb, int plain) {
    int p = 0;
    return plain ? radix_plain(a, m, k, nb, &p) : radix_inplace(a, m, k, nb, &p);
}
extern "C" u32 orc_bucket_kth(const u32* a, u64 m, u64 k, int nbuckets) {
    int p = 0;
    return bucket_kth(a, m, m - k, (u64)nbuckets, &p);
}
extern "C" void orc_bucket_bounds(u32 lo, u32 hi, int nbuckets, u64 b, u32* blo, u32* bhi) {
    *blo = bucket_lo(b, lo, hi, (u64)nbuckets); *bhi = bucket_hi(b, lo, hi, (u64)nbuckets);
}
extern "C" u64 orc_bucket_of(u32 v, u32 lo, u32 hi, int nbuckets) { return bucket_of(v, lo, hi, (u64)nbuckets); }

// ---------------------------------------------------------------------------
// 4. CPU baseline: nth_element per chunk on a private copy (copy not timed), nth_element over the
//    P*k candidates, sort of the top k (descending), then the index pass (lowest index wins).
//    Returns seconds.
// ---------------------------------------------------------------------------
extern "C" double orc_cpu_baseline(const u32* keys, u64 n, u64 k, int threads, u32* kth_out,
                                   u32* out_keys, i64* out_idx) {
    if (k == 0 || k > n) return -1.0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
    int P = threads > 0 ? threads : omp_get_max_threads();
#else
    int P = 1;
#endif
    std::vector<u32> work(keys, keys + n);
    auto t0 = std::chrono::steady_clock::now();
    u64 chunk = (n + P - 1) / P;
    std::vector<u32> cand;
    std::vector<u64> cstart(P + 1, 0);
    for (int c = 0; c < P; ++c) {
        u64 b = std::min(n, (u64)c * chunk), e = std::min(n, b + chunk);
        cstart[c + 1] = cstart[c] + std::min<u64>(k, e - b);
    }
    cand.resize(cstart[P]);
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < P; ++c) {
        u64 b = std::min(n, (u64)c * chunk), e = std::min(n, b + chunk), len = e - b;
        u64 kk = std::min<u64>(k, len);
        if (kk == 0) continue;
        std::nth_element(work.begin() + b, work.begin() + (e - kk), work.begin() + e);
        std::copy(work.begin() + (e - kk), work.begin() + e, cand.begin() + cstart[c]);
    }
    u64 C = cand.size();
    std::nth_element(cand.begin(), cand.begin() + (C - k), cand.end());
    u32 K = cand[C - k];
    std::sort(cand.begin() + (C - k), cand.end(), std::greater<u32>());
    // index pass: per-chunk counts, then ordered writes
    std::vector<u64> gtc(P, 0), eqc(P, 0);
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < P; ++c) {
        u64 b = std::min(n, (u64)c * chunk), e = std::min(n, b + chunk), g = 0, q = 0;
        for (u64 i = b; i < e; ++i) { g += keys[i] > K; q += keys[i] == K; }
        gtc[c] = g; eqc[c] = q;
    }
    u64 gt = 0;
    for (int c = 0; c < P; ++c) gt += gtc[c];
    u64 r = k - gt;
    std::vector<u64> wofs(P + 1, 0), tofs(P + 1, 0);
    for (int c = 0; c < P; ++c) {
        tofs[c + 1] = tofs[c] + eqc[c];
        u64 tk = tofs[c] >= r ? 0 : std::min(eqc[c], r - tofs[c]);
        wofs[c + 1] = wofs[c] + gtc[c] + tk;
    }
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < P; ++c) {
        u64 b = std::min(n, (u64)c * chunk), e = std::min(n, b + chunk), w = wofs[c], t = tofs[c];
        for (u64 i = b; i < e; ++i) {
            u32 v = keys[i];
            bool take = v > K || (v == K && t++ < r);
            if (take) { if (out_keys) out_keys[w] = v; if (out_idx) out_idx[w] = (i64)i; ++w; }
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (kth_out) *kth_out = K;
    return std::chrono::duration<double>(t1 - t0).count();
}

extern "C" int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// ---------------------------------------------------------------------------
// 5. Exact checker of a finished top-k (tests only).  The k-th largest K is the one key with
//    #(v > K) < k <= #(v >= K); the top-k set is every index with v > K plus the r = k - #(v > K)
//    lowest indices with v == K, in ascending index order.  One pass over V counts keys > K and
//    == K per chunk, a second walks each chunk and compares every taken index with the output at
//    its position.  Keys, kth and outputs are raw bits of the given type (float32: ordered by the
//    sign flip).  Returns 0 when (kth, out_keys, out_idx) is exactly that result; otherwise the
//    first failed condition: 1 bad arguments, 2 K is not the k-th largest, 3 an output index
//    differs, 4 an output key differs.  out_keys / out_idx may be NULL (not checked).
// ---------------------------------------------------------------------------
static int verify_core(const u32* raw, u64 n, u64 k, u32 kth_raw, const u32* out_keys, const i64* out_idx,
                       int f32, int P) {
    if (!raw || k == 0 || k > n) return 1;
    const u32 K = f32 ? f2o(kth_raw) : kth_raw;
    if (P < 1) P = 1;
    const u64 chunk = (n + P - 1) / P;
    std::vector<u64> gtc(P, 0), eqc(P, 0);
#pragma omp parallel for schedule(static, 1) if (P > 1)
    for (int c = 0; c < P; ++c) {
        const u64 b = std::min(n, (u64)c * chunk), e = std::min(n, b + chunk);
        u64 g = 0, q = 0;
        for (u64 i = b; i < e; ++i) {
            const u32 v = f32 ? f2o(raw[i]) : raw[i];
            g += v > K;
            q += v == K;
        }
        gtc[c] = g;
        eqc[c] = q;
    }
    u64 gt = 0, eq = 0;
    for (int c = 0; c < P; ++c) { gt += gtc[c]; eq += eqc[c]; }
    if (!(gt < k && k <= gt + eq)) return 2;
    if (!out_keys && !out_idx) return 0;
    const u64 r = k - gt;
    std::vector<u64> wofs(P + 1, 0), tofs(P + 1, 0);
    for (int c = 0; c < P; ++c) {
        tofs[c + 1] = tofs[c] + eqc[c];
        const u64 tk = tofs[c] >= r ? 0 : std::min(eqc[c], r - tofs[c]);
        wofs[c + 1] = wofs[c] + gtc[c] + tk;
    }
    std::vector<int> bad(P, 0);
#pragma omp parallel for schedule(static, 1) if (P > 1)
    for (int c = 0; c < P; ++c) {
        const u64 b = std::min(n, (u64)c * chunk), e = std::min(n, b + chunk);
        u64 w = wofs[c], t = tofs[c];
        for (u64 i = b; i < e && !bad[c]; ++i) {
            const u32 v = f32 ? f2o(raw[i]) : raw[i];
            const bool take = v > K || (v == K && t++ < r);
            if (!take) continue;
            if (out_idx && out_idx[w] != (i64)i) bad[c] = 3;
            else if (out_keys && out_keys[w] != raw[i]) bad[c] = 4;
            ++w;
        }
    }
    for (int c = 0; c < P; ++c)
        if (bad[c]) return bad[c];
    return 0;
}

extern "C" int orc_verify_topk(const u32* keys, u64 n, u64 k, u32 kth_raw, const u32* out_keys, const i64* out_idx,
                               int f32, int threads) {
#ifdef _OPENMP
    const int P = threads > 0 ? threads : omp_get_max_threads();
#else
    const int P = 1;
#endif
    return verify_core(keys, n, k, kth_raw, out_keys, out_idx, f32, std::max(1, std::min<int>(P, (int)std::min<u64>(n, 4096))));
}

// The checker over many rows of one key buffer, rows in parallel: row r holds n[r] keys at
// offset off[r], its k-th key is kth[r] and its k outputs sit at out_off[r].  codes[r] gets the
// row's result (0 = exact).  Returns the number of rows that failed.
extern "C" u64 orc_verify_rows(const u32* base, u64 nrows, const u64* off, const u64* n, const u64* k,
                               const u64* out_off, const int* f32, const u32* kth, const u32* out_keys,
                               const i64* out_idx, int* codes) {
    u64 nbad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nbad)
    for (i64 r = 0; r < (i64)nrows; ++r) {
        const int c = verify_core(base + off[r], n[r], k[r], kth[r], out_keys ? out_keys + out_off[r] : nullptr,
                                  out_idx ? out_idx + out_off[r] : nullptr, f32[r], 1);
        codes[r] = c;
        nbad += c != 0;
    }
    return nbad;
}
