// api.hip -- host orchestration and the C ABI (include/drtopk.h).
//
// The reference drives each phase from the host with a cudaDeviceSynchronize, a 1 KB
// device-to-host copy and a host scan per radix digit (radixselect.cuh:1427-1445,
// :192-220).  Here the whole pipeline is one stream-ordered sequence of kernels: every
// digit / bucket decision is made by k_select on the device and read by the next kernel
// from the control block, so a call issues no host synchronisation (unless the caller
// asks for stats) and can be captured in a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/drtopk.h"
#include "common.hpp"
#include "host_util.hpp"
#include "kernels.hpp"

using namespace drtopk;

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace drtopk {
drtopk_status fail(drtopk_status s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return s;
}
void clear_error() { g_last_error.clear(); }
}  // namespace drtopk

namespace {

constexpr u64 MAX_N = (1ull << 31) - 1;

// drtopk_testing_set_segment_cap: entries a concatenation segment may hold (0: its capacity).
std::atomic<u64> g_test_segcap{0};

u64 real_delegates(u64 n, int alpha, int beta) {
    u64 L = 1ull << alpha, S = cdiv(n, L), last = n - (S - 1) * L;
    return (u64)beta * (S - 1) + std::min<u64>((u64)beta, last);
}

// Resolved problem + workspace layout.
struct Plan {
    drtopk_algo algo;
    int f32;
    u64 n, k;
    int alpha, beta, path, nb, B;
    u32 flags;
    u64 L, S, D;
    u64 segcap, nseg;  // segmented concatenation (path 1: gather tiles, path 2: beta tiles)
    u64 ext_nseg;      // extraction segments
    u64 range_nt;      // range-selection tiles
    int hist_bins;     // histogram buffer words
    // byte offsets
    u64 o_ctrl, o_hist, o_zero_end;
    u64 o_rcnt, o_roff, o_ecnt, o_eoff, o_scnt;
    u64 npart, o_part;  // beta path: per-workgroup records of the concatenation
    u64 o_dvals, o_doff, o_sel, o_skeys, o_sidx, total;
    u64 npad1, npad2, o_bufA, o_bufB, o_dsel;  // bitonic family
};

#include "single_tuning.inc"

// The measured (alpha, beta) of the radix family for (floor log2 n, ceil log2 k), if any.
bool single_tuned(u64 n, u64 k, int* alpha, int* beta) {
    const int ln = floor_log2(n), lk = ceil_log2(k);
    for (const auto& c : kSingleTuned)
        if (c[2] != 0 && c[0] == ln && c[1] == lk) {
            *alpha = c[2];
            *beta = c[3];
            return true;
        }
    return false;
}

// Parameter policy: radix alpha = floor(0.5*(log2|V| - log2 k + 3)), beta path iff alpha<=5
// (baseline+filter+beta+shuffle/main.cu:83-89); bucket: alpha++ when alpha<5
// (bucket_select/main.cu:66-67); bitonic: user alpha, auto = radix formula in [5,16].
drtopk_status resolve(drtopk_algo algo, drtopk_dtype dtype, u64 n, u64 k, const drtopk_params* p, Plan* pl) {
    if (algo < DRTOPK_ALGO_RADIX || algo > DRTOPK_ALGO_BITONIC) return fail(DRTOPK_ERR_INVALID_ARG, "bad algo %d", (int)algo);
    if (dtype != DRTOPK_U32 && dtype != DRTOPK_F32) return fail(DRTOPK_ERR_INVALID_ARG, "bad dtype %d", (int)dtype);
    if (!p) return fail(DRTOPK_ERR_INVALID_ARG, "params is NULL");
    if (p->reserved[0] || p->reserved[1] || p->reserved[2]) return fail(DRTOPK_ERR_INVALID_ARG, "reserved params must be 0");
    if (n == 0) return fail(DRTOPK_ERR_INVALID_ARG, "n == 0");
    if (n > MAX_N) return fail(DRTOPK_ERR_UNSUPPORTED, "n = %llu exceeds 2^31-1", (unsigned long long)n);
    if (k == 0 || k > n) return fail(DRTOPK_ERR_INVALID_ARG, "k = %llu outside [1, n = %llu]", (unsigned long long)k, (unsigned long long)n);
    const int nb = p->digit_bits;
    if (nb != 1 && nb != 2 && nb != 4 && nb != 8 && nb != 16) return fail(DRTOPK_ERR_INVALID_ARG, "digit_bits %d not in {1,2,4,8,16}", nb);
    if (p->num_buckets < 4 || p->num_buckets > 4096) return fail(DRTOPK_ERR_INVALID_ARG, "num_buckets %d not in [4, 4096]", p->num_buckets);
    if (p->beta < 1 || p->beta > 4) return fail(DRTOPK_ERR_INVALID_ARG, "beta %d not in [1, 4]", p->beta);
    if (p->alpha < -2 || p->alpha > 30) return fail(DRTOPK_ERR_INVALID_ARG, "alpha %d not in [-2, 30]", p->alpha);
    if (algo == DRTOPK_ALGO_BITONIC && (k & (k - 1)))
        return fail(DRTOPK_ERR_INVALID_ARG, "bitonic top-k needs a power-of-two k (got %llu)", (unsigned long long)k);
    const int Ln = floor_log2(n);
    int alpha = p->alpha;
    // measured choice: its beta replaces p->beta, and beta 1 means one delegate per subrange
    int tuned_beta = 0;
    if (alpha == DRTOPK_ALPHA_TUNED) {
        int ta = 0, tb = 0;
        if (algo == DRTOPK_ALGO_RADIX && single_tuned(n, k, &ta, &tb)) {
            alpha = ta;
            tuned_beta = tb;
        } else {
            alpha = DRTOPK_ALPHA_AUTO;
        }
    }
    const bool autoa = alpha < 0 || tuned_beta > 0;  // a tuned alpha may be lowered like the policy's
    if (tuned_beta > 0 && alpha > Ln) alpha = Ln;
    if (alpha < 0) {
        alpha = (Ln + 3 - ceil_log2(k)) / 2;
        if (alpha < 0) alpha = 0;
        if (algo == DRTOPK_ALGO_BUCKET && alpha < 5) alpha++;
        if (algo == DRTOPK_ALGO_BITONIC) alpha = std::min(16, std::max(5, alpha));
        if (alpha > Ln) alpha = Ln;
    }
    if (alpha > Ln) return fail(DRTOPK_ERR_INVALID_ARG, "alpha %d larger than log2(n) = %d", alpha, Ln);
    const int beta_in = tuned_beta > 0 ? tuned_beta : p->beta;
    const bool beta_flag = tuned_beta > 0 ? tuned_beta >= 2 : (p->flags & DRTOPK_FLAG_BETA) != 0;
    auto want_beta = [&](int a) {
        return algo != DRTOPK_ALGO_BITONIC && beta_flag && a <= 5 && beta_in >= 2;
    };
    auto valid = [&](int a) -> bool {
        if (a == 0) return true;
        u64 L = 1ull << a, S = cdiv(n, L);
        if (algo == DRTOPK_ALGO_BITONIC) return S > k;
        if (want_beta(a)) return beta_in <= (int)L && real_delegates(n, a, beta_in) >= k;
        return S >= k;
    };
    if (!valid(alpha)) {
        if (!autoa)
            return fail(DRTOPK_ERR_INVALID_ARG, "alpha %d leaves fewer than k = %llu delegates", alpha,
                        (unsigned long long)k);
        while (alpha > 0 && !valid(alpha)) --alpha;
    }
    int beta = 1, path = alpha == 0 ? 0 : 1;
    if (alpha > 0 && want_beta(alpha)) { beta = beta_in; path = 2; }
    pl->algo = algo;
    pl->f32 = dtype == DRTOPK_F32;
    pl->n = n; pl->k = k;
    pl->alpha = alpha; pl->beta = beta; pl->path = path; pl->nb = nb; pl->B = p->num_buckets;
    pl->flags = p->flags;
    pl->L = 1ull << alpha;
    pl->S = alpha ? cdiv(n, pl->L) : 0;
    pl->D = (u64)beta * pl->S;
    // layout
    pl->segcap = 0;
    pl->nseg = 0;
    pl->npad1 = pl->npad2 = 0;
    const bool bit = algo == DRTOPK_ALGO_BITONIC;
    auto pow2 = [](u64 x) { u64 p = 1; while (p < x) p <<= 1; return p; };
    if (bit) {
        pl->npad1 = path ? pow2(pl->S) : 0;
        pl->npad2 = pow2(path ? (k << alpha) : n);
    } else if (path == 1) {
        pl->segcap = gather_segcap(k * pl->L);
        pl->nseg = gather_nseg(k * pl->L);
    } else if (path == 2) {
        pl->segcap = beta_segcap(pl->S, alpha);
        pl->nseg = beta_nseg(pl->S, alpha);
    }
    pl->ext_nseg = bit ? 0 : (path == 0 ? cdiv(n, DIRECT_TILE) : pl->nseg);
    pl->range_nt = (path == 1 && !bit) ? range_tiles(pl->S) : 0;
    pl->hist_bins = std::max<int>(HIST_REPS * 256, std::max<int>(1 << nb, pl->B));
    u64 o = 0;
    pl->o_ctrl = o; o = align_up(o + sizeof(Ctrl), 256);
    pl->o_hist = o; o = align_up(o + 4ull * pl->hist_bins, 256);
    pl->o_zero_end = o;
    pl->o_rcnt = o; o = align_up(o + 8ull * pl->range_nt, 256);
    pl->o_roff = o; o = align_up(o + 8ull * pl->range_nt, 256);
    pl->o_ecnt = o; o = align_up(o + 8ull * pl->ext_nseg, 256);
    pl->o_eoff = o; o = align_up(o + 8ull * pl->ext_nseg, 256);
    pl->o_scnt = o; o = align_up(o + 4ull * pl->nseg, 256);
    pl->npart = (path == 2 && !bit) ? beta_grid(pl->S, alpha) : 0;
    pl->o_part = o; o = align_up(o + 32ull * pl->npart, 256);
    pl->o_dvals = o; o = align_up(o + 4ull * pl->D, 256);
    const u64 offb = path == 2 ? (beta == 2 ? 1 : beta == 3 ? 2 : 4) : 0;
    pl->o_doff = o; o = align_up(o + offb * pl->S, 256);
    pl->o_sel = o; o = align_up(o + ((path == 1 && !bit) ? 4ull * k : 0), 256);
    pl->o_skeys = o; o = align_up(o + 4ull * pl->segcap * pl->nseg, 256);
    pl->o_sidx = o; o = align_up(o + 4ull * pl->segcap * pl->nseg, 256);
    const u64 nbuf = std::max(pl->npad1, pl->npad2);
    pl->o_bufA = o; o = align_up(o + 8ull * nbuf, 256);
    pl->o_bufB = o; o = align_up(o + 8ull * nbuf, 256);
    pl->o_dsel = o; o = align_up(o + (bit && path ? 8ull * k : 0), 256);
    pl->total = o;
    return DRTOPK_OK;
}

SelState make_sel(int mode, u64 m, u64 k, int nb, int B, int stop_digit, int stop_on_one, int max_passes) {
    SelState s;
    std::memset(&s, 0, sizeof(s));
    s.m = m;
    s.k = k;
    if (mode == 0) {
        s.nb = (u32)nb;
        s.digit = 32 / nb - 1;
        s.stop_digit = (u32)stop_digit;
        if (s.digit < stop_digit) { s.done = 1; s.result = 0; }
    } else {
        s.nb = (u32)B;
        s.lo = 0xFFFFFFFFu;
        s.hi = 0;
        s.max_passes = (u32)max_passes;
    }
    s.stop_on_one = (u32)stop_on_one;
    return s;
}

int bucket_max_levels(int B) {
    return (int)std::ceil(32.0 / std::log2((double)(B - 1))) + 2;
}

}  // namespace

extern "C" {

void drtopk_params_default(drtopk_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->alpha = DRTOPK_ALPHA_AUTO;
    p->beta = 2;
    p->digit_bits = 8;
    p->num_buckets = 256;
    p->flags = DRTOPK_FLAGS_DEFAULT;
}

const char* drtopk_last_error(void) { return g_last_error.c_str(); }

const char* drtopk_version(void) { return "0.1 gfx950"; }

drtopk_status drtopk_workspace_size(drtopk_algo algo, drtopk_dtype dtype, uint64_t n, uint64_t k,
                                    const drtopk_params* p, size_t* bytes) {
    if (!bytes) return fail(DRTOPK_ERR_INVALID_ARG, "bytes is NULL");
    Plan pl;
    drtopk_status s = resolve(algo, dtype, n, k, p, &pl);
    if (s != DRTOPK_OK) return s;
    *bytes = (size_t)pl.total;
    return DRTOPK_OK;
}

}  // extern "C"

namespace {

// The whole pipeline of one top-k on a resolved plan, into a workspace of at least pl.total
// bytes (checked by the caller).
drtopk_status select_impl(const Plan& pl, const void* d_keys, void* d_workspace, void* d_kth, void* d_topk_keys,
                          int64_t* d_topk_idx, drtopk_stats* h_stats, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    char* ws = (char*)d_workspace;
    Ctrl* ctrl = (Ctrl*)(ws + pl.o_ctrl);
    u32* hist = (u32*)(ws + pl.o_hist);
    u64* rcnt = (u64*)(ws + pl.o_rcnt);
    u64* roff = (u64*)(ws + pl.o_roff);
    u64* ecnt = (u64*)(ws + pl.o_ecnt);
    u64* eoff = (u64*)(ws + pl.o_eoff);
    u32* scnt = (u32*)(ws + pl.o_scnt);
    u32* dvals = (u32*)(ws + pl.o_dvals);
    void* doff = (void*)(ws + pl.o_doff);
    u32* sel_ids = (u32*)(ws + pl.o_sel);
    u32* skeys = (u32*)(ws + pl.o_skeys);
    u32* sidx = (u32*)(ws + pl.o_sidx);
    u64* part = pl.npart ? (u64*)(ws + pl.o_part) : nullptr;
    const u32* keys = (const u32*)d_keys;
    const bool want_out = d_topk_keys || d_topk_idx;
    const bool radix = pl.algo == DRTOPK_ALGO_RADIX;
    const int mode = radix ? 0 : 1;
    const int nb = pl.nb, B = pl.B;
    const int D0 = 32 / nb - 1;
    const int bins = radix ? (1 << nb) : B;
    const bool fuse_hist = radix && nb <= 8;
    // The first digit's histogram rides on the delegate pass, except for alpha 6..8: there the
    // lane-group sampler produces one delegate per 16..64 lanes and the per-wave histogram update
    // costs it more (25-60 us at 2^30 keys) than a separate pass over the delegates (11-17 us).
    const bool fuse_first = fuse_hist && !(pl.alpha >= 6 && pl.alpha <= 8);
    // beta path, radix, uint32 keys, 8-bit digits, beta 2, FILTER, fused first digit: the
    // threshold is a top-digit boundary and no pass reads the delegates but the concatenation,
    // which needs the second delegate only against it (lonely delegates below it are dropped),
    // so the delegates are stored as the first one plus one 16-bit word holding the second's top
    // byte and the first's offset (6 bytes per subrange, not 8 + 1).  Float keys keep full
    // delegates: their top byte (sign and seven exponent bits) often leaves a dense threshold,
    // which takes one digit more.
    const bool top_form = pl.path == 2 && radix && !pl.f32 && nb == 8 && pl.beta == 2 &&
                          (pl.flags & DRTOPK_FLAG_FILTER) && fuse_first;
    const int Lmax = bucket_max_levels(B);
    const bool timing = (pl.flags & DRTOPK_FLAG_TIMING) && h_stats;

    hipEvent_t ev[6] = {};
    if (timing)
        for (int i = 0; i < 6; ++i) HIPCHK(hipEventCreate(&ev[i]));
    auto mark = [&](int i) -> hipError_t { return timing ? hipEventRecord(ev[i], st) : hipSuccess; };

    // ---- control block
    Ctrl init;
    std::memset(&init, 0, sizeof(init));
    if (pl.path == 0) {
        init.second = make_sel(mode, pl.n, pl.k, nb, B, 0, radix ? 0 : 1, 1 << 20);
    } else {
        if (pl.path == 1)
            init.first = make_sel(mode, pl.S, pl.k, nb, B, 0, 1, 1 << 20);
        else {
            init.first = make_sel(mode, pl.D, pl.k, nb, B, 3, 0, 2);
            // full delegates: one more digit over a dense threshold
            if (radix && !top_form) init.first.dense_limit = 4 * pl.k;
        }
        init.second = make_sel(mode, 0, pl.k, nb, B, 0, radix ? 0 : 1, 1 << 20);
    }
    HIPCHK(launch_init(ctrl, init, (u32*)(ws + pl.o_hist), (pl.o_zero_end - pl.o_hist) / 4, st));
    HIPCHK(mark(0));

    if (pl.algo == DRTOPK_ALGO_BITONIC) {
        // ---- bitonic family (bitonic/sampleBitonic.cu:283-308)
        if (pl.path != 0) {
            SampleArgs sa;
            sa.keys = keys; sa.n = pl.n; sa.alpha = pl.alpha; sa.beta = 1; sa.f32 = pl.f32;
            sa.staged = 0; sa.posmode = 0; sa.fuse = 0;
            sa.dvals = dvals; sa.dpos = nullptr; sa.ghist = nullptr; sa.hshift = 0; sa.bins = 0; sa.ctrl = ctrl;
            HIPCHK(launch_sample(sa, st));
        }
        HIPCHK(mark(1));
        BitonicArgs ba{keys, pl.n, pl.alpha, pl.f32, pl.k, pl.S, dvals, ctrl, (u64*)(ws + pl.o_bufA),
                       (u64*)(ws + pl.o_bufB), (u64*)(ws + pl.o_dsel), pl.npad1, pl.npad2, d_topk_keys, d_topk_idx,
                       d_kth};
        HIPCHK(launch_bitonic(ba, st));
        for (int i = 2; i <= 4; ++i) HIPCHK(mark(i));
    } else if (pl.path == 0) {
        // ---- no delegates: radix_select / bucket_select over V (radixselect.cuh:1473,
        //      bucketselect_combined.cuh:473)
        HIPCHK(mark(1));
        HIPCHK(mark(2));
        SelState* sec = &ctrl->second;
        const unsigned g = (unsigned)std::min<u64>(2048, cdiv(pl.n, 4096) + 1);
        if (radix) {
            for (int d = 0; d <= D0; ++d) {
                HIPCHK(launch_hist_raw(0, pl.f32, keys, sec, hist, bins, g, st));
                HIPCHK(launch_select(0, hist, bins, sec, ctrl, st));
            }
        } else {
            HIPCHK(launch_minmax_raw(pl.f32, keys, sec, ctrl, g, st));
            HIPCHK(launch_bucket_begin(sec, ctrl, 0, st));
            for (int l = 0; l < Lmax; ++l) {
                HIPCHK(launch_hist_raw(1, pl.f32, keys, sec, hist, bins, g, st));
                HIPCHK(launch_select(1, hist, bins, sec, ctrl, st));
            }
            HIPCHK(launch_find(keys, pl.f32 ? 1 : 0, sec, g, st));
        }
        HIPCHK(mark(3));
        if (want_out) {  // the extraction also writes the k-th key and the counts
            ExtractArgs ea{Seg{keys, nullptr, nullptr, DIRECT_TILE, pl.ext_nseg, pl.n}, 1, pl.f32, pl.k, sec, ctrl,
                           d_topk_keys, d_topk_idx, ecnt, eoff, true, d_kth};
            HIPCHK(launch_extract(ea, st));
        } else {
            HIPCHK(launch_finalize(sec, ctrl, pl.f32, d_kth, pl.k, st));
        }
        HIPCHK(mark(4));
    } else {
        SelState* fst = &ctrl->first;
        SelState* sec = &ctrl->second;
        // ---- delegate pass (sampleMax / sampleMax_NoReduction_Beta*, radixselect.cuh:1787-1809)
        SampleArgs sa;
        sa.keys = keys; sa.n = pl.n; sa.alpha = pl.alpha; sa.beta = pl.beta; sa.f32 = pl.f32;
        sa.staged = (pl.flags & DRTOPK_FLAG_SHUFFLE) && pl.alpha <= 5;
        sa.posmode = (pl.path == 2 && want_out) ? 2 : 0;
        sa.fuse = radix ? (fuse_first ? 1 : 0) : 2;
        sa.dvals = dvals; sa.dpos = doff; sa.ghist = hist;
        sa.hshift = 32 - nb; sa.bins = bins; sa.ctrl = ctrl;
        // 16-bit words after the first delegates, 16-byte aligned, inside the 8*S-byte area
        u8* dtop = top_form ? (u8*)(dvals + align_up(pl.S, 4)) : nullptr;
        sa.dtop = dtop;
        HIPCHK(launch_sample(sa, st));
        HIPCHK(mark(1));
        // ---- first top-k over the delegates
        const u64 Dm = pl.path == 1 ? pl.S : pl.D;
        const unsigned gd = (unsigned)std::min<u64>(2048, cdiv(Dm, 4096) + 1);
        if (radix) {
            // default path: digits until the k-digit bucket holds one delegate
            // (radix_select_inplace_firsttopk :1667-1724); beta path: digits >= 3 (:1747)
            // (beta path with full delegates: one digit more, run only when the threshold is dense)
            const int ndig = pl.path == 1 ? D0 + 1 : (D0 >= 3 ? D0 - 3 + 1 + (top_form ? 0 : 1) : 0);
            for (int d = 0; d < ndig; ++d) {
                if (!(d == 0 && fuse_first)) HIPCHK(launch_hist(0, dvals, fst, hist, bins, gd, st));
                HIPCHK(launch_select(0, hist, bins, fst, ctrl, st));
            }
        } else {
            HIPCHK(launch_bucket_begin(fst, ctrl, 1, st));
            const int lv = pl.path == 1 ? Lmax : 2;
            for (int l = 0; l < lv; ++l) {
                HIPCHK(launch_hist(1, dvals, fst, hist, bins, gd, st));
                HIPCHK(launch_select(1, hist, bins, fst, ctrl, st));
            }
            if (pl.path == 1) HIPCHK(launch_find(dvals, 0, fst, gd, st));
        }
        HIPCHK(mark(2));
        // ---- range selection + concatenation (segmented, one segment per tile)
        u32* h2 = fuse_hist ? hist : nullptr;
        const u64 seg_limit = g_test_segcap.load(std::memory_order_relaxed);
        u32* sidx_or_null = want_out ? sidx : nullptr;
        if (pl.path == 1) {
            RangeArgs ra{dvals, pl.S, pl.k, radix ? 0 : 1, fst, ctrl, sel_ids, rcnt, roff};
            HIPCHK(launch_range(ra, st));
            GatherArgs ga{keys, pl.n, pl.alpha, pl.f32, (pl.flags & DRTOPK_FLAG_FILTER) ? 1 : 0, sel_ids, pl.k,
                          ctrl, skeys, sidx_or_null, scnt, h2, 32 - nb, bins, seg_limit};
            HIPCHK(launch_gather(ga, st));
        } else {
            // filter bit 1: keep keys >= threshold; bit 2 (radix family): lonely delegates below the
            // threshold are counted, not written
            const int bfilter = (pl.flags & DRTOPK_FLAG_FILTER) ? (radix ? 3 : 1) : 0;
            BetaGatherArgs ba{keys, pl.n, pl.alpha, pl.beta, pl.f32, bfilter,
                              dvals, doff, ctrl, skeys, sidx_or_null, scnt, h2, 32 - nb, bins, dtop, seg_limit, part};
            HIPCHK(launch_beta_gather(ba, st));
        }
        // with the first digit's histogram fused into the concatenation, the first select also
        // totals the segments; otherwise a one-workgroup pass does before the first histogram
        const bool total_in_select = radix && fuse_hist;
        if (!total_in_select) HIPCHK(launch_seg_total(scnt, pl.nseg, ctrl, st, part, pl.npart));
        HIPCHK(mark(3));
        // ---- second top-k over the concatenation (radix_select_inplace :1891 / bucket_select :1355)
        const Seg cseg{skeys, sidx, scnt, pl.segcap, pl.nseg, 0};
        const unsigned gc = (unsigned)std::min<u64>(2048, pl.nseg);
        if (radix) {
            for (int d = 0; d <= D0; ++d) {
                if (!(d == 0 && fuse_hist)) HIPCHK(launch_hist_seg(0, cseg, sec, hist, bins, gc, st));
                if (d == 0 && total_in_select)
                    HIPCHK(launch_select(0, hist, bins, sec, ctrl, st, scnt, pl.nseg, part, pl.npart));
                else HIPCHK(launch_select(0, hist, bins, sec, ctrl, st));
            }
        } else {
            HIPCHK(launch_minmax_seg(cseg, sec, gc, st));
            HIPCHK(launch_bucket_begin(sec, ctrl, 0, st));
            for (int l = 0; l < Lmax; ++l) {
                HIPCHK(launch_hist_seg(1, cseg, sec, hist, bins, gc, st));
                HIPCHK(launch_select(1, hist, bins, sec, ctrl, st));
            }
            HIPCHK(launch_find_seg(cseg, sec, gc, st));
        }
        HIPCHK(mark(4));
        if (want_out) {  // the extraction also writes the k-th key and the counts
            ExtractArgs ea{cseg, 0, pl.f32, pl.k, sec, ctrl, d_topk_keys, d_topk_idx, ecnt, eoff, true, d_kth};
            HIPCHK(launch_extract(ea, st));
        } else {
            HIPCHK(launch_finalize(sec, ctrl, pl.f32, d_kth, pl.k, st));
        }
    }
    HIPCHK(mark(5));

    if (h_stats) {
        Ctrl hc;
        HIPCHK(hipMemcpyAsync(&hc, ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        drtopk_stats o;
        std::memset(&o, 0, sizeof(o));
        o.algo = (int)pl.algo; o.alpha = pl.alpha; o.beta = pl.beta; o.path = pl.path;
        o.digit_bits = nb; o.num_buckets = B;
        o.first_passes = pl.path == 0 ? 0 : (int)hc.first.passes;
        o.second_passes = (int)hc.second.passes;
        o.num_subranges = pl.S; o.num_delegates = pl.D;
        o.threshold = pl.path == 0 ? 0 : hc.first.result;
        o.kth_bits = hc.kth;
        o.selected_subranges = hc.selected;
        o.lonely_subranges = hc.lonely_sub;
        o.lonely_elements = hc.lonely_el;
        o.concat_size = pl.path == 0 ? pl.n : hc.concat_size;
        o.count_greater = hc.count_gt;
        o.ties_taken = hc.ties;
        if (timing) {
            // phases: sample [0,1], first [1,2], concatenate [2,3], second [3,4], extract [4,5]
            for (int i = 0; i < 5; ++i) {
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
                o.phase_ms[i] = ms;
            }
            for (int i = 0; i < 6; ++i) hipEventDestroy(ev[i]);
        }
        *h_stats = o;
        if (hc.head.error)
            return fail(DRTOPK_ERR_INTERNAL, "device-side check failed (code 0x%x)", hc.head.error);
    }
    return DRTOPK_OK;
}

drtopk_status check_keys(const void* d_keys) {
    if (!d_keys) return fail(DRTOPK_ERR_INVALID_ARG, "d_keys is NULL");
    if (((uintptr_t)d_keys & 15) != 0) return fail(DRTOPK_ERR_INVALID_ARG, "d_keys must be 16-byte aligned");
    return DRTOPK_OK;
}

// ---- scratch of drtopk_select_static: one file-scope device array per GPU.  In host code the
// name `scratch` is only a stand-in for the symbol; its device address, different on every
// GPU, comes from hipGetSymbolAddress and is what every kernel argument uses.
constexpr size_t STATIC_SCRATCH_BYTES = (size_t)DRTOPK_STATIC_SCRATCH_BYTES;
__device__ __align__(256) u8 scratch[STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::atomic<void*> g_scratch_ptr[MAX_DEVICES];
std::mutex g_scratch_mu;

// Device address of the current device's scratch, looked up once per device.
drtopk_status static_scratch(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(DRTOPK_ERR_UNSUPPORTED, "device ordinal %d >= %d", dev, MAX_DEVICES);
    void* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
        if (!p) {
            void* q = nullptr;
            HIPCHK(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
            if (!q || ((uintptr_t)q & 255) != 0)
                return fail(DRTOPK_ERR_HIP, "static scratch address %p is not 256-byte aligned", q);
            g_scratch_ptr[dev].store(q, std::memory_order_release);
            p = q;
        }
    }
    *out = p;
    return DRTOPK_OK;
}

}  // namespace

extern "C" {

drtopk_status drtopk_select(drtopk_algo algo, drtopk_dtype dtype, const void* d_keys, uint64_t n, uint64_t k,
                            const drtopk_params* p, void* d_workspace, size_t workspace_bytes, void* d_kth,
                            void* d_topk_keys, int64_t* d_topk_idx, drtopk_stats* h_stats, void* stream) {
    clear_error();
    Plan pl;
    drtopk_status s = resolve(algo, dtype, n, k, p, &pl);
    if (s != DRTOPK_OK) return s;
    if ((s = check_keys(d_keys)) != DRTOPK_OK) return s;
    if (!d_workspace) return fail(DRTOPK_ERR_INVALID_ARG, "d_workspace is NULL");
    if (((uintptr_t)d_workspace & 255) != 0) return fail(DRTOPK_ERR_INVALID_ARG, "d_workspace must be 256-byte aligned");
    if (workspace_bytes < pl.total)
        return fail(DRTOPK_ERR_WORKSPACE_TOO_SMALL, "workspace %zu bytes < %llu needed", workspace_bytes,
                    (unsigned long long)pl.total);
    return select_impl(pl, d_keys, d_workspace, d_kth, d_topk_keys, d_topk_idx, h_stats, stream);
}

drtopk_status drtopk_select_static(drtopk_algo algo, drtopk_dtype dtype, const void* d_keys, uint64_t n,
                                   uint64_t k, const drtopk_params* p, void* d_kth, void* d_topk_keys,
                                   int64_t* d_topk_idx, drtopk_stats* h_stats, void* stream) {
    clear_error();
    Plan pl;
    drtopk_status s = resolve(algo, dtype, n, k, p, &pl);
    if (s != DRTOPK_OK) return s;
    if ((s = check_keys(d_keys)) != DRTOPK_OK) return s;
    if (pl.total > STATIC_SCRATCH_BYTES)
        return fail(DRTOPK_ERR_WORKSPACE_TOO_SMALL,
                    "n = %llu, k = %llu needs %llu bytes of scratch; the static scratch holds %llu "
                    "(use drtopk_select with a workspace)",
                    (unsigned long long)n, (unsigned long long)k, (unsigned long long)pl.total,
                    (unsigned long long)STATIC_SCRATCH_BYTES);
    if (pl.flags & DRTOPK_FLAG_TIMING)
        return fail(DRTOPK_ERR_INVALID_ARG, "DRTOPK_FLAG_TIMING creates HIP events; not available here");
    void* ws = nullptr;
    if ((s = static_scratch(&ws)) != DRTOPK_OK) return s;
    return select_impl(pl, d_keys, ws, d_kth, d_topk_keys, d_topk_idx, h_stats, stream);
}

drtopk_status drtopk_workspace_check(const void* d_workspace, void* stream, uint32_t* h_error) {
    clear_error();
    if (h_error) *h_error = 0;
    if (!d_workspace || ((uintptr_t)d_workspace & 255) != 0)
        return fail(DRTOPK_ERR_INVALID_ARG, "d_workspace must be a 256-byte aligned device pointer");
    hipStream_t st = (hipStream_t)stream;
    WsHead h;
    HIPCHK(hipMemcpyAsync(&h, d_workspace, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (h.magic != WS_MAGIC) return fail(DRTOPK_ERR_INVALID_ARG, "the workspace holds no call's result");
    const u32 bits = h.sticky;  // every failed check also sets the call's own word
    if (h_error) *h_error = bits;
    if (h.sticky) {  // read and clear: the next check reports only calls made after this one
        static const u32 zero = 0;
        HIPCHK(hipMemcpyAsync((char*)d_workspace + offsetof(WsHead, sticky), &zero, sizeof(zero),
                              hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    if (bits)
        return fail(DRTOPK_ERR_INTERNAL, "device-side check failed in a call on this workspace (bits 0x%x:%s%s)",
                    bits, (bits & ERR_CONCAT_SHORT) ? " concatenation shorter than k" : "",
                    (bits & ERR_CAPACITY) ? " a buffer bound was exceeded" : "");
    return DRTOPK_OK;
}

drtopk_status drtopk_testing_set_segment_cap(uint64_t cap) {
    clear_error();
    g_test_segcap.store(cap, std::memory_order_relaxed);
    return DRTOPK_OK;
}

drtopk_status drtopk_sample_delegates(drtopk_dtype dtype, const void* d_keys, uint64_t n, int32_t alpha,
                                      int32_t beta, uint32_t flags, void* d_vals, uint32_t* d_pos, void* stream) {
    clear_error();
    if (dtype != DRTOPK_U32 && dtype != DRTOPK_F32) return fail(DRTOPK_ERR_INVALID_ARG, "bad dtype");
    if (!d_keys || !d_vals) return fail(DRTOPK_ERR_INVALID_ARG, "NULL pointer");
    if (((uintptr_t)d_keys & 15) != 0) return fail(DRTOPK_ERR_INVALID_ARG, "d_keys must be 16-byte aligned");
    if (n == 0) return fail(DRTOPK_ERR_INVALID_ARG, "n == 0");
    if (n > MAX_N) return fail(DRTOPK_ERR_UNSUPPORTED, "n exceeds 2^31-1");
    if (alpha < 1 || alpha > 30 || (1ull << alpha) > std::max<u64>(n, 1)) return fail(DRTOPK_ERR_INVALID_ARG, "alpha %d out of range", alpha);
    if (beta < 1 || beta > 4 || beta > (1 << alpha)) return fail(DRTOPK_ERR_INVALID_ARG, "beta %d out of range", beta);
    hipStream_t st = (hipStream_t)stream;
    SampleArgs sa;
    sa.keys = (const u32*)d_keys; sa.n = n; sa.alpha = alpha; sa.beta = beta; sa.f32 = dtype == DRTOPK_F32;
    sa.staged = (flags & DRTOPK_FLAG_SHUFFLE) && alpha <= 5;
    sa.posmode = d_pos ? 1 : 0;
    sa.fuse = 0;
    sa.dvals = (u32*)d_vals; sa.dpos = d_pos; sa.ghist = nullptr; sa.hshift = 0; sa.bins = 0; sa.ctrl = nullptr;
    HIPCHK(launch_sample(sa, st));
    if (sa.f32) {
        const u64 S = cdiv(n, 1ull << alpha);
        HIPCHK(launch_to_user((const u32*)d_vals, S * (u64)beta, 1, d_vals, st));
    }
    return DRTOPK_OK;
}

}  // extern "C"
