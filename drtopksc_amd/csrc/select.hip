// select.hip -- the device-resident selection machinery of Dr.Top-k.
//
//  k_init        writes the control block (keeping the workspace's sticky error bits) and
//                clears histograms and counters (a kernel, not hipMemsetAsync, so the pipeline
//                can be graph-captured).
//  k_hist        one histogram pass: radix digit of keys matching the current prefix
//                (CalculateOccurence_inplace[_firsttopk], radixselect.cuh:353-410) or
//                linear bucket of keys inside the current range (assignBuckets,
//                bucketselect_combined.cuh:167-197, integer map).
//  k_select      one workgroup: cumulative counts, the k-digit / k-bucket, state update
//                (CumulateCount_inplace radixselect.cuh:192-220 and the host loop
//                :1427-1445; FindKBucket bucketselect_combined.cuh:200-216).  No D2H copy.
//  *_seg         the same passes over a segmented concatenation (compact.hip).
// Range selection, concatenation and top-k extraction live in compact.hip.
#include <algorithm>
#include <atomic>

#include "common.hpp"
#include "kernels.hpp"

namespace drtopk {

constexpr int MAX_LDS_BINS = 4096;

// ---------------------------------------------------------------- init
__global__ void k_init(Ctrl* ctrl, Ctrl init, u32* zero, u64 words) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const WsHead h = ctrl->head;  // the sticky error bits survive the call's reset
        *ctrl = init;
        ctrl->head = h;
        ws_head_begin(&ctrl->head);
    }
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) zero[i] = 0;
}

hipError_t launch_init(Ctrl* ctrl, const Ctrl& init, u32* zero_begin, u64 zero_words, hipStream_t st) {
    unsigned grid = (unsigned)std::min<u64>(1024, (zero_words + WG - 1) / WG + 1);
    hipLaunchKernelGGL(k_init, dim3(grid), dim3(WG), 0, st, ctrl, init, zero_begin, zero_words);
    return hipGetLastError();
}

// ---------------------------------------------------------------- bucket map
// bucket(v) = floor((v - lo) * (B-1) / (hi - lo)), computed with a double estimate and an
// exact integer correction (the estimate is off by at most one).
__device__ __forceinline__ u32 bucket_of(u32 v, u32 lo, u32 W, u32 Bm1, double inv) {
    u64 t = (u64)(v - lo) * Bm1;
    u64 q = (u64)((double)t * inv);
    if (q * W > t) --q;
    else if ((q + 1) * W <= t) ++q;
    return (u32)(q > Bm1 ? Bm1 : q);
}
__device__ __forceinline__ u32 bucket_lo(u64 b, u32 lo, u32 hi, u64 B) {
    u64 W = hi - lo;
    return lo + (u32)((b * W + (B - 2)) / (B - 1));
}
__device__ __forceinline__ u32 bucket_hi(u64 b, u32 lo, u32 hi, u64 B) {
    return b + 1 == B ? hi : bucket_lo(b + 1, lo, hi, B) - 1;
}

// ---------------------------------------------------------------- element iteration
// Calls f(v) for this workgroup's share of a flat array (grid-stride, 16-byte loads) or of
// a segmented array (one segment per workgroup at a time).
// Four 16-byte loads per thread are issued before any is used (indices clamped in bounds, the
// extra ones masked), so a thread keeps four in flight instead of one.
constexpr int HU = 4;
template <class F>
__device__ __forceinline__ void for_flat(const u32* arr, u64 m, F&& f) {
    const u64 nv = m >> 2;
    const u32x4* a4 = (const u32x4*)arr;
    const u64 stride = (u64)gridDim.x * WG * HU;
    for (u64 i0 = (u64)blockIdx.x * WG * HU + threadIdx.x; i0 < nv; i0 += stride) {
        u32x4 r[HU];
#pragma unroll
        for (int u = 0; u < HU; ++u) {
            const u64 i = i0 + (u64)u * WG;
            r[u] = a4[i < nv ? i : 0];
        }
#pragma unroll
        for (int u = 0; u < HU; ++u)
            if (i0 + (u64)u * WG < nv) { f(r[u].x); f(r[u].y); f(r[u].z); f(r[u].w); }
    }
    if (blockIdx.x == 0)
        for (u64 i = (nv << 2) + threadIdx.x; i < m; i += WG) f(arr[i]);
}
template <class F>
__device__ __forceinline__ void for_seg(const Seg& g, F&& f) {
    for (u64 s = blockIdx.x; s < g.nseg; s += gridDim.x) {
        const u64 len = g.cnt ? (u64)g.cnt[s] : std::min<u64>(g.segcap, g.n - s * g.segcap);
        const u32* base = g.keys + s * g.segcap;  // segcap is a multiple of 4: 16-byte aligned
        const u64 nv = len >> 2;
        const u32x4* a4 = (const u32x4*)base;
        for (u64 i0 = threadIdx.x; i0 < nv; i0 += (u64)WG * HU) {
            u32x4 r[HU];
#pragma unroll
            for (int u = 0; u < HU; ++u) {
                const u64 i = i0 + (u64)u * WG;
                r[u] = a4[i < nv ? i : 0];
            }
#pragma unroll
            for (int u = 0; u < HU; ++u)
                if (i0 + (u64)u * WG < nv) { f(r[u].x); f(r[u].y); f(r[u].z); f(r[u].w); }
        }
        for (u64 i = (nv << 2) + threadIdx.x; i < len; i += WG) f(base[i]);
    }
}

// ---------------------------------------------------------------- histogram pass
template <int MODE, bool GLOBALH, class K, bool RAW, bool SEG>
__global__ __launch_bounds__(WG) void k_hist(const u32* __restrict__ arr, Seg seg, SelState* sel, u32* ghist,
                                             int bins) {
    __shared__ u32 lh[GLOBALH ? 1 : MAX_LDS_BINS];
    if (sel->done) return;
    u32 flag = 0, mask = 0, lo = 0, W = 0, Bm1 = 0;
    int shright = 0;
    double inv = 0.0;
    if (MODE == 0) {
        const int nb = (int)sel->nb, d = sel->digit;
        const int shleft = (d + 1) * nb;
        shright = d * nb;
        flag = sel->flag;
        mask = shleft >= 32 ? 0u : ~((1u << shleft) - 1u);
    } else {
        lo = sel->lo;
        W = sel->hi - sel->lo;
        Bm1 = (u32)bins - 1;
        inv = (double)1.0 / (double)W;
    }
    if (!GLOBALH) {
        for (int i = threadIdx.x; i < bins; i += WG) lh[i] = 0;
        __syncthreads();
    }
    u32* h = GLOBALH ? ghist : lh;
    const u32 dmask = (u32)bins - 1;
    // LDS histograms: one peel round, so a wave of equal digits (skewed or tied keys) costs
    // one atomic; the global-memory histogram of 16-bit digits counts key by key
    auto count = [&](u32 raw) {
        const u32 v = RAW ? K::to_ord(raw) : raw;
        if (MODE == 0) {
            const bool hit = (v & mask) == flag;
            const u32 b = (v >> shright) & dmask;
            if (GLOBALH) {
                if (hit) atomicAdd(&h[b], 1u);
            } else {
                wave_hist_add<1>(h, hit, b);
            }
        } else {
            const bool hit = v >= lo && v - lo <= W;
            wave_hist_add<1>(h, hit, hit ? bucket_of(v, lo, W, Bm1, inv) : 0u);
        }
    };
    if (SEG) for_seg(seg, count);
    else for_flat(arr, sel->m, count);
    if (!GLOBALH) {
        __syncthreads();
        u32* g = ghist + (size_t)(blockIdx.x % hist_reps(bins)) * bins;
        for (int i = threadIdx.x; i < bins; i += WG) {
            u32 c = lh[i];
            if (c) atomicAdd(&g[i], c);
        }
    }
}

template <class K, bool RAW, bool SEG>
static hipError_t launch_hist_t(int mode, const u32* arr, const Seg& seg, SelState* sel, u32* ghist, int bins,
                                unsigned grid, hipStream_t st) {
    const bool g = bins > MAX_LDS_BINS;
    if (mode == 0 && !g)
        hipLaunchKernelGGL((k_hist<0, false, K, RAW, SEG>), dim3(grid), dim3(WG), 0, st, arr, seg, sel, ghist, bins);
    else if (mode == 0)
        hipLaunchKernelGGL((k_hist<0, true, K, RAW, SEG>), dim3(grid), dim3(WG), 0, st, arr, seg, sel, ghist, bins);
    else if (!g)
        hipLaunchKernelGGL((k_hist<1, false, K, RAW, SEG>), dim3(grid), dim3(WG), 0, st, arr, seg, sel, ghist, bins);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}
hipError_t launch_hist(int mode, const u32* arr, SelState* sel, u32* ghist, int bins, unsigned grid, hipStream_t st) {
    return launch_hist_t<KeyU32, false, false>(mode, arr, Seg{}, sel, ghist, bins, grid, st);
}
hipError_t launch_hist_raw(int mode, int f32, const u32* keys, SelState* sel, u32* ghist, int bins, unsigned grid,
                           hipStream_t st) {
    return f32 ? launch_hist_t<KeyF32, true, false>(mode, keys, Seg{}, sel, ghist, bins, grid, st)
               : launch_hist_t<KeyU32, true, false>(mode, keys, Seg{}, sel, ghist, bins, grid, st);
}
hipError_t launch_hist_seg(int mode, const Seg& seg, SelState* sel, u32* ghist, int bins, unsigned grid,
                           hipStream_t st) {
    return launch_hist_t<KeyU32, false, true>(mode, nullptr, seg, sel, ghist, bins, grid, st);
}

// ---------------------------------------------------------------- select step
// Target: the first bin i whose cumulative count (within this pass) reaches T = m-k+1-below,
// i.e. the bin holding the key of ascending rank m-k.  One 256-thread workgroup: with <= 256
// bins each thread owns one bin and issues all of its HIST_REPS replica loads at once; larger
// (bucket) histograms have one replica and a few bins per thread.
constexpr int SEL_T = 256;

template <int MODE>
__global__ __launch_bounds__(SEL_T) void k_select(u32* ghist, int bins, SelState* sel, Ctrl* ctrl,
                                                 const u32* __restrict__ segcnt, u64 nseg,
                                                 const u64* __restrict__ part, u64 npart) {
    __shared__ u64 scratch[SEL_T / WAVE + 1];
    __shared__ u64 found[3];  // bin, count before it, count in it
    u64 seg_m = 0;  // with segcnt: the concatenation's size (every thread holds it)
    if (segcnt) {
        __shared__ u64 pl[4][SEL_T / WAVE];
        if (part) beta_part_total<SEL_T>(part, npart, ctrl, pl);
        // first pass over a segmented concatenation: its size from the segment counts first
        // (what a separate one-workgroup total would do, see k_seg_total)
        u64 s = 0;
#pragma unroll 8
        for (u64 i = threadIdx.x; i < nseg; i += SEL_T) s += segcnt[i];
        u64 tot;
        block_excl_scan_u64<SEL_T>(s, scratch, &tot);
        seg_m = tot;
        if (threadIdx.x == 0) {
            ctrl->concat_size = tot + ctrl->lonely_dropped;  // the dropped lonely delegates count too
            sel->m = tot;
            if (tot < sel->k) {
                ctrl_fail(ctrl, ERR_CONCAT_SHORT);
                sel->done = 1;
            }
        }
        __syncthreads();
        if (tot < sel->k) return;
    }
    if (sel->done) return;
    constexpr int MAXPER = MAX_LDS_BINS / SEL_T;  // 16
    const int per = (bins + SEL_T - 1) / SEL_T;
    const int b0 = min(bins, (int)threadIdx.x * per), b1 = min(bins, b0 + per);
    u32 cnt[MAXPER];
    u64 sum = 0;
    if (bins <= SEL_T) {  // 16 replicas, one bin per thread
        u32 c = 0;
        if (b0 < b1) {
            u32 r16[HIST_REPS];
#pragma unroll
            for (int r = 0; r < HIST_REPS; ++r) r16[r] = ghist[(size_t)r * bins + b0];
#pragma unroll
            for (int r = 0; r < HIST_REPS; ++r) c += r16[r];
        }
        cnt[0] = c;
        sum = c;
    } else if (per <= MAXPER) {  // one replica, up to 16 bins per thread
#pragma unroll
        for (int j = 0; j < MAXPER; ++j) {
            cnt[j] = (b0 + j < b1) ? ghist[b0 + j] : 0u;
            sum += cnt[j];
        }
    } else {  // 16-bit digits: one replica, 256 bins per thread, read again below
        for (int b = b0; b < b1; ++b) sum += ghist[b];
    }
    u64 tot;
    const u64 excl = block_excl_scan_u64<SEL_T>(sum, scratch, &tot);
    // sel->m as just written by thread 0 need not be visible yet to the others: use seg_m
    const u64 m = segcnt ? seg_m : sel->m, k = sel->k, below = sel->below;
    const u64 T = m - k + 1 - below;
    if (threadIdx.x == 0) found[0] = ~0ull;
    __syncthreads();
    if (excl < T && excl + sum >= T) {
        u64 run = excl;
        if (per <= MAXPER) {
#pragma unroll
            for (int j = 0; j < MAXPER; ++j) {
                if (b0 + j >= b1) break;
                const u64 c = cnt[j];
                if (run + c >= T) {
                    found[0] = (u64)(b0 + j);
                    found[1] = run;
                    found[2] = c;
                    break;
                }
                run += c;
            }
        } else {
            for (int b = b0; b < b1; ++b) {
                const u64 c = ghist[b];
                if (run + c >= T) {
                    found[0] = (u64)b;
                    found[1] = run;
                    found[2] = c;
                    break;
                }
                run += c;
            }
        }
    }
    __syncthreads();
    const int reps = hist_reps(bins);
    for (int r = 0; r < reps; ++r)  // ready for the next pass
        for (int b = b0; b < b1; ++b) ghist[(size_t)r * bins + b] = 0;
    if (threadIdx.x != 0) return;
    if (found[0] == ~0ull) {
        ctrl_fail(ctrl, ERR_CONCAT_SHORT);
        sel->done = 1;
        sel->result = 0;
        return;
    }
    const u32 i = (u32)found[0];
    sel->below = below + found[1];
    sel->kcount = found[2];
    sel->passes += 1;
    if (MODE == 0) {
        const int d = sel->digit;
        const u32 flag = sel->flag | (i << (d * (int)sel->nb));
        sel->flag = flag;
        sel->digit = d - 1;
        bool fin = d - 1 < (int)sel->stop_digit;
        // a dense threshold (more than dense_limit keys at or above it: float keys, whose top
        // byte is the sign and seven exponent bits, or skewed keys) takes one digit more
        if (fin && sel->dense_limit && d - 1 == (int)sel->stop_digit - 1 && m - (below + found[1]) > sel->dense_limit)
            fin = false;
        if (fin || (sel->stop_on_one && found[2] == 1)) {
            sel->done = 1;
            sel->result = flag;
        }
    } else {
        const u32 lo = sel->lo, hi = sel->hi;
        const u64 B = (u64)bins;
        const u32 nlo = bucket_lo(i, lo, hi, B), nhi = bucket_hi(i, lo, hi, B);
        sel->lo = nlo;
        sel->hi = nhi;
        if (sel->stop_on_one && found[2] == 1) {
            sel->need_find = 1;
            sel->done = 1;
            sel->result = nlo;  // replaced by the find pass
        } else if (nlo == nhi) {
            sel->done = 1;
            sel->result = nlo;
        } else if (sel->passes >= sel->max_passes) {
            sel->done = 1;
            sel->result = nlo;
        }
    }
}

hipError_t launch_select(int mode, u32* ghist, int bins, SelState* sel, Ctrl* ctrl, hipStream_t st,
                         const u32* segcnt, u64 nseg, const u64* part, u64 npart) {
    if (mode == 0)
        hipLaunchKernelGGL(k_select<0>, dim3(1), dim3(SEL_T), 0, st, ghist, bins, sel, ctrl, segcnt, nseg, part, npart);
    else
        hipLaunchKernelGGL(k_select<1>, dim3(1), dim3(SEL_T), 0, st, ghist, bins, sel, ctrl, segcnt, nseg, part, npart);
    return hipGetLastError();
}

// ---------------------------------------------------------------- bucket range set-up
// min/max over an array (second top-k), into sel->lo / sel->hi (initialised to ~0 / 0).
template <class K, bool RAW, bool SEG>
__global__ __launch_bounds__(WG) void k_minmax(const u32* __restrict__ arr, Seg seg, SelState* sel) {
    __shared__ u32 red[16];
    u32 mn = 0xFFFFFFFFu, mx = 0;
    auto f = [&](u32 raw) {
        const u32 v = RAW ? K::to_ord(raw) : raw;
        mn = min(mn, v);
        mx = max(mx, v);
    };
    if (SEG) for_seg(seg, f);
    else for_flat(arr, sel->m, f);
    mn = wave_min_u32(mn);
    mx = wave_max_u32(mx);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[w] = mn; red[8 + w] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < WG / WAVE; ++i) { mn = min(mn, red[i]); mx = max(mx, red[8 + i]); }
        mn = min(mn, red[0]);
        mx = max(mx, red[8]);
        atomicMin(&sel->lo, mn);
        atomicMax(&sel->hi, mx);
    }
}
hipError_t launch_minmax(const u32* arr, SelState* sel, Ctrl*, unsigned grid, hipStream_t st) {
    hipLaunchKernelGGL((k_minmax<KeyU32, false, false>), dim3(grid), dim3(WG), 0, st, arr, Seg{}, sel);
    return hipGetLastError();
}
hipError_t launch_minmax_raw(int f32, const u32* keys, SelState* sel, Ctrl*, unsigned grid, hipStream_t st) {
    if (f32) hipLaunchKernelGGL((k_minmax<KeyF32, true, false>), dim3(grid), dim3(WG), 0, st, keys, Seg{}, sel);
    else hipLaunchKernelGGL((k_minmax<KeyU32, true, false>), dim3(grid), dim3(WG), 0, st, keys, Seg{}, sel);
    return hipGetLastError();
}
hipError_t launch_minmax_seg(const Seg& seg, SelState* sel, unsigned grid, hipStream_t st) {
    hipLaunchKernelGGL((k_minmax<KeyU32, false, true>), dim3(grid), dim3(WG), 0, st, nullptr, seg, sel);
    return hipGetLastError();
}

// After min/max: a single-valued array is already selected (Max_Min + max==min,
// bucketselect_combined.cuh:1141-1153).
__global__ void k_bucket_begin(SelState* sel) {
    if (sel->done) return;
    if (sel->lo == sel->hi) {
        sel->done = 1;
        sel->result = sel->lo;
        sel->below = 0;
        sel->kcount = sel->m;
    }
}
hipError_t launch_bucket_begin(SelState* sel, Ctrl*, int, hipStream_t st) {
    hipLaunchKernelGGL(k_bucket_begin, dim3(1), dim3(1), 0, st, sel);
    return hipGetLastError();
}

// The k-bucket holds exactly one key: fetch it (GetKValue, bucketselect_combined.cuh:271-298;
// unique, so there is a single writer).
template <class K, bool RAW, bool SEG>
__global__ __launch_bounds__(WG) void k_find(const u32* __restrict__ arr, Seg seg, SelState* sel) {
    if (!sel->need_find) return;
    const u32 lo = sel->lo, hi = sel->hi;
    auto f = [&](u32 raw) {
        const u32 v = RAW ? K::to_ord(raw) : raw;
        if (v >= lo && v <= hi) sel->result = v;
    };
    if (SEG) for_seg(seg, f);
    else for_flat(arr, sel->m, f);
}
hipError_t launch_find(const u32* arr, int raw_f32, SelState* sel, unsigned grid, hipStream_t st) {
    if (raw_f32 == 1) hipLaunchKernelGGL((k_find<KeyF32, true, false>), dim3(grid), dim3(WG), 0, st, arr, Seg{}, sel);
    else hipLaunchKernelGGL((k_find<KeyU32, false, false>), dim3(grid), dim3(WG), 0, st, arr, Seg{}, sel);
    return hipGetLastError();
}
hipError_t launch_find_seg(const Seg& seg, SelState* sel, unsigned grid, hipStream_t st) {
    hipLaunchKernelGGL((k_find<KeyU32, false, true>), dim3(grid), dim3(WG), 0, st, nullptr, seg, sel);
    return hipGetLastError();
}

// ---------------------------------------------------------------- finalize
__global__ void k_finalize(SelState* sel, Ctrl* ctrl, int f32, u32* d_kth, u64 k) {
    const u32 K0 = sel->result;
    const u64 gt = sel->m - sel->below - sel->kcount;
    ctrl->kth = K0;
    ctrl->count_gt = gt;
    ctrl->ties = k - gt;
    if (d_kth) *d_kth = f32 ? KeyF32::from_ord(K0) : K0;
}
hipError_t launch_finalize(SelState* sel, Ctrl* ctrl, int f32, void* d_kth, u64 k, hipStream_t st) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1), 0, st, sel, ctrl, f32, (u32*)d_kth, k);
    return hipGetLastError();
}

// ordered keys -> user dtype
__global__ void k_to_user(const u32* ord, u64 m, int f32, u32* out) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride)
        out[i] = f32 ? KeyF32::from_ord(ord[i]) : ord[i];
}
hipError_t launch_to_user(const u32* ord, u64 m, int f32, void* out, hipStream_t st) {
    unsigned grid = (unsigned)std::max<u64>(1, std::min<u64>((m + WG - 1) / WG, 4096));
    hipLaunchKernelGGL(k_to_user, dim3(grid), dim3(WG), 0, st, ord, m, f32, (u32*)out);
    return hipGetLastError();
}

unsigned device_cus() {
    static std::atomic<unsigned> cache[64];  // 0 until set; concurrent callers store equal values
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    unsigned c = cache[dev].load(std::memory_order_relaxed);
    if (!c) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        c = (unsigned)v;
        cache[dev].store(c, std::memory_order_relaxed);
    }
    return c;
}

}  // namespace drtopk
