// select.hip -- the device-resident selection machinery of Dr.Top-k.
//
//  k_init        writes the control block and clears histograms / look-back state
//                (a kernel, not hipMemsetAsync, so the pipeline can be graph-captured).
//  k_hist        one histogram pass: radix digit of keys matching the current prefix
//                (CalculateOccurence_inplace[_firsttopk], radixselect.cuh:353-410) or
//                linear bucket of keys inside the current range (assignBuckets,
//                bucketselect_combined.cuh:167-197, integer map).
//  k_select      one workgroup: cumulative counts, the k-digit / k-bucket, state update
//                (CumulateCount_inplace radixselect.cuh:192-220 and the host loop
//                :1427-1445; FindKBucket bucketselect_combined.cuh:200-216).  No D2H copy.
//  k_range_default  FindTopKRanges + FillWithKElementsID (radixselect.cuh:413-427,
//                :603-618) as one ordered compaction: delegates above the threshold, then
//                the lowest-id delegates equal to it, exactly k ids in ascending order.
//  k_gather_default CreateConcatenateRange[_optimized] (radixselect.cuh:1546-1607): the
//                keys (>= threshold with the filter) of the selected subranges, with their
//                indices, in index order.
//  k_range_gather_beta FindTopKRanges_beta + CreateConcatenateRange_Beta_optimized
//                (radixselect.cuh:472-600, :1610-1639) fused: per subrange, all beta
//                delegates >= threshold -> its keys, else if the maximum >= threshold ->
//                its delegates ("lonely").
//  k_extract     findTopk + FillTopk (MultiGPU/radixselect.cuh:1355-1383): keys above the
//                k-th largest plus the lowest-index keys equal to it, in index order.
// Ordered compactions use the decoupled look-back of common.hpp over 4096-element tiles.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace drtopk {

constexpr int TILE_E = 16;              // elements per thread per tile
constexpr int TILE = WG * TILE_E;       // 4096 elements per tile
constexpr int MAX_LDS_BINS = 4096;

// ---------------------------------------------------------------- init
__global__ void k_init(Ctrl* ctrl, Ctrl init, u32* zero, u64 words) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *ctrl = init;
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

// ---------------------------------------------------------------- histogram pass
template <int MODE, bool GLOBALH, class K, bool RAW>
__global__ __launch_bounds__(WG) void k_hist(const u32* __restrict__ arr, SelState* sel, u32* ghist, int bins) {
    __shared__ u32 lh[GLOBALH ? 1 : MAX_LDS_BINS];
    if (sel->done) return;
    const u64 m = sel->m;
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
    auto count = [&](u32 raw) {
        const u32 v = RAW ? K::to_ord(raw) : raw;
        if (MODE == 0) {
            if ((v & mask) == flag) atomicAdd(&h[(v >> shright) & dmask], 1u);
        } else {
            if (v >= lo && v - lo <= W) atomicAdd(&h[bucket_of(v, lo, W, Bm1, inv)], 1u);
        }
    };
    const u64 nv = m >> 2;
    const u32x4* a4 = (const u32x4*)arr;
    const u64 stride = (u64)gridDim.x * WG;
    for (u64 i = (u64)blockIdx.x * WG + threadIdx.x; i < nv; i += stride) {
        u32x4 r = a4[i];
        count(r.x); count(r.y); count(r.z); count(r.w);
    }
    if (blockIdx.x == 0)
        for (u64 i = (nv << 2) + threadIdx.x; i < m; i += WG) count(arr[i]);
    if (!GLOBALH) {
        __syncthreads();
        for (int i = threadIdx.x; i < bins; i += WG) {
            u32 c = lh[i];
            if (c) atomicAdd(&ghist[i], c);
        }
    }
}

template <class K, bool RAW>
static hipError_t launch_hist_t(int mode, const u32* arr, SelState* sel, u32* ghist, int bins, unsigned grid,
                                hipStream_t st) {
    const bool g = bins > MAX_LDS_BINS;
    if (mode == 0 && !g) hipLaunchKernelGGL((k_hist<0, false, K, RAW>), dim3(grid), dim3(WG), 0, st, arr, sel, ghist, bins);
    else if (mode == 0) hipLaunchKernelGGL((k_hist<0, true, K, RAW>), dim3(grid), dim3(WG), 0, st, arr, sel, ghist, bins);
    else if (!g) hipLaunchKernelGGL((k_hist<1, false, K, RAW>), dim3(grid), dim3(WG), 0, st, arr, sel, ghist, bins);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}
hipError_t launch_hist(int mode, const u32* arr, SelState* sel, u32* ghist, int bins, unsigned grid, hipStream_t st) {
    return launch_hist_t<KeyU32, false>(mode, arr, sel, ghist, bins, grid, st);
}
hipError_t launch_hist_raw(int mode, int f32, const u32* keys, SelState* sel, u32* ghist, int bins, unsigned grid,
                           hipStream_t st) {
    return f32 ? launch_hist_t<KeyF32, true>(mode, keys, sel, ghist, bins, grid, st)
               : launch_hist_t<KeyU32, true>(mode, keys, sel, ghist, bins, grid, st);
}

// ---------------------------------------------------------------- select step
// Target: the first bin i whose cumulative count (within this pass) reaches T = m-k+1-below,
// i.e. the bin holding the key of ascending rank m-k.
constexpr int SEL_T = 1024;

template <int MODE>
__global__ __launch_bounds__(SEL_T) void k_select(u32* ghist, int bins, SelState* sel, Ctrl* ctrl) {
    __shared__ u64 scratch[SEL_T / WAVE + 1];
    __shared__ u64 found[3];  // bin, count before it, count in it
    if (sel->done) return;
    const int per = (bins + SEL_T - 1) / SEL_T;
    const int b0 = min(bins, (int)threadIdx.x * per), b1 = min(bins, b0 + per);
    u64 sum = 0;
    for (int b = b0; b < b1; ++b) sum += ghist[b];
    u64 tot;
    const u64 excl = block_excl_scan_u64<SEL_T>(sum, scratch, &tot);
    const u64 m = sel->m, k = sel->k, below = sel->below;
    const u64 T = m - k + 1 - below;
    if (threadIdx.x == 0) found[0] = ~0ull;
    __syncthreads();
    if (excl < T && excl + sum >= T) {
        u64 run = excl;
        for (int b = b0; b < b1; ++b) {
            u64 c = ghist[b];
            if (run + c >= T) { found[0] = (u64)b; found[1] = run; found[2] = c; break; }
            run += c;
        }
    }
    __syncthreads();
    for (int b = b0; b < b1; ++b) ghist[b] = 0;  // ready for the next pass
    if (threadIdx.x != 0) return;
    if (found[0] == ~0ull) {
        atomicOr(&ctrl->error, ERR_CONCAT_SHORT);
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
        if (d - 1 < (int)sel->stop_digit || (sel->stop_on_one && found[2] == 1)) {
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

hipError_t launch_select(int mode, u32* ghist, int bins, SelState* sel, Ctrl* ctrl, hipStream_t st) {
    if (mode == 0) hipLaunchKernelGGL(k_select<0>, dim3(1), dim3(SEL_T), 0, st, ghist, bins, sel, ctrl);
    else hipLaunchKernelGGL(k_select<1>, dim3(1), dim3(SEL_T), 0, st, ghist, bins, sel, ctrl);
    return hipGetLastError();
}

// ---------------------------------------------------------------- bucket range set-up
// min/max over an array (second top-k), into sel->lo / sel->hi (initialised to ~0 / 0).
template <class K, bool RAW>
__global__ __launch_bounds__(WG) void k_minmax(const u32* __restrict__ arr, SelState* sel) {
    __shared__ u32 red[16];
    const u64 m = sel->m;
    u32 mn = 0xFFFFFFFFu, mx = 0;
    const u64 stride = (u64)gridDim.x * WG;
    for (u64 i = (u64)blockIdx.x * WG + threadIdx.x; i < m; i += stride) {
        u32 v = RAW ? K::to_ord(arr[i]) : arr[i];
        mn = min(mn, v);
        mx = max(mx, v);
    }
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
    hipLaunchKernelGGL((k_minmax<KeyU32, false>), dim3(grid), dim3(WG), 0, st, arr, sel);
    return hipGetLastError();
}
hipError_t launch_minmax_raw(int f32, const u32* keys, SelState* sel, Ctrl*, unsigned grid, hipStream_t st) {
    if (f32) hipLaunchKernelGGL((k_minmax<KeyF32, true>), dim3(grid), dim3(WG), 0, st, keys, sel);
    else hipLaunchKernelGGL((k_minmax<KeyU32, true>), dim3(grid), dim3(WG), 0, st, keys, sel);
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
template <class K, bool RAW>
__global__ __launch_bounds__(WG) void k_find(const u32* __restrict__ arr, SelState* sel) {
    if (!sel->need_find) return;
    const u64 m = sel->m;
    const u32 lo = sel->lo, hi = sel->hi;
    const u64 stride = (u64)gridDim.x * WG;
    for (u64 i = (u64)blockIdx.x * WG + threadIdx.x; i < m; i += stride) {
        u32 v = RAW ? K::to_ord(arr[i]) : arr[i];
        if (v >= lo && v <= hi) sel->result = v;
    }
}
hipError_t launch_find(const u32* arr, int raw_f32, SelState* sel, unsigned grid, hipStream_t st) {
    if (raw_f32 == 1) hipLaunchKernelGGL((k_find<KeyF32, true>), dim3(grid), dim3(WG), 0, st, arr, sel);
    else hipLaunchKernelGGL((k_find<KeyU32, false>), dim3(grid), dim3(WG), 0, st, arr, sel);
    return hipGetLastError();
}

// ---------------------------------------------------------------- range selection (default path)
__global__ __launch_bounds__(WG) void k_range_default(const u32* __restrict__ dvals, u64 S, u64 k, int fullcmp,
                                                      SelState* sel, Ctrl* ctrl, u32* __restrict__ sel_ids,
                                                      u64* status, u32* ticket) {
    __shared__ u64 scratch[WG / WAVE + 1];
    __shared__ u64 tile_excl;
    __shared__ u32 tslot;
    const u32 thr = sel->result;
    const int shleft = fullcmp ? 0 : (sel->digit + 1) * (int)sel->nb;
    const u32 mask = shleft >= 32 ? 0u : ~((1u << shleft) - 1u);
    const u64 total_gt = sel->m - sel->below - sel->kcount;
    const u64 r1 = k - total_gt;
    const u64 ntiles = (S + TILE - 1) / TILE;
    for (;;) {
        const u32 t = next_tile(ticket, &tslot);
        if (t >= ntiles) break;
        const u64 e0 = (u64)t * TILE + (u64)threadIdx.x * TILE_E;
        u32 v[TILE_E];
        u32 gt = 0, eq = 0;
#pragma unroll
        for (int j = 0; j < TILE_E; ++j) {
            v[j] = e0 + j < S ? (dvals[e0 + j] & mask) : 0u;
            bool in = e0 + j < S;
            gt += in && v[j] > thr;
            eq += in && v[j] == thr;
        }
        u64 tot;
        const u64 ex = block_excl_scan_u64<WG>(pack2(gt, eq), scratch, &tot);
        if (threadIdx.x == 0) tile_excl = lookback(status, t, tot, &ctrl->error);
        __syncthreads();
        const u64 te = tile_excl;
        u64 g = unpack_hi(te) + unpack_hi(ex), q = unpack_lo(te) + unpack_lo(ex);
#pragma unroll
        for (int j = 0; j < TILE_E; ++j) {
            if (e0 + j >= S) break;
            u64 pos = ~0ull;
            if (v[j] > thr) {
                pos = g + (q < r1 ? q : r1);
                ++g;
            } else if (v[j] == thr) {
                if (q < r1) pos = g + q;
                ++q;
            }
            if (pos != ~0ull) {
                if (pos < k) sel_ids[pos] = (u32)(e0 + j);
                else atomicOr(&ctrl->error, ERR_CAPACITY);
            }
        }
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctrl->selected = k;
}

hipError_t launch_range_default(const RangeArgs& a, hipStream_t st) {
    const u64 ntiles = (a.S + TILE - 1) / TILE;
    unsigned grid = (unsigned)std::min<u64>(ntiles, 2048);
    hipLaunchKernelGGL(k_range_default, dim3(grid), dim3(WG), 0, st, a.dvals, a.S, a.k, a.mode, a.sel, a.ctrl,
                       a.sel_ids, a.status, a.ticket);
    return hipGetLastError();
}

// ---------------------------------------------------------------- gather (default path)
template <class K, bool HIST2>
__global__ __launch_bounds__(WG) void k_gather_default(const u32* __restrict__ keys, u64 n, int alpha, int filter,
                                                       int vec_ok, const u32* __restrict__ sel_ids, u64 nsel,
                                                       Ctrl* ctrl, u32* __restrict__ ckeys, u32* __restrict__ cidx,
                                                       u64 cap, u64* status, u32* ticket, u32* ghist2, int hshift2,
                                                       int bins2) {
    __shared__ u64 scratch[WG / WAVE + 1];
    __shared__ u64 tile_excl;
    __shared__ u32 tslot;
    __shared__ u32 lh[HIST2 ? 256 : 1];
    if (HIST2) {
        for (int i = threadIdx.x; i < bins2; i += WG) lh[i] = 0;
    }
    const u32 thr = ctrl->first.result;
    const u64 L = 1ull << alpha, Lm = L - 1;
    const u64 V = nsel << alpha;
    const u64 ntiles = (V + TILE - 1) / TILE;
    for (;;) {
        const u32 t = next_tile(ticket, &tslot);
        if (t >= ntiles) break;
        const u64 v0 = (u64)t * TILE + (u64)threadIdx.x * TILE_E;
        u32 kv[TILE_E];
        u32 gi[TILE_E];
        u32 cnt = 0;
#pragma unroll
        for (int q = 0; q < TILE_E / 4; ++q) {
            const u64 vp = v0 + 4 * q;
            u32 r[4] = {0, 0, 0, 0};
            u64 g0 = 0;
            bool any = vp < V;
            if (any) {
                if (L >= 4) {
                    const u64 sid = sel_ids[vp >> alpha];
                    g0 = (sid << alpha) + (vp & Lm);
                    if (vec_ok && g0 + 3 < n) {
                        u32x4 x = *(const u32x4*)(keys + g0);
                        r[0] = x.x; r[1] = x.y; r[2] = x.z; r[3] = x.w;
                    } else {
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            if (g0 + c < n) r[c] = keys[g0 + c];
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const u64 vpc = vp + c;
                u64 g = g0 + c;
                bool ok = vpc < V;
                if (ok && L < 4) {
                    const u64 sid = sel_ids[vpc >> alpha];
                    g = (sid << alpha) + (vpc & Lm);
                    if (g < n) r[c] = keys[g];
                }
                ok = ok && g < n;
                const u32 o = K::to_ord(r[c]);
                const bool take = ok && (!filter || o >= thr);
                kv[4 * q + c] = take ? o : 0u;
                gi[4 * q + c] = take ? (u32)g : 0xFFFFFFFFu;
                cnt += take;
            }
        }
        u64 tot;
        const u64 ex = block_excl_scan_u64<WG>((u64)cnt, scratch, &tot);
        if (threadIdx.x == 0) tile_excl = lookback(status, t, tot, &ctrl->error);
        __syncthreads();
        u64 pos = tile_excl + ex;
#pragma unroll
        for (int j = 0; j < TILE_E; ++j) {
            if (gi[j] != 0xFFFFFFFFu) {
                if (pos < cap) {
                    ckeys[pos] = kv[j];
                    if (cidx) cidx[pos] = gi[j];
                } else {
                    atomicOr(&ctrl->error, ERR_CAPACITY);
                }
                if (HIST2) atomicAdd(&lh[kv[j] >> hshift2], 1u);
                ++pos;
            }
        }
        if (t == ntiles - 1 && threadIdx.x == 0) {
            u64 C = tile_excl + tot;
            ctrl->concat_size = C;
            if (C > cap) {
                atomicOr(&ctrl->error, ERR_CAPACITY);
                C = cap;
            }
            ctrl->second.m = C;
            if (C < ctrl->second.k) atomicOr(&ctrl->error, ERR_CONCAT_SHORT);
        }
        __syncthreads();
    }
    if (HIST2) {
        __syncthreads();
        for (int i = threadIdx.x; i < bins2; i += WG) {
            u32 c = lh[i];
            if (c) atomicAdd(&ghist2[i], c);
        }
    }
}

hipError_t launch_gather_default(const GatherArgs& a, hipStream_t st) {
    const u64 V = a.nsel << a.alpha;
    const u64 ntiles = (V + TILE - 1) / TILE;
    unsigned grid = (unsigned)std::min<u64>(ntiles, 4096);
    const int vec_ok = ((uintptr_t)a.keys & 15) == 0;
#define GD_ARGS a.keys, a.n, a.alpha, a.filter, vec_ok, a.sel_ids, a.nsel, a.ctrl, a.ckeys, a.cidx, a.cap, a.status, \
                a.ticket, a.ghist2, a.hshift2, a.bins2
    if (a.f32) {
        if (a.ghist2) hipLaunchKernelGGL((k_gather_default<KeyF32, true>), dim3(grid), dim3(WG), 0, st, GD_ARGS);
        else hipLaunchKernelGGL((k_gather_default<KeyF32, false>), dim3(grid), dim3(WG), 0, st, GD_ARGS);
    } else {
        if (a.ghist2) hipLaunchKernelGGL((k_gather_default<KeyU32, true>), dim3(grid), dim3(WG), 0, st, GD_ARGS);
        else hipLaunchKernelGGL((k_gather_default<KeyU32, false>), dim3(grid), dim3(WG), 0, st, GD_ARGS);
    }
#undef GD_ARGS
    return hipGetLastError();
}

// ---------------------------------------------------------------- range select + gather (beta path)
constexpr int BETA_SPT = 4;                 // subranges per thread
constexpr int BETA_TILE = WG * BETA_SPT;    // 1024 subranges per tile

template <class K, int BETA, bool POS, bool HIST2>
__global__ __launch_bounds__(WG) void k_range_gather_beta(const u32* __restrict__ keys, u64 n, int alpha, int filter,
                                                          const u32* __restrict__ dvals, const void* __restrict__ doff,
                                                          Ctrl* ctrl, u32* __restrict__ ckeys, u32* __restrict__ cidx,
                                                          u64 cap, u64* status, u32* ticket, u32* ghist2, int hshift2,
                                                          int bins2) {
    __shared__ u64 scratch[WG / WAVE + 1];
    __shared__ u64 tile_excl;
    __shared__ u32 tslot;
    __shared__ u32 lh[HIST2 ? 256 : 1];
    __shared__ u64 red[3][WG / WAVE];
    if (HIST2) {
        for (int i = threadIdx.x; i < bins2; i += WG) lh[i] = 0;
    }
    const u32 thr = ctrl->first.result;
    const u64 L = 1ull << alpha;
    const u64 S = (n + L - 1) >> alpha;
    const u64 ntiles = (S + BETA_TILE - 1) / BETA_TILE;
    u64 nsel = 0, nlon = 0, nlel = 0;
    for (;;) {
        const u32 t = next_tile(ticket, &tslot);
        if (t >= ntiles) break;
        const u64 s0 = (u64)t * BETA_TILE + (u64)threadIdx.x * BETA_SPT;
        // kind: 0 none, 1 selected (whole subrange, filtered), 2 lonely
        int kind[BETA_SPT];
        u32 cnt = 0;
#pragma unroll
        for (int j = 0; j < BETA_SPT; ++j) {
            const u64 s = s0 + j;
            kind[j] = 0;
            if (s >= S) continue;
            const u64 base = s << alpha;
            const u64 len = min(L, n - base);
            const int m = (int)std::min<u64>(BETA, len);
            u32 d[BETA];
#pragma unroll
            for (int b = 0; b < BETA; ++b) d[b] = dvals[s * BETA + b];
            if (d[0] < thr) continue;
            bool all = true;
#pragma unroll
            for (int b = 0; b < BETA; ++b) all = all && (b >= m || d[b] >= thr);
            if (all) {
                kind[j] = 1;
                ++nsel;
                if (!filter) cnt += (u32)len;
                else
                    for (u64 i = base; i < base + len; ++i) cnt += K::to_ord(keys[i]) >= thr;
            } else {
                kind[j] = 2;
                ++nlon;
                nlel += (u64)m;
                cnt += (u32)m;
            }
        }
        u64 tot;
        const u64 ex = block_excl_scan_u64<WG>((u64)cnt, scratch, &tot);
        if (threadIdx.x == 0) tile_excl = lookback(status, t, tot, &ctrl->error);
        __syncthreads();
        u64 pos = tile_excl + ex;
        auto put = [&](u32 o, u32 gidx) {
            if (pos < cap) {
                ckeys[pos] = o;
                if (POS) cidx[pos] = gidx;
            } else {
                atomicOr(&ctrl->error, ERR_CAPACITY);
            }
            if (HIST2) atomicAdd(&lh[o >> hshift2], 1u);
            ++pos;
        };
#pragma unroll
        for (int j = 0; j < BETA_SPT; ++j) {
            const u64 s = s0 + j;
            if (kind[j] == 0) continue;
            const u64 base = s << alpha;
            const u64 len = min(L, n - base);
            if (kind[j] == 1) {
                for (u64 i = base; i < base + len; ++i) {
                    const u32 o = K::to_ord(keys[i]);
                    if (!filter || o >= thr) put(o, (u32)i);
                }
            } else {
                const int m = (int)std::min<u64>(BETA, len);
                u32 d[BETA], p[BETA];
                u32 packed = 0;
                if (POS) {
                    if (BETA == 2) packed = ((const uint8_t*)doff)[s];
                    else if (BETA == 3) packed = ((const uint16_t*)doff)[s];
                    else packed = ((const u32*)doff)[s];
                }
#pragma unroll
                for (int b = 0; b < BETA; ++b) {
                    d[b] = dvals[s * BETA + b];
                    const bool ge = b < m && d[b] >= thr;  // only these can reach the output
                    p[b] = (POS && ge && b < BETA - 1) ? (u32)(base + ((packed >> (8 * b)) & 0xFFu)) : 0xFFFFFFFFu;
                }
                // delegates >= threshold in index order, then the rest (never selected later)
#pragma unroll
                for (int x = 0; x < BETA; ++x)
#pragma unroll
                    for (int y = 0; y + 1 < BETA - x; ++y)
                        if (p[y + 1] < p[y]) {
                            u32 tp = p[y]; p[y] = p[y + 1]; p[y + 1] = tp;
                            u32 td = d[y]; d[y] = d[y + 1]; d[y + 1] = td;
                        }
                for (int b = 0; b < BETA; ++b)
                    if (b < m) put(d[b], p[b]);
            }
        }
        if (t == ntiles - 1 && threadIdx.x == 0) {
            u64 C = tile_excl + tot;
            ctrl->concat_size = C;
            if (C > cap) {
                atomicOr(&ctrl->error, ERR_CAPACITY);
                C = cap;
            }
            ctrl->second.m = C;
            if (C < ctrl->second.k) atomicOr(&ctrl->error, ERR_CONCAT_SHORT);
        }
        __syncthreads();
    }
    // counters
    {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        for (int o = 32; o > 0; o >>= 1) {
            nsel += shfl_xor_u64(nsel, o);
            nlon += shfl_xor_u64(nlon, o);
            nlel += shfl_xor_u64(nlel, o);
        }
        if (lane == 0) { red[0][w] = nsel; red[1][w] = nlon; red[2][w] = nlel; }
        __syncthreads();
        if (threadIdx.x == 0) {
            u64 a = 0, b = 0, c = 0;
            for (int i = 0; i < WG / WAVE; ++i) { a += red[0][i]; b += red[1][i]; c += red[2][i]; }
            if (a) atomicAdd((unsigned long long*)&ctrl->selected, (unsigned long long)a);
            if (b) atomicAdd((unsigned long long*)&ctrl->lonely_sub, (unsigned long long)b);
            if (c) atomicAdd((unsigned long long*)&ctrl->lonely_el, (unsigned long long)c);
        }
    }
    if (HIST2) {
        __syncthreads();
        for (int i = threadIdx.x; i < bins2; i += WG) {
            u32 c = lh[i];
            if (c) atomicAdd(&ghist2[i], c);
        }
    }
}

template <class K, int BETA>
static hipError_t launch_rgb_t(const BetaGatherArgs& a, hipStream_t st) {
    const u64 L = 1ull << a.alpha;
    const u64 S = (a.n + L - 1) >> a.alpha;
    const u64 ntiles = (S + BETA_TILE - 1) / BETA_TILE;
    unsigned grid = (unsigned)std::min<u64>(ntiles, 4096);
#define RGB_ARGS a.keys, a.n, a.alpha, a.filter, a.dvals, a.doff, a.ctrl, a.ckeys, a.cidx, a.cap, a.status, a.ticket, \
                 a.ghist2, a.hshift2, a.bins2
    const bool pos = a.cidx != nullptr, h = a.ghist2 != nullptr;
    if (pos && h) hipLaunchKernelGGL((k_range_gather_beta<K, BETA, true, true>), dim3(grid), dim3(WG), 0, st, RGB_ARGS);
    else if (pos) hipLaunchKernelGGL((k_range_gather_beta<K, BETA, true, false>), dim3(grid), dim3(WG), 0, st, RGB_ARGS);
    else if (h) hipLaunchKernelGGL((k_range_gather_beta<K, BETA, false, true>), dim3(grid), dim3(WG), 0, st, RGB_ARGS);
    else hipLaunchKernelGGL((k_range_gather_beta<K, BETA, false, false>), dim3(grid), dim3(WG), 0, st, RGB_ARGS);
#undef RGB_ARGS
    return hipGetLastError();
}
template <class K>
static hipError_t launch_rgb_k(const BetaGatherArgs& a, hipStream_t st) {
    switch (a.beta) {
        case 2: return launch_rgb_t<K, 2>(a, st);
        case 3: return launch_rgb_t<K, 3>(a, st);
        case 4: return launch_rgb_t<K, 4>(a, st);
    }
    return hipErrorInvalidValue;
}
hipError_t launch_range_gather_beta(const BetaGatherArgs& a, hipStream_t st) {
    return a.f32 ? launch_rgb_k<KeyF32>(a, st) : launch_rgb_k<KeyU32>(a, st);
}

// ---------------------------------------------------------------- extraction
template <class K, bool DIRECT>
__global__ __launch_bounds__(WG) void k_extract(const u32* __restrict__ ckeys, const u32* __restrict__ cidx, u64 k,
                                                SelState* sel, Ctrl* ctrl, u32* __restrict__ out_keys,
                                                i64* __restrict__ out_idx, u64* status, u32* ticket) {
    __shared__ u64 scratch[WG / WAVE + 1];
    __shared__ u64 tile_excl;
    __shared__ u32 tslot;
    const u64 C = sel->m;
    const u32 K0 = sel->result;
    const u64 total_gt = sel->m - sel->below - sel->kcount;
    const u64 r = k - total_gt;
    const u64 ntiles = (C + TILE - 1) / TILE;
    for (;;) {
        const u32 t = next_tile(ticket, &tslot);
        if (t >= ntiles) break;
        const u64 e0 = (u64)t * TILE + (u64)threadIdx.x * TILE_E;
        u32 v[TILE_E];
        u32 gt = 0, eq = 0;
#pragma unroll
        for (int j = 0; j < TILE_E; ++j) {
            const bool in = e0 + j < C;
            v[j] = in ? (DIRECT ? K::to_ord(ckeys[e0 + j]) : ckeys[e0 + j]) : 0u;
            gt += in && v[j] > K0;
            eq += in && v[j] == K0;
        }
        u64 tot;
        const u64 ex = block_excl_scan_u64<WG>(pack2(gt, eq), scratch, &tot);
        if (threadIdx.x == 0) tile_excl = lookback(status, t, tot, &ctrl->error);
        __syncthreads();
        const u64 te = tile_excl;
        u64 g = unpack_hi(te) + unpack_hi(ex), q = unpack_lo(te) + unpack_lo(ex);
#pragma unroll
        for (int j = 0; j < TILE_E; ++j) {
            if (e0 + j >= C) break;
            u64 pos = ~0ull;
            if (v[j] > K0) {
                pos = g + (q < r ? q : r);
                ++g;
            } else if (v[j] == K0) {
                if (q < r) pos = g + q;
                ++q;
            }
            if (pos != ~0ull) {
                if (pos < k) {
                    if (out_keys) out_keys[pos] = K::from_ord(v[j]);
                    if (out_idx) out_idx[pos] = DIRECT ? (i64)(e0 + j) : (i64)cidx[e0 + j];
                } else {
                    atomicOr(&ctrl->error, ERR_CAPACITY);
                }
            }
        }
        __syncthreads();
    }
}

hipError_t launch_extract(const ExtractArgs& a, hipStream_t st) {
    const u64 ntiles = (a.cap + TILE - 1) / TILE;
    unsigned grid = (unsigned)std::max<u64>(1, std::min<u64>(ntiles, 4096));
#define EX_ARGS a.ckeys, a.cidx, a.k, a.sel, a.ctrl, (u32*)a.out_keys, a.out_idx, a.status, a.ticket
    if (a.f32) {
        if (a.direct) hipLaunchKernelGGL((k_extract<KeyF32, true>), dim3(grid), dim3(WG), 0, st, EX_ARGS);
        else hipLaunchKernelGGL((k_extract<KeyF32, false>), dim3(grid), dim3(WG), 0, st, EX_ARGS);
    } else {
        if (a.direct) hipLaunchKernelGGL((k_extract<KeyU32, true>), dim3(grid), dim3(WG), 0, st, EX_ARGS);
        else hipLaunchKernelGGL((k_extract<KeyU32, false>), dim3(grid), dim3(WG), 0, st, EX_ARGS);
    }
#undef EX_ARGS
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

}  // namespace drtopk
