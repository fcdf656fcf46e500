// compact.hip -- range selection, concatenation and top-k extraction of Dr.Top-k.
//
// Every ordered compaction here runs as
//     count pass (workgroups independent)  ->  one-workgroup scan  ->  write pass
// with no workgroup ever waiting on another.  (A single-pass decoupled look-back was
// measured first: with ~1500 tiles in flight and ~2 us per cross-XCD status round trip,
// each tile spent tens of microseconds walking back; the passes below do not wait.)
//
// The concatenation is kept as a segmented array: tile t of the range/gather pass writes
// its survivors, in order, at [t*segcap, t*segcap + cnt[t]).  The second top-k reads the
// segments directly, and the extraction's own count/scan/write orders them globally, so
// the survivors are written once and never copied.
//
//  k_range_count / k_range_write  FindTopKRanges + FillWithKElementsID
//        (baseline+filter+beta+shuffle/radixselect.cuh:413-427, :603-618): delegates whose
//        prefix exceeds the first top-k flag, then the lowest-id delegates equal to it, up
//        to k ids in ascending order (the reference's fill is a racy first-come append).
//  k_gather  CreateConcatenateRange[_optimized] (radixselect.cuh:1546-1607): keys of the
//        selected subranges (>= threshold with the filter) with their indices.
//  k_beta_gather  FindTopKRanges_beta + CreateConcatenateRange_Beta_optimized
//        (radixselect.cuh:472-600, :1610-1639) fused: a subrange whose beta delegates are
//        all >= threshold contributes its keys, one whose maximum alone reaches it
//        contributes its delegates ("lonely").
//  k_extract_count / k_extract_write  findTopk + FillTopk (MultiGPU/radixselect.cuh:
//        1355-1383): keys above the k-th largest plus the lowest-index keys equal to it.
#include <algorithm>
#include <atomic>

#include "common.hpp"
#include "kernels.hpp"

namespace drtopk {

constexpr int TE = 16;             // consecutive elements per thread
constexpr u64 TT = WG * TE;        // 4096 elements per tile / chunk
constexpr int SCAN_T = 1024;
constexpr u64 CGRID = 8192;        // persistent grid bound of the count/write passes

// ---------------------------------------------------------------- scan
__global__ __launch_bounds__(SCAN_T) void k_scan(const u64* __restrict__ in, u64 nseg, u64* __restrict__ out,
                                                 u64* total) {
    __shared__ u64 scratch[SCAN_T / WAVE + 1];
    const u64 per = (nseg + SCAN_T - 1) / SCAN_T;
    const u64 b0 = std::min<u64>(nseg, threadIdx.x * per), b1 = std::min<u64>(nseg, b0 + per);
    u64 sum = 0;
    for (u64 i = b0; i < b1; ++i) sum += in[i];
    u64 tot;
    u64 run = block_excl_scan_u64<SCAN_T>(sum, scratch, &tot);
    for (u64 i = b0; i < b1; ++i) {
        const u64 v = in[i];
        out[i] = run;
        run += v;
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

hipError_t launch_scan(const u64* in, u64 nseg, u64* out_excl, u64* total, hipStream_t st) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(SCAN_T), 0, st, in, nseg, out_excl, total);
    return hipGetLastError();
}

// Load the TE consecutive elements of this thread starting at p (16-byte aligned), zero past m.
template <class K, bool RAW>
__device__ __forceinline__ void load_te(const u32* p, u64 e0, u64 m, u32* v) {
    if (e0 + TE <= m) {
        const u32x4* q = (const u32x4*)p;
#pragma unroll
        for (int j = 0; j < TE / 4; ++j) {
            const u32x4 r = q[j];
            v[4 * j] = RAW ? K::to_ord(r.x) : r.x;
            v[4 * j + 1] = RAW ? K::to_ord(r.y) : r.y;
            v[4 * j + 2] = RAW ? K::to_ord(r.z) : r.z;
            v[4 * j + 3] = RAW ? K::to_ord(r.w) : r.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < TE; ++j) v[j] = e0 + j < m ? (RAW ? K::to_ord(p[j]) : p[j]) : 0u;
    }
}

// Counts of keys > K and == K of thread-owned runs of TE keys; `whole` (uniform) skips the
// per-key bounds checks of a chunk that lies inside the segment.
template <int N>
__device__ __forceinline__ void count_gt_eq(const u32 (&v)[N], u32 K0, bool whole, u32 nvalid, u32& gt, u32& eq) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const bool in = whole || (u32)j < nvalid;
        gt += (in && v[j] > K0) ? 1u : 0u;
        eq += (in && v[j] == K0) ? 1u : 0u;
    }
}

// ---------------------------------------------------------------- LDS staging
// Survivors of a 4096-element chunk go to LDS at their rank, then the workgroup copies the
// chunk's run out with coalesced stores (a thread's own survivors sit at scattered global
// positions, so storing them directly costs one cache line per lane per store).
constexpr u32 STAGE = 4096;
constexpr u32 GSTAGE = 4096;  // survivors of a default-path gather chunk staged in LDS

template <bool IDX, bool HIST2>
__device__ __forceinline__ void flush_stage(const u32* sk, const u32* si, u32 cnt, u32* __restrict__ gk,
                                            u32* __restrict__ gi, u32* lh, int hshift2) {
    for (u32 b = 0; b < cnt; b += WG) {  // uniform bound: every lane reaches wave_hist_add
        const u32 i = b + threadIdx.x;
        const bool act = i < cnt;
        u32 v = 0;
        if (act) {
            v = sk[i];
            gk[i] = v;
            if (IDX) gi[i] = si[i];
        }
        if (HIST2) wave_hist_add(lh, act, v >> hshift2);
    }
}

// ---------------------------------------------------------------- range selection (default path)
struct RangeRule {
    u32 thr, mask;
    __device__ __forceinline__ RangeRule(const SelState* sel, int fullcmp) {
        thr = sel->result;
        const int shleft = fullcmp ? 0 : (sel->digit + 1) * (int)sel->nb;
        mask = shleft >= 32 ? 0u : ~((1u << shleft) - 1u);
    }
};

__global__ __launch_bounds__(WG) void k_range_count(const u32* __restrict__ dvals, u64 S, int fullcmp, SelState* sel,
                                                    u64* __restrict__ cnt) {
    __shared__ u32 scratch[WG / WAVE];
    const RangeRule rr(sel, fullcmp);
    const u64 ntiles = range_tiles(S);
    for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const u64 e0 = t * TT + (u64)threadIdx.x * TE;
        const bool whole = (t + 1) * TT <= S;  // uniform
        const u32 nvalid = e0 < S ? (u32)std::min<u64>(S - e0, TE) : 0u;
        u32 v[TE];
        load_te<KeyU32, false>(dvals + e0, e0, S, v);
#pragma unroll
        for (int j = 0; j < TE; ++j) v[j] &= rr.mask;
        u32 gt = 0, eq = 0;
        count_gt_eq(v, rr.thr, whole, nvalid, gt, eq);
        u32 tot;  // packed (gt << 16) | eq: at most TT = 4096 each
        block_excl_scan_u32<WG>((gt << 16) | eq, scratch, &tot);
        if (threadIdx.x == 0) cnt[t] = pack2(tot >> 16, tot & 0xFFFFu);
    }
}

// Positions inside a tile are 32-bit and relative to its first slot, as in k_extract_write.
// off: the tiles' exclusive prefixes of packed (gt, eq) counts, or null: each workgroup then sums
// the counts of the tiles before its own.
__global__ __launch_bounds__(WG) void k_range_write(const u32* __restrict__ dvals, u64 S, u64 k, int fullcmp,
                                                    SelState* sel, const u64* __restrict__ off,
                                                    const u64* __restrict__ cnt, u32* __restrict__ ids, Ctrl* ctrl) {
    __shared__ u32 scratch[WG / WAVE];
    __shared__ u64 red[WG / WAVE];
    __shared__ u32 sk[STAGE];
    const RangeRule rr(sel, fullcmp);
    const u64 total_gt = sel->m - sel->below - sel->kcount;
    const u64 r1 = k - total_gt;  // delegates equal to the flag that are taken, lowest ids first
    const u64 ntiles = range_tiles(S);
    for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const u64 e0 = t * TT + (u64)threadIdx.x * TE;
        const bool whole = (t + 1) * TT <= S;  // uniform
        const u32 nvalid = e0 < S ? (u32)std::min<u64>(S - e0, TE) : 0u;
        u32 v[TE];
        load_te<KeyU32, false>(dvals + e0, e0, S, v);
        u64 base;
        if (off) {
            base = off[t];
        } else {
            u64 part = 0;  // packed fields add without carries: each total is below 2^31
            for (u64 i = threadIdx.x; i < t; i += WG) part += cnt[i];
            part = wave_sum_u64(part);
            if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = part;
            __syncthreads();
            base = 0;
#pragma unroll
            for (int w = 0; w < WG / WAVE; ++w) base += red[w];
        }
#pragma unroll
        for (int j = 0; j < TE; ++j) v[j] &= rr.mask;
        u32 gt = 0, eq = 0;
        count_gt_eq(v, rr.thr, whole, nvalid, gt, eq);
        u32 tot;
        const u32 ex = block_excl_scan_u32<WG>((gt << 16) | eq, scratch, &tot);
        const u64 run_q = unpack_lo(base);
        const u32 rem = (u32)std::min<u64>(r1 - std::min<u64>(run_q, r1), TT);
        // the tile's ids occupy [P0, P0 + c): every id before position P0 comes from earlier tiles
        const u64 P0 = unpack_hi(base) + std::min<u64>(run_q, r1);
        const u32 c = (tot >> 16) + std::min<u32>(tot & 0xFFFFu, rem);
        u32 lg = ex >> 16, lq = ex & 0xFFFFu;
#pragma unroll
        for (int j = 0; j < TE; ++j) {
            const bool in = whole || (u32)j < nvalid;
            const bool isg = in && v[j] > rr.thr, ise = in && v[j] == rr.thr;
            const u32 sl = isg ? lg + min(lq, rem) : lg + lq;
            if (isg || (ise && lq < rem)) sk[sl] = (u32)(e0 + j);
            lg += isg ? 1u : 0u;
            lq += ise ? 1u : 0u;
        }
        __syncthreads();
        if (P0 + c > k) {
            if (threadIdx.x == 0) ctrl_fail(ctrl, ERR_CAPACITY);
        } else {
            u32* __restrict__ o = ids + P0;
            for (u32 i = threadIdx.x; i < c; i += WG) o[i] = sk[i];
        }
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctrl->selected = k;
}

hipError_t launch_range(const RangeArgs& a, hipStream_t st) {
    const u64 nt = range_tiles(a.S);
    const unsigned grid = (unsigned)std::min<u64>(nt, CGRID);
    hipLaunchKernelGGL(k_range_count, dim3(grid), dim3(WG), 0, st, a.dvals, a.S, a.fullcmp, a.sel, a.cnt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // up to 2048 tiles every workgroup sums the counts before its tile itself
    const bool fused = nt <= 2048;
    if (!fused) {
        e = launch_scan(a.cnt, nt, a.off, nullptr, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_range_write, dim3(grid), dim3(WG), 0, st, a.dvals, a.S, a.k, a.fullcmp, a.sel,
                       fused ? nullptr : a.off, a.cnt, a.sel_ids, a.ctrl);
    return hipGetLastError();
}

// ---------------------------------------------------------------- histogram fusion helpers
template <bool HIST2>
__device__ __forceinline__ void h2_init(u32* lh, int bins2) {
    if (HIST2)
        for (int i = threadIdx.x; i < bins2; i += WG) lh[i] = 0;
    __syncthreads();
}
template <bool HIST2>
__device__ __forceinline__ void h2_flush(u32* lh, u32* ghist2, int bins2) {
    if (HIST2) {
        __syncthreads();
        u32* g = ghist2 + (size_t)(blockIdx.x % hist_reps(bins2)) * bins2;
        for (int i = threadIdx.x; i < bins2; i += WG) {
            const u32 c = lh[i];
            if (c) atomicAdd(&g[i], c);
        }
    }
}

// Direct (unstaged) stores of the concatenation count their top-digit bin one key at a time.
// Every entry is >= the threshold, so skewed keys put most of them in the threshold's own bin:
// those are counted in a register and added once per wave at the end, the rest with LDS atomics.
struct H2Direct {
    u32 tbin, own = 0;
    __device__ __forceinline__ void add(u32* lh, u32 b) {
        if (b == tbin) ++own;
        else atomicAdd(&lh[b], 1u);
    }
    // every lane of the workgroup, before h2_flush
    __device__ __forceinline__ void flush(u32* lh) {
        const u32 w = (u32)wave_sum_u64(own);
        if ((threadIdx.x & (WAVE - 1)) == 0 && w) atomicAdd(&lh[tbin], w);
    }
};

// Chunks per segment: enough segments to fill the chip (~2048), at most 16 chunks each.
constexpr u64 CPS_MAX = 16;
u64 chunks_per_segment(u64 nchunks) { return std::max<u64>(1, std::min<u64>(CPS_MAX, nchunks / 2048)); }

// ---------------------------------------------------------------- default-path concatenation
// alpha == 1: one subrange id per key pair; the plain loop below.
template <class K, bool HIST2, bool IDX>
__global__ __launch_bounds__(WG) void k_gather_small(const u32* __restrict__ keys, u64 n, int alpha, int filter, int vec_ok,
                                               const u32* __restrict__ sel_ids, u64 nsel, u64 cps, Ctrl* ctrl,
                                               u32* __restrict__ skeys, u32* __restrict__ sidx, u32* __restrict__ cnt,
                                               u32* ghist2, int hshift2, int bins2, u32 cap) {
    __shared__ u64 scratch[WG / WAVE + 1];
    __shared__ u32 lh[HIST2 ? 256 : 1];
    __shared__ u32 sk[STAGE];
    __shared__ u32 si[IDX ? STAGE : 1];
    h2_init<HIST2>(lh, bins2);
    const u32 thr = ctrl->first.result;
    const u64 L = 1ull << alpha, Lm = L - 1;
    const u64 V = nsel << alpha;
    const u64 nchunks = (V + TT - 1) / TT;
    const u64 nseg = (nchunks + cps - 1) / cps;
    const u64 segcap = cps * TT;
    for (u64 s = blockIdx.x; s < nseg; s += gridDim.x) {
        u64 run = 0;
        for (u64 ch = s * cps; ch < std::min<u64>((s + 1) * cps, nchunks); ++ch) {
            const u64 v0 = ch * TT + (u64)threadIdx.x * TE;
            u32 kv[TE];
            u64 g0[TE / 4];
            u32 take = 0;  // bit j: element j passes
#pragma unroll
            for (int q = 0; q < TE / 4; ++q) {
                const u64 vp = v0 + 4 * q;
                u32 r[4] = {0, 0, 0, 0};
                u64 gq = 0;
                if (vp < V && L >= 4) {
                    gq = ((u64)sel_ids[vp >> alpha] << alpha) + (vp & Lm);
                    if (vec_ok && gq + 3 < n) {
                        const u32x4 x = *(const u32x4*)(keys + gq);
                        r[0] = x.x; r[1] = x.y; r[2] = x.z; r[3] = x.w;
                    } else {
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            if (gq + c < n) r[c] = keys[gq + c];
                    }
                }
                g0[q] = gq;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const u64 vpc = vp + c;
                    u64 g = gq + c;
                    bool ok = vpc < V;
                    if (ok && L < 4) {
                        g = ((u64)sel_ids[vpc >> alpha] << alpha) + (vpc & Lm);
                        if (g < n) r[c] = keys[g];
                    }
                    ok = ok && g < n;
                    const u32 o = K::to_ord(r[c]);
                    kv[4 * q + c] = o;
                    if (ok && (!filter || o >= thr)) take |= 1u << (4 * q + c);
                }
            }
            u64 tot;
            u32 slot = (u32)block_excl_scan_u64<WG>((u64)__popc(take), scratch, &tot);
            if (run + tot > cap) {  // uniform: the segment would overflow its bound
                if (threadIdx.x == 0) ctrl_fail(ctrl, ERR_CAPACITY);
                take = 0;
                tot = 0;
            }
#pragma unroll
            for (int j = 0; j < TE; ++j) {
                if (take & (1u << j)) {
                    sk[slot] = kv[j];
                    if (IDX) {
                        u64 g = g0[j >> 2] + (j & 3);
                        if (L < 4) {
                            const u64 vpc = v0 + j;
                            g = ((u64)sel_ids[vpc >> alpha] << alpha) + (vpc & Lm);
                        }
                        si[slot] = (u32)g;
                    }
                    ++slot;
                }
            }
            __syncthreads();
            flush_stage<IDX, HIST2>(sk, si, (u32)tot, skeys + (s * segcap + run),
                                    IDX ? sidx + (s * segcap + run) : nullptr, lh, hshift2);
            __syncthreads();
            run += tot;
        }
        if (threadIdx.x == 0) cnt[s] = (u32)run;
    }
    h2_flush<HIST2>(lh, ghist2, bins2);
}


// alpha >= 2: each thread owns TE = 16 consecutive slots of the virtual concatenation (the
// selected subranges back to back), 4 granules of 4 keys that each lie in one subrange.  The
// chunk loop is a register pipeline: subrange ids are loaded two chunks ahead and keys one chunk
// ahead, so their latency overlaps the current chunk's scan, staging and copy-out (barriers on
// gfx950 wait for LDS only, not for these loads).
struct GatherIds {
    u32 v[TE / 4];
};
struct GatherKeys {
    u32x4 v[TE / 4];
};

// Index math is 32-bit: n < 2^31, so key ids, subrange ids and virtual positions fit in u32.
// Loads are issued unconditionally from clamped in-bounds addresses: a load under a per-lane
// condition makes the compiler wait for it right where the branches join, which would serialize
// the loads of a thread and defeat the prefetch.  Lanes past the end mask their keys later.
__device__ __forceinline__ void gather_load_ids(const u32* __restrict__ sel_ids, u32 V, int alpha, u32 ch, GatherIds& g) {
    const u32 v0 = ch * (u32)TT + threadIdx.x * TE;
#pragma unroll
    for (int q = 0; q < TE / 4; ++q) {
        const u32 vp = v0 + 4 * q;
        g.v[q] = sel_ids[vp < V ? vp >> alpha : 0u];
    }
}

// VEC: 16-byte aligned keys and n >= 4.  The last granule of the input, when n is not a multiple
// of 4, is loaded from key 0 here and re-read key by key where it is consumed (gather_fix).
template <bool VEC>
__device__ __forceinline__ void gather_load_keys(const u32* __restrict__ keys, u32 n, u32 V, int alpha,
                                                 u32 ch, const GatherIds& g, GatherKeys& k) {
    const u32 v0 = ch * (u32)TT + threadIdx.x * TE, Lm = (1u << alpha) - 1;
#pragma unroll
    for (int q = 0; q < TE / 4; ++q) {
        const u32 vp = v0 + 4 * q;
        const u32 gq = (g.v[q] << alpha) + (vp & Lm);
        if (VEC) {
            k.v[q] = *(const u32x4*)(keys + ((vp < V && gq + 4 <= n) ? gq : 0u));
        } else {
            u32x4 r = {0, 0, 0, 0};
            if (vp < V) {
                if (gq + 0 < n) r.x = keys[gq + 0];
                if (gq + 1 < n) r.y = keys[gq + 1];
                if (gq + 2 < n) r.z = keys[gq + 2];
                if (gq + 3 < n) r.w = keys[gq + 3];
            }
            k.v[q] = r;
        }
    }
}

template <class K, bool HIST2, bool IDX, bool VEC>
__global__ __launch_bounds__(WG) void k_gather(const u32* __restrict__ keys, u64 n64, int alpha, int filter, int vec_ok,
                                               const u32* __restrict__ sel_ids, u64 nsel, u64 cps64, Ctrl* ctrl,
                                               u32* __restrict__ skeys, u32* __restrict__ sidx, u32* __restrict__ cnt,
                                               u32* ghist2, int hshift2, int bins2, u32 cap) {
    __shared__ u32 scratch[WG / WAVE];
    __shared__ u32 lh[HIST2 ? 256 : 1];
    __shared__ u32 sk[GSTAGE];
    __shared__ u32 si[IDX ? GSTAGE : 1];
    h2_init<HIST2>(lh, bins2);
    const u32 thr = ctrl->first.result;
    H2Direct h2d{thr >> hshift2};
    const u32 n = (u32)n64;
    const u32 Lm = (1u << alpha) - 1;
    const u32 V = (u32)(nsel << alpha);
    const u32 cps = (u32)cps64;
    const u32 nchunks = (V + (u32)TT - 1) / (u32)TT;
    const u32 nseg = (nchunks + cps - 1) / cps;
    const u64 segcap = (u64)cps * TT;
    for (u32 s = blockIdx.x; s < nseg; s += gridDim.x) {
        u32 run = 0;
        const u32 cb = s * cps, ce = min((s + 1) * cps, nchunks);
        GatherIds ia, ib;
        GatherKeys ka;
        gather_load_ids(sel_ids, V, alpha, cb, ia);
        if (cb + 1 < ce) gather_load_ids(sel_ids, V, alpha, cb + 1, ib);
        gather_load_keys<VEC>(keys, n, V, alpha, cb, ia, ka);
        for (u32 ch = cb; ch < ce; ++ch) {
            const u32 v0 = ch * (u32)TT + threadIdx.x * TE;
            u32 kv[TE];
            u32 take = 0;  // bit j: element j passes
#pragma unroll
            for (int q = 0; q < TE / 4; ++q) {
                const u32 vp = v0 + 4 * q;
                const u32 gq = (ia.v[q] << alpha) + (vp & Lm);
                const bool gwhole = vp + 4 <= V && gq + 4 <= n;
                if (VEC && !gwhole && vp < V && gq < n)  // the input's partial last granule
                    for (int c = 0; c < 4; ++c) ka.v[q][c] = gq + c < n ? keys[gq + c] : 0u;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const u32 o = K::to_ord(ka.v[q][c]);
                    kv[4 * q + c] = o;
                    const bool ok = gwhole || (vp + c < V && gq + c < n);
                    if (ok && (!filter || o >= thr)) take |= 1u << (4 * q + c);
                }
            }
            // prefetch after this chunk's keys are consumed (a later wait would cover them too)
            GatherIds ic;
            GatherKeys kb;
            if (ch + 2 < ce) gather_load_ids(sel_ids, V, alpha, ch + 2, ic);
            if (ch + 1 < ce) gather_load_keys<VEC>(keys, n, V, alpha, ch + 1, ib, kb);
            u32 tot;
            u32 slot = block_excl_scan_u32<WG>((u32)__popc(take), scratch, &tot);
            if (run + tot > cap) {  // uniform: the segment would overflow its bound
                if (threadIdx.x == 0) ctrl_fail(ctrl, ERR_CAPACITY);
                take = 0;
                tot = 0;
            }
            const u64 gpos = (u64)s * segcap + run;
            const bool staged = tot <= GSTAGE;  // uniform; else survivors are stored directly
            u32* __restrict__ gk = skeys + gpos;
            u32* __restrict__ gi = IDX ? sidx + gpos : nullptr;
#pragma unroll
            for (int j = 0; j < TE; ++j) {
                if (take & (1u << j)) {
                    const u32 gid = (ia.v[j >> 2] << alpha) + ((v0 + j) & Lm);
                    if (staged) {
                        sk[slot] = kv[j];
                        if (IDX) si[slot] = gid;
                    } else {
                        gk[slot] = kv[j];
                        if (IDX) gi[slot] = gid;
                        if (HIST2) h2d.add(lh, kv[j] >> hshift2);
                    }
                    ++slot;
                }
            }
            __syncthreads();
            if (staged) flush_stage<IDX, HIST2>(sk, si, tot, gk, gi, lh, hshift2);
            __syncthreads();
            run += tot;
            ia = ib;
            ib = ic;
            ka = kb;
        }
        if (threadIdx.x == 0) cnt[s] = run;
    }
    if (HIST2) h2d.flush(lh);
    h2_flush<HIST2>(lh, ghist2, bins2);
}

hipError_t launch_gather(const GatherArgs& a, hipStream_t st) {
    const u64 V = a.nsel << a.alpha;
    const u64 nch = (V + TT - 1) / TT;
    const u64 cps = chunks_per_segment(nch);
    const u64 nseg = (nch + cps - 1) / cps;
    const unsigned grid = (unsigned)std::min<u64>(nseg, CGRID);
    const int vec_ok = ((uintptr_t)a.keys & 15) == 0;
    const u64 segcap = cps * TT;
    const u32 cap = (u32)std::min<u64>(segcap, a.seg_limit ? a.seg_limit : segcap);
#define G_ARGS a.keys, a.n, a.alpha, a.filter, vec_ok, a.sel_ids, a.nsel, cps, a.ctrl, a.skeys, a.sidx, a.cnt, \
               a.ghist2, a.hshift2, a.bins2, cap
#define G_LAUNCH(K, H, I)                                                                       \
    do {                                                                                        \
        if (a.alpha >= 2 && vec_ok && a.n >= 4)                                                 \
            hipLaunchKernelGGL((k_gather<K, H, I, true>), dim3(grid), dim3(WG), 0, st, G_ARGS); \
        else if (a.alpha >= 2)                                                                  \
            hipLaunchKernelGGL((k_gather<K, H, I, false>), dim3(grid), dim3(WG), 0, st, G_ARGS); \
        else hipLaunchKernelGGL((k_gather_small<K, H, I>), dim3(grid), dim3(WG), 0, st, G_ARGS);  \
    } while (0)
    const bool h = a.ghist2 != nullptr, i = a.sidx != nullptr;
    if (a.f32) {
        if (h && i) G_LAUNCH(KeyF32, true, true); else if (h) G_LAUNCH(KeyF32, true, false);
        else if (i) G_LAUNCH(KeyF32, false, true); else G_LAUNCH(KeyF32, false, false);
    } else {
        if (h && i) G_LAUNCH(KeyU32, true, true); else if (h) G_LAUNCH(KeyU32, true, false);
        else if (i) G_LAUNCH(KeyU32, false, true); else G_LAUNCH(KeyU32, false, false);
    }
#undef G_LAUNCH
#undef G_ARGS
    return hipGetLastError();
}
u64 gather_segcap(u64 V) { return chunks_per_segment((V + TT - 1) / TT) * TT; }
u64 gather_nseg(u64 V) {
    const u64 nch = (V + TT - 1) / TT;
    const u64 cps = chunks_per_segment(nch);
    return (nch + cps - 1) / cps;
}

// ---------------------------------------------------------------- beta-path range selection + concatenation
constexpr int BSPT = BETA_SUBS_PER_THREAD;  // subranges per thread of the beta gather

// Per chunk of 1024 subranges: (1) each thread classifies its 4 subranges from their
// delegates and packed offsets and the selected ones are listed in LDS in id order;
// (2) the listed subranges' keys are copied to LDS with 16-byte loads (re-read from global
// memory instead when they exceed the buffer) and their kept keys counted, 2^alpha/4 lanes
// per subrange; (3) a block scan places every subrange's output (selected: its kept keys,
// lonely: its delegates, in index order); (4) each listed key goes to its slot by a prefix
// sum over its subrange's lanes.  Survivors are staged in LDS and leave with coalesced stores.
constexpr u32 BSTAGE = 1024;  // staged survivors per chunk (more: direct stores); at 5 waves per SIMD
                              // 2-4% faster than 2048 at 4 waves
constexpr u32 BKBUF = 1024;   // keys of listed subranges held in LDS at 5 waves per SIMD (2048,
                              // 4096: slower)

template <int BETA, bool WHOLE = false>
__device__ __forceinline__ void load_deleg(const u32* __restrict__ dvals, u64 s0, u64 S, u32 (&d)[BSPT][BETA]) {
    if (WHOLE || s0 + BSPT <= S) {
        const u32x4* p = (const u32x4*)(dvals + s0 * BETA);  // BSPT*BETA words, 16-byte aligned
        u32 w[BSPT * BETA];
#pragma unroll
        for (int q = 0; q < BSPT * BETA / 4; ++q) {
            const u32x4 r = p[q];
            w[4 * q] = r.x; w[4 * q + 1] = r.y; w[4 * q + 2] = r.z; w[4 * q + 3] = r.w;
        }
#pragma unroll
        for (int j = 0; j < BSPT; ++j)
#pragma unroll
            for (int b = 0; b < BETA; ++b) d[j][b] = w[j * BETA + b];
    } else {
#pragma unroll
        for (int j = 0; j < BSPT; ++j)
#pragma unroll
            for (int b = 0; b < BETA; ++b) d[j][b] = s0 + j < S ? dvals[(s0 + j) * BETA + b] : 0u;
    }
}

// Packed in-subrange offsets of the first beta-1 delegates of the thread's BSPT subranges
// (u8 / u16 / u32 per subrange for beta 2 / 3 / 4): BSPT * {1,2,4} bytes at an offset that is a
// multiple of that size, read with one aligned load.
template <int BETA, bool WHOLE = false>
__device__ __forceinline__ void load_doff(const void* __restrict__ doff, u64 s0, u64 S, u32 (&pk)[BSPT]) {
    static_assert(BSPT == 4 || BSPT == 8, "packed-offset loads assume 4 or 8 subranges per thread");
    constexpr int EB = BETA == 2 ? 1 : (BETA == 3 ? 2 : 4);  // bytes per subrange
    if (WHOLE || s0 + BSPT <= S) {
        const unsigned char* p = (const unsigned char*)doff + s0 * EB;
        if constexpr (EB * BSPT == 4) {
            const u32 q = *(const u32*)p;
#pragma unroll
            for (int j = 0; j < BSPT; ++j) pk[j] = (q >> (8 * j)) & 0xFFu;
        } else if constexpr (EB * BSPT == 8) {
            const u64 q = *(const u64*)p;
#pragma unroll
            for (int j = 0; j < BSPT; ++j) pk[j] = (u32)(q >> (8 * EB * j)) & (EB == 1 ? 0xFFu : 0xFFFFu);
        } else {
            u32 w[EB * BSPT / 4];
#pragma unroll
            for (int q = 0; q < EB * BSPT / 16; ++q) {
                const u32x4 r = *(const u32x4*)(p + 16 * q);
                w[4 * q] = r.x; w[4 * q + 1] = r.y; w[4 * q + 2] = r.z; w[4 * q + 3] = r.w;
            }
#pragma unroll
            for (int j = 0; j < BSPT; ++j) {
                if (EB == 4) pk[j] = w[j];
                else pk[j] = (w[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < BSPT; ++j) {
            const u64 s = s0 + j;
            pk[j] = 0;
            if (s < S) {
                if (BETA == 2) pk[j] = ((const uint8_t*)doff)[s];
                else if (BETA == 3) pk[j] = ((const uint16_t*)doff)[s];
                else pk[j] = ((const u32*)doff)[s];
            }
        }
    }
}

// The kernel is bound by instruction issue rather than by HBM (about 1.3 GB moved per launch at
// k = 2^24, alpha 4), so everything below works in 32 bits: n < 2^31 keeps key and subrange ids,
// and chunk-local offsets, inside u32.  A chunk whose subranges are all whole (every chunk but
// the last) skips the per-subrange and per-key bounds checks.
template <class K, int BETA, bool POS, bool HIST2, int WAVES, u32 KBUF>
__global__ __launch_bounds__(WG, WAVES) void k_beta_gather(const u32* __restrict__ keys, u64 n64, int alpha, int filter,
                                                    const u32* __restrict__ dvals, const void* __restrict__ doff,
                                                    u64 cps64, Ctrl* ctrl, u32* __restrict__ skeys,
                                                    u32* __restrict__ sidx, u32* __restrict__ cnt, u32* ghist2,
                                                    int hshift2, int bins2, const u8* __restrict__ dtop, u32 cap,
                                                    u64* __restrict__ part) {
    __shared__ u32 scratch[WG / WAVE];
    __shared__ u32 lh[HIST2 ? 256 : 1];
    __shared__ u64 red[4][WG / WAVE];
    __shared__ u32 sk[BSTAGE];
    __shared__ u32 si[POS ? BSTAGE : 1];
    __shared__ uint16_t slist[BETA_TILE_SUBS];  // selected subranges of the chunk (chunk-local ids), ascending
    __shared__ u32 spos[BETA_TILE_SUBS];   // their kept-key counts, then their first output slot
    __shared__ __align__(16) u32 kbuf[KBUF];
    h2_init<HIST2>(lh, bins2);
    const u32 n = (u32)n64;
    const u32 thr = ctrl->first.result;
    const u32 L = 1u << alpha;
    const u32 S = (u32)((n64 + L - 1) >> alpha);
    const u32 cps = (u32)cps64;
    const u32 nchunks = (S + BETA_TILE_SUBS - 1) / BETA_TILE_SUBS;
    const u32 nseg = (nchunks + cps - 1) / cps;
    const u64 segcap = (u64)cps * ((u64)BETA_TILE_SUBS << alpha);
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const int G = alpha >= 2 ? (1 << (alpha - 2)) : 1;  // lanes per listed subrange (alpha <= 8)
    const int gl = lane & (G - 1), grp = lane / G, per_wave = WAVE / G;
    const u32 l0 = threadIdx.x * BSPT;  // chunk-local id of the thread's first subrange
    u32 nsel = 0, nlon = 0, nlel = 0, ndrop = 0;
    // 5 waves per SIMD leave room for the compact delegates of the next chunk (6 VGPRs): they
    // are loaded while this chunk is processed (at 8 waves they would spill)
    constexpr bool PF = WAVES <= 6;
    u32x4 pf_q = {0, 0, 0, 0};
    u64 pf_t = 0;
    for (u32 sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
        u32 run = 0;
        const u32 ch_end = min((sg + 1) * cps, nchunks);
        for (u32 ch = sg * cps; ch < ch_end; ++ch) {
            const u32 c0 = ch * BETA_TILE_SUBS;  // first subrange of the chunk
            const u32 kb = c0 << alpha;          // its first key
            const bool full = ((u64)(c0 + BETA_TILE_SUBS) << alpha) <= n64;  // uniform
            u32 d[BSPT][BETA], pk[BSPT];
            if (BETA == 2 && dtop) {
                static_assert(BSPT == 4, "the top-byte delegate form is read as 4 subranges per thread");
                // first delegates and the second's top bytes: the threshold's low 24 bits are 0,
                // so (top << 24) >= threshold exactly when the second delegate is
                u32 m1[BSPT];
                u64 tw;  // per subrange: the second's top byte, then the first's offset byte
                if (full) {
                    if constexpr (PF) {
                        // this chunk's words were prefetched during the previous chunk of the
                        // segment (from a clamped, whole chunk: always in bounds)
                        u32x4 q;
                        if (ch == sg * cps) {
                            q = *(const u32x4*)(dvals + c0 + l0);
                            tw = *(const u64*)(dtop + 2 * (c0 + l0));
                        } else {
                            q = pf_q;
                            tw = pf_t;
                        }
                        const u32 cn = (ch + 1 < ch_end && ((u64)(c0 + 2 * BETA_TILE_SUBS) << alpha) <= n64)
                                           ? c0 + BETA_TILE_SUBS : c0;
                        pf_q = *(const u32x4*)(dvals + cn + l0);
                        pf_t = *(const u64*)(dtop + 2 * (cn + l0));
                        m1[0] = q.x; m1[1] = q.y; m1[2] = q.z; m1[3] = q.w;
                    } else {
                        const u32x4 q = *(const u32x4*)(dvals + c0 + l0);
                        m1[0] = q.x; m1[1] = q.y; m1[2] = q.z; m1[3] = q.w;
                        tw = *(const u64*)(dtop + 2 * (c0 + l0));
                    }
                } else {
                    tw = 0;
#pragma unroll
                    for (int j = 0; j < BSPT; ++j) {
                        const u32 sj = c0 + l0 + j;
                        m1[j] = sj < S ? dvals[sj] : 0u;
                        tw |= (sj < S ? (u64)((const uint16_t*)dtop)[sj] : 0ull) << (16 * j);
                    }
                }
#pragma unroll
                for (int j = 0; j < BSPT; ++j) pk[j] = (u32)(tw >> (16 * j + 8)) & 0xFFu;
#pragma unroll
                for (int j = 0; j < BSPT; ++j) {
                    d[j][0] = m1[j];
                    d[j][BETA - 1] = ((u32)(tw >> (16 * j)) & 0xFFu) << 24;
                }
            } else if (full) {
                load_deleg<BETA, true>(dvals, c0 + l0, S, d);
                if (POS) load_doff<BETA, true>(doff, c0 + l0, S, pk);
            } else {
                load_deleg<BETA>(dvals, c0 + l0, S, d);
                if (POS) load_doff<BETA>(doff, c0 + l0, S, pk);
            }
            int kind[BSPT];  // 0 none, 1 selected (its keys), 2 lonely (its delegates)
            u32 mj[BSPT];    // lonely: delegates written (with filter bit 2 only those >= threshold)
            u32 ns_t = 0;
#pragma unroll
            for (int j = 0; j < BSPT; ++j) {
                const u32 s = c0 + l0 + j;
                kind[j] = 0;
                mj[j] = 0;
                if ((!full && s >= S) || d[j][0] < thr) continue;
                const int m = full ? BETA : (int)min<u32>(BETA, n - (s << alpha));
                bool all = true;
#pragma unroll
                for (int b = 0; b < BETA; ++b) all = all && (b >= m || d[j][b] >= thr);
                mj[j] = (u32)m;
                if (all) {
                    kind[j] = 1;
                    ++ns_t;
                } else {
                    kind[j] = 2;
                    ++nlon;
                    nlel += (u32)m;
                    if (filter & 2) {
                        // a lonely delegate below the threshold is below the k-th largest key:
                        // counted in the concatenation's size, not written (radix family; the
                        // bucket family's second pass maps the concatenation's min..max range, so
                        // there every delegate is written as the reference does)
                        u32 ge = 0;
#pragma unroll
                        for (int b = 0; b < BETA; ++b) ge += (b < m && d[j][b] >= thr) ? 1u : 0u;
                        mj[j] = ge;
                        ndrop += (u32)m - ge;
                    }
                }
            }
            nsel += ns_t;
            // (1) list the selected subranges
            u32 ns;
            const u32 ex_s = block_excl_scan_u32<WG>(ns_t, scratch, &ns);
            {
                u32 q = ex_s;
#pragma unroll
                for (int j = 0; j < BSPT; ++j)
                    if (kind[j] == 1) slist[q++] = (uint16_t)(l0 + j);
            }
            __syncthreads();
            // uniform; with ns == 0 nothing is listed (slist holds no entry of this chunk)
            const bool inl = alpha >= 2 && ns > 0 && (ns << alpha) <= KBUF;
            if (inl) {
                const u32 ngran = (ns << alpha) >> 2;  // <= KBUF / 4 = KU * WG granules
                const int lg = alpha - 2;
                constexpr int KU = KBUF / (4 * WG);
                u32x4 r[KU];
                u32 e[KU];
#pragma unroll
                for (int u = 0; u < KU; ++u) {
                    const u32 x = u * WG + threadIdx.x;
                    const u32 xc = x < ngran ? x : 0u;
                    // the mask keeps the address inside this chunk's keys whatever slist holds
                    const u32 sub = (u32)slist[xc >> lg] & (BETA_TILE_SUBS - 1);
                    e[u] = kb + (sub << alpha) + ((xc & ((1u << lg) - 1)) << 2);
                }
                if (full) {
                    // unconditional in-bounds loads (past ngran they re-read listed subrange 0):
                    // a load under a per-lane condition is waited for where the branches join
#pragma unroll
                    for (int u = 0; u < KU; ++u) r[u] = *(const u32x4*)(keys + e[u]);
                } else {
#pragma unroll
                    for (int u = 0; u < KU; ++u) {
                        const u32 x = u * WG + threadIdx.x;
                        r[u] = u32x4{0, 0, 0, 0};
                        if (x < ngran) {
                            if (e[u] + 4 <= n) r[u] = *(const u32x4*)(keys + e[u]);
                            else
                                for (int jj = 0; jj < 4; ++jj)
                                    if (e[u] + jj < n) r[u][jj] = keys[e[u] + jj];
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < KU; ++u) {
                    const u32 x = u * WG + threadIdx.x;
                    if (x < ngran) {
                        u32x4 v = r[u];
                        v.x = K::to_ord(v.x); v.y = K::to_ord(v.y); v.z = K::to_ord(v.z); v.w = K::to_ord(v.w);
                        *(u32x4*)(kbuf + 4 * x) = v;
                    }
                }
                __syncthreads();
            }
            // ordered keys and kept mask of granule gl (4 keys) of listed subrange i
            // (i < ns for the caller's use of the result; i >= ns reads listed subrange 0)
            auto fetch4 = [&](u32 i, u32* kv, u32* e_out) -> u32 {
                const u32 ic = i < ns ? i : 0u;
                const u32 e = kb + (((u32)slist[ic] & (BETA_TILE_SUBS - 1)) << alpha) + 4u * (u32)gl;
                *e_out = e;
                u32 okm = 0xF;
                if (inl) {
                    const u32x4 v = *(const u32x4*)(kbuf + (ic << alpha) + 4 * gl);
                    kv[0] = v.x; kv[1] = v.y; kv[2] = v.z; kv[3] = v.w;
                    if (!full)
                        for (int jj = 0; jj < 4; ++jj)
                            if (e + jj >= n) okm &= ~(1u << jj);
                } else if (full || e + 4 <= n) {
                    const u32x4 r = *(const u32x4*)(keys + e);
                    kv[0] = K::to_ord(r.x); kv[1] = K::to_ord(r.y); kv[2] = K::to_ord(r.z); kv[3] = K::to_ord(r.w);
                } else {
                    for (int jj = 0; jj < 4; ++jj) {
                        kv[jj] = e + jj < n ? K::to_ord(keys[e + jj]) : 0u;
                        if (e + jj >= n) okm &= ~(1u << jj);
                    }
                }
                if (filter)
                    for (int jj = 0; jj < 4; ++jj)
                        if (kv[jj] < thr) okm &= ~(1u << jj);
                return okm;
            };
            // (2) kept keys of each listed subrange
            if (alpha >= 2) {
                for (u32 i0 = (u32)wv * per_wave; i0 < ns; i0 += (WG / WAVE) * per_wave) {  // uniform per wave
                    const u32 i = i0 + grp;
                    u32 kv[4], e;
                    const u32 okm = fetch4(i, kv, &e);  // ns >= 1 here, so listed subrange 0 exists
                    u32 c = i < ns ? (u32)__popc(okm) : 0u;
                    for (int o = 1; o < G; o <<= 1) c += (u32)__shfl_xor((int)c, o, WAVE);
                    if (gl == 0 && i < ns) spos[i] = c;
                }
            } else {
                for (u32 i = threadIdx.x; i < ns; i += WG) {
                    const u32 base = kb + ((u32)slist[i] << alpha);
                    const u32 len = min(L, n - base);
                    u32 c = 0;
                    for (u32 e = base; e < base + len; ++e) c += (!filter || K::to_ord(keys[e]) >= thr) ? 1u : 0u;
                    spos[i] = c;
                }
            }
            __syncthreads();
            // (3) output slots
            u32 c = 0;
            {
                u32 q = ex_s;
#pragma unroll
                for (int j = 0; j < BSPT; ++j) {
                    if (kind[j] == 1) c += spos[q++];
                    else if (kind[j] == 2) c += mj[j];
                }
            }
            u32 tot;
            const u32 ex = block_excl_scan_u32<WG>(c, scratch, &tot);
            const bool over = run + tot > cap;  // uniform: the segment would overflow its bound
            if (over) {
                if (threadIdx.x == 0) ctrl_fail(ctrl, ERR_CAPACITY);
                tot = 0;
            }
            const bool staged = tot <= BSTAGE;  // uniform
            const u64 gbase = (u64)sg * segcap + run;
            u32* __restrict__ gk = skeys + gbase;
            u32* __restrict__ gi = POS ? sidx + gbase : nullptr;
            auto put = [&](u32 slot, u32 o, u32 gidx) {
                if (over) return;
                if (staged) {
                    sk[slot] = o;
                    if (POS) si[slot] = gidx;
                } else {
                    gk[slot] = o;
                    if (POS) gi[slot] = gidx;
                    if (HIST2) atomicAdd(&lh[o >> hshift2], 1u);
                }
            };
            if (BETA == 2 && (filter & 2)) {
                // beta 2 with lonely delegates below the threshold dropped: a lonely subrange
                // writes exactly its first delegate (the second is below the threshold)
                u32 slot = ex;
                u32 q = ex_s;
#pragma unroll
                for (int j = 0; j < BSPT; ++j) {
                    const bool sel = kind[j] == 1, lon = kind[j] == 2;
                    u32 cj = lon ? 1u : 0u;
                    if (sel) {
                        cj = spos[q];
                        spos[q] = slot;
                    }
                    if (lon) put(slot, d[j][0], kb + ((l0 + j) << alpha) + (POS ? (pk[j] & 0xFFu) : 0u));
                    slot += cj;
                    q += sel ? 1u : 0u;
                }
            } else {
                u32 slot = ex;
                u32 q = ex_s;
#pragma unroll
                for (int j = 0; j < BSPT; ++j) {
                    if (kind[j] == 1) {
                        const u32 cj = spos[q];
                        spos[q] = slot;
                        slot += cj;
                        ++q;
                    } else if (kind[j] == 2) {
                        const u32 base = kb + ((l0 + j) << alpha);
                        u32 dv[BETA], p[BETA];
                        const u32 packed = POS ? pk[j] : 0u;
#pragma unroll
                        for (int b = 0; b < BETA; ++b) {
                            dv[b] = d[j][b];
                            const bool ge = b < (int)mj[j] && dv[b] >= thr;  // only these can reach the output
                            p[b] = (POS && ge && b < BETA - 1) ? base + ((packed >> (8 * b)) & 0xFFu) : 0xFFFFFFFFu;
                        }
                        // delegates >= threshold in index order, then the rest (never selected later)
#pragma unroll
                        for (int x = 0; x < BETA; ++x)
#pragma unroll
                            for (int y = 0; y + 1 < BETA - x; ++y)
                                if (p[y + 1] < p[y]) {
                                    const u32 tp = p[y]; p[y] = p[y + 1]; p[y + 1] = tp;
                                    const u32 td = dv[y]; dv[y] = dv[y + 1]; dv[y + 1] = td;
                                }
#pragma unroll
                        for (int b = 0; b < BETA; ++b)
                            if (b < (int)mj[j]) put(slot++, dv[b], p[b]);
                    }
                }
            }
            __syncthreads();
            // (4) keys of the listed subranges to their slots
            if (alpha >= 2) {
                for (u32 i0 = (u32)wv * per_wave; i0 < ns; i0 += (WG / WAVE) * per_wave) {
                    const u32 i = i0 + grp;
                    u32 kv[4], e;
                    u32 okm = fetch4(i, kv, &e);
                    if (i >= ns) okm = 0;
                    const u32 c4 = (u32)__popc(okm);
                    u32 incl = c4;
                    for (int o = 1; o < G; o <<= 1) {
                        const u32 tt = (u32)__shfl_up((int)incl, o, WAVE);
                        if (gl >= o) incl += tt;
                    }
                    if (i < ns) {
                        u32 slot = spos[i] + (incl - c4);
                        for (int jj = 0; jj < 4; ++jj)
                            if ((okm >> jj) & 1u) put(slot++, kv[jj], e + jj);
                    }
                }
            } else {
                for (u32 i = threadIdx.x; i < ns; i += WG) {
                    const u32 base = kb + ((u32)slist[i] << alpha);
                    const u32 len = min(L, n - base);
                    u32 slot = spos[i];
                    for (u32 e = base; e < base + len; ++e) {
                        const u32 o = K::to_ord(keys[e]);
                        if (!filter || o >= thr) put(slot++, o, e);
                    }
                }
            }
            __syncthreads();
            if (staged) flush_stage<POS, HIST2>(sk, si, tot, gk, gi, lh, hshift2);
            __syncthreads();
            run += tot;
        }
        if (threadIdx.x == 0) cnt[sg] = run;
    }
    {
        const int w = threadIdx.x >> 6;
        u64 a4[4] = {nsel, nlon, nlel, ndrop};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a4[q] = wave_sum_u64(a4[q]);
            if (lane == 0) red[q][w] = a4[q];
        }
        __syncthreads();
        if (threadIdx.x == 0) {  // this workgroup's record (summed by beta_part_total)
            u64 a = 0, b = 0, cc = 0, dd = 0;
            for (int i = 0; i < WG / WAVE; ++i) { a += red[0][i]; b += red[1][i]; cc += red[2][i]; dd += red[3][i]; }
            u64* r = part + 4ull * blockIdx.x;
            r[0] = a;
            r[1] = b;
            r[2] = cc;
            r[3] = dd;
        }
    }
    h2_flush<HIST2>(lh, ghist2, bins2);
}

// Waves per SIMD of the beta gather: 32-key subranges (alpha 5) run faster at 8 (64 VGPRs); at
// alpha <= 4 the random 16-byte subrange reads are the limit and 5 do better.
static int beta_waves(int alpha) { return alpha >= 5 ? 8 : 5; }

// Resident workgroups of the beta gather on the current device, from the occupancy of one fixed
// instantiation (uint32 keys, beta 2, positions, fused histogram) so that planning and launch
// agree whatever the call's outputs; cached per device.
template <int W, u32 KB>
static u64 beta_slots_t() {
    static std::atomic<int> cache[64];  // 0 until set; calls from several threads store equal values
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int nb = cache[dev].load(std::memory_order_relaxed);
    if (nb == 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, reinterpret_cast<const void*>(&k_beta_gather<KeyU32, 2, true, true, W, KB>), WG, 0) != hipSuccess ||
            nb < 1)
            nb = W;  // one workgroup per wave slot of a SIMD
        cache[dev].store(nb, std::memory_order_relaxed);
    }
    return (u64)nb * device_cus();
}
static u64 beta_slots(int alpha) {
    return beta_waves(alpha) == 8 ? beta_slots_t<8, 1024>() : beta_slots_t<5, BKBUF>();
}

// Chunks per beta-gather segment.  A workgroup runs one segment at a time, so equal segments
// finish in rounds of the resident workgroups; once segments would exceed 16 chunks' worth of
// rounds, one segment per resident workgroup avoids a partly filled last round (2^30 keys,
// alpha 4: 4096 segments would be 2.7 rounds of the 1536 resident workgroups).
static u64 beta_cps(u64 nch, int alpha) {
    const u64 cps = chunks_per_segment(nch);
    const u64 slots = beta_slots(alpha);
    if (nch > slots * cps) return (nch + slots - 1) / slots;
    return cps;
}
u64 beta_nseg(u64 S, int alpha) {
    const u64 nch = (S + BETA_TILE_SUBS - 1) / BETA_TILE_SUBS;
    const u64 cps = beta_cps(nch, alpha);
    return (nch + cps - 1) / cps;
}
u64 beta_grid(u64 S, int alpha) { return std::min<u64>(beta_nseg(S, alpha), CGRID); }
u64 beta_segcap(u64 S, int alpha) {
    return beta_cps((S + BETA_TILE_SUBS - 1) / BETA_TILE_SUBS, alpha) * (BETA_TILE_SUBS << alpha);
}

template <class K, int BETA>
static hipError_t launch_beta_t(const BetaGatherArgs& a, hipStream_t st) {
    const u64 L = 1ull << a.alpha;
    const u64 S = (a.n + L - 1) >> a.alpha;
    const u64 nch = (S + BETA_TILE_SUBS - 1) / BETA_TILE_SUBS;
    const u64 cps = beta_cps(nch, a.alpha);
    const u64 nseg = (nch + cps - 1) / cps;
    const unsigned grid = (unsigned)std::min<u64>(nseg, CGRID);
    if (grid != beta_grid(S, a.alpha)) return hipErrorInvalidValue;  // a.part holds beta_grid records
    const u64 segcap = cps * (BETA_TILE_SUBS << a.alpha);
    const u32 cap = (u32)std::min<u64>(segcap, a.seg_limit ? a.seg_limit : segcap);
#define B_ARGS a.keys, a.n, a.alpha, a.filter, a.dvals, a.doff, cps, a.ctrl, a.skeys, a.sidx, a.cnt, a.ghist2, \
               a.hshift2, a.bins2, BETA == 2 ? a.dtop : nullptr, cap, a.part
    const bool pos = a.sidx != nullptr, h = a.ghist2 != nullptr;
#define B_LAUNCH(W, KB)                                                                                          \
    do {                                                                                                         \
        if (pos && h) hipLaunchKernelGGL((k_beta_gather<K, BETA, true, true, W, KB>), dim3(grid), dim3(WG), 0, st, B_ARGS);   \
        else if (pos) hipLaunchKernelGGL((k_beta_gather<K, BETA, true, false, W, KB>), dim3(grid), dim3(WG), 0, st, B_ARGS);  \
        else if (h) hipLaunchKernelGGL((k_beta_gather<K, BETA, false, true, W, KB>), dim3(grid), dim3(WG), 0, st, B_ARGS);    \
        else hipLaunchKernelGGL((k_beta_gather<K, BETA, false, false, W, KB>), dim3(grid), dim3(WG), 0, st, B_ARGS);          \
    } while (0)
    if (beta_waves(a.alpha) == 8) B_LAUNCH(8, 1024);
    else B_LAUNCH(5, BKBUF);
#undef B_LAUNCH
#undef B_ARGS
    return hipGetLastError();
}
template <class K>
static hipError_t launch_beta_k(const BetaGatherArgs& a, hipStream_t st) {
    switch (a.beta) {
        case 2: return launch_beta_t<K, 2>(a, st);
        case 3: return launch_beta_t<K, 3>(a, st);
        case 4: return launch_beta_t<K, 4>(a, st);
    }
    return hipErrorInvalidValue;
}
hipError_t launch_beta_gather(const BetaGatherArgs& a, hipStream_t st) {
    // the top-byte form holds only what a threshold on the top 8-bit digit with lonely
    // delegates below it dropped can use
    if (a.dtop && !(a.beta == 2 && (a.filter & 2))) return hipErrorInvalidValue;
    return a.f32 ? launch_beta_k<KeyF32>(a, st) : launch_beta_k<KeyU32>(a, st);
}

// ---------------------------------------------------------------- concatenation size
__global__ __launch_bounds__(SCAN_T) void k_seg_total(const u32* __restrict__ cnt, u64 nseg, Ctrl* ctrl,
                                                     const u64* __restrict__ part, u64 npart) {
    __shared__ u64 scratch[SCAN_T / WAVE + 1];
    __shared__ u64 pl[4][SCAN_T / WAVE];
    if (part) beta_part_total<SCAN_T>(part, npart, ctrl, pl);
    u64 s = 0;
#pragma unroll 8
    for (u64 i = threadIdx.x; i < nseg; i += SCAN_T) s += cnt[i];
    u64 tot;
    block_excl_scan_u64<SCAN_T>(s, scratch, &tot);
    if (threadIdx.x == 0) {
        ctrl->concat_size = tot + ctrl->lonely_dropped;  // the dropped lonely delegates count too
        ctrl->second.m = tot;
        if (tot < ctrl->second.k) {
            ctrl_fail(ctrl, ERR_CONCAT_SHORT);
            ctrl->second.done = 1;  // nothing sensible to select; the call reports the error
        }
    }
}
hipError_t launch_seg_total(const u32* cnt, u64 nseg, Ctrl* ctrl, hipStream_t st, const u64* part, u64 npart) {
    hipLaunchKernelGGL(k_seg_total, dim3(1), dim3(SCAN_T), 0, st, cnt, nseg, ctrl, part, npart);
    return hipGetLastError();
}

// ---------------------------------------------------------------- extraction
template <bool DIRECT>
__device__ __forceinline__ u64 seg_len(const Seg& g, u64 s) {
    return DIRECT ? std::min<u64>(g.segcap, g.n - s * g.segcap) : (u64)g.cnt[s];
}

template <class K, bool DIRECT>
__global__ __launch_bounds__(WG) void k_extract_count(Seg g, SelState* sel, u64* __restrict__ cnt) {
    __shared__ u64 scratch[WG / WAVE + 1];
    const u32 K0 = sel->result;
    for (u64 s = blockIdx.x; s < g.nseg; s += gridDim.x) {
        const u64 len = seg_len<DIRECT>(g, s);
        const u32* base = g.keys + s * g.segcap;
        u32 gt = 0, eq = 0;
        for (u64 c0 = 0; c0 < len; c0 += TT) {
            const u64 e0 = c0 + (u64)threadIdx.x * TE;
            const bool whole = c0 + TT <= len;
            u32 v[TE];
            load_te<K, DIRECT>(base + e0, e0, len, v);
            count_gt_eq(v, K0, whole, e0 < len ? (u32)std::min<u64>(len - e0, TE) : 0u, gt, eq);
        }
        u64 tot;
        block_excl_scan_u64<WG>(pack2(gt, eq), scratch, &tot);
        if (threadIdx.x == 0) cnt[s] = tot;
    }
}

// Per chunk of TT keys the outputs are placed relative to the chunk's first slot P0 in 32 bits:
// with lg / lq the keys > K / == K before a key inside the chunk and rem the ties still to take
// when the chunk starts, a key > K goes to lg + min(lq, rem) and a key == K with lq < rem to lg + lq.
// off: the segments' exclusive prefixes of packed (gt, eq) counts, or null: then each
// workgroup sums the counts `cnt` of the segments before its own (few segments: this replaces
// a one-workgroup scan launch).
template <class K, bool DIRECT, bool KEYS, bool IDX>
__global__ __launch_bounds__(WG) void k_extract_write(Seg g, SelState* sel, u64 k, const u64* __restrict__ off,
                                                      const u64* __restrict__ cnt, u32* __restrict__ out_keys,
                                                      i64* __restrict__ out_idx, Ctrl* ctrl, int fin, u32* d_kth) {
    __shared__ u32 scratch[WG / WAVE];
    __shared__ u64 red[WG / WAVE];
    __shared__ u32 sk[KEYS ? STAGE : 1];
    __shared__ u32 si[IDX ? STAGE : 1];
    const u32 K0 = sel->result;
    const u64 total_gt = sel->m - sel->below - sel->kcount;
    const u64 r = k - total_gt;  // ties taken, lowest index first
    if (fin && blockIdx.x == 0 && threadIdx.x == 0) {  // k_finalize's work, saving its launch
        ctrl->kth = K0;
        ctrl->count_gt = total_gt;
        ctrl->ties = r;
        if (d_kth) *d_kth = K::from_ord(K0);
    }
    for (u64 s = blockIdx.x; s < g.nseg; s += gridDim.x) {
        const u64 len = seg_len<DIRECT>(g, s);
        const u32* base = g.keys + s * g.segcap;
        u64 o;
        if (off) {
            o = off[s];
        } else {
            u64 part = 0;  // packed fields add without carries: each total is below 2^31
            for (u64 t = threadIdx.x; t < s; t += WG) part += cnt[t];
            part = wave_sum_u64(part);
            if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = part;
            __syncthreads();
            o = 0;
#pragma unroll
            for (int w = 0; w < WG / WAVE; ++w) o += red[w];
            __syncthreads();
        }
        u64 run_g = unpack_hi(o), run_q = unpack_lo(o);
        for (u64 c0 = 0; c0 < len; c0 += TT) {
            const u64 e0 = c0 + (u64)threadIdx.x * TE;
            const bool whole = c0 + TT <= len;  // uniform
            const u32 nvalid = e0 < len ? (u32)std::min<u64>(len - e0, TE) : 0u;
            u32 v[TE];
            u32 vi[(IDX && !DIRECT) ? TE : 1];
            load_te<K, DIRECT>(base + e0, e0, len, v);
            // the concatenation's indices, read with the keys (16-byte loads) rather than one by one later
            if (IDX && !DIRECT) load_te<KeyU32, false>(g.idx + s * g.segcap + e0, e0, len, vi);
            u32 gt = 0, eq = 0;
            count_gt_eq(v, K0, whole, nvalid, gt, eq);
            u32 tot;  // packed (gt << 16) | eq: at most TT = 4096 each
            const u32 ex = block_excl_scan_u32<WG>((gt << 16) | eq, scratch, &tot);
            const u64 rem64 = r - std::min<u64>(run_q, r);
            const u32 rem = (u32)std::min<u64>(rem64, TT);
            // the chunk's outputs occupy [P0, P0 + c) of the top-k arrays
            const u64 P0 = run_g + std::min<u64>(run_q, r);
            const u32 c = (tot >> 16) + std::min<u32>(tot & 0xFFFFu, rem);
            u32 lg = ex >> 16, lq = ex & 0xFFFFu;
#pragma unroll
            for (int j = 0; j < TE; ++j) {
                const bool in = whole || (u32)j < nvalid;
                const bool isg = in && v[j] > K0, ise = in && v[j] == K0;
                const u32 sl = isg ? lg + min(lq, rem) : lg + lq;
                if (isg || (ise && lq < rem)) {
                    if (KEYS) sk[sl] = K::from_ord(v[j]);
                    if (IDX) si[sl] = DIRECT ? (u32)(s * g.segcap + e0 + j) : vi[(IDX && !DIRECT) ? j : 0];
                }
                lg += isg ? 1u : 0u;
                lq += ise ? 1u : 0u;
            }
            __syncthreads();
            if (P0 + c > k) {
                if (threadIdx.x == 0) ctrl_fail(ctrl, ERR_CAPACITY);
            } else {
                u32* __restrict__ ok = KEYS ? out_keys + P0 : nullptr;
                i64* __restrict__ oi = IDX ? out_idx + P0 : nullptr;
                for (u32 i = threadIdx.x; i < c; i += WG) {
                    if (KEYS) ok[i] = sk[i];
                    if (IDX) oi[i] = (i64)si[i];
                }
            }
            __syncthreads();
            run_g += tot >> 16;
            run_q += tot & 0xFFFFu;
        }
    }
}

hipError_t launch_extract(const ExtractArgs& a, hipStream_t st) {
    const unsigned grid = (unsigned)std::max<u64>(1, std::min<u64>(a.seg.nseg, CGRID));
#define EXC(K, D) hipLaunchKernelGGL((k_extract_count<K, D>), dim3(grid), dim3(WG), 0, st, a.seg, a.sel, a.cnt)
    if (a.f32) { if (a.direct) EXC(KeyF32, true); else EXC(KeyF32, false); }
    else { if (a.direct) EXC(KeyU32, true); else EXC(KeyU32, false); }
#undef EXC
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // up to 2048 segments every workgroup sums the counts before its segment itself
    const bool fused = a.seg.nseg <= 2048;
    if (!fused) {
        e = launch_scan(a.cnt, a.seg.nseg, a.off, nullptr, st);
        if (e != hipSuccess) return e;
    }
    const u64* off = fused ? nullptr : a.off;
#define EXW(K, D, KK, II)                                                                                        \
    hipLaunchKernelGGL((k_extract_write<K, D, KK, II>), dim3(grid), dim3(WG), 0, st, a.seg, a.sel, a.k, off,     \
                       a.cnt, (u32*)a.out_keys, a.out_idx, a.ctrl, a.finalize ? 1 : 0, (u32*)a.d_kth)
#define EXW2(K, D)                                                                                                \
    do {                                                                                                          \
        if (a.out_keys && a.out_idx) EXW(K, D, true, true);                                                       \
        else if (a.out_keys) EXW(K, D, true, false);                                                              \
        else EXW(K, D, false, true);                                                                              \
    } while (0)
    if (a.f32) { if (a.direct) EXW2(KeyF32, true); else EXW2(KeyF32, false); }
    else { if (a.direct) EXW2(KeyU32, true); else EXW2(KeyU32, false); }
#undef EXW2
#undef EXW
    return hipGetLastError();
}

}  // namespace drtopk
