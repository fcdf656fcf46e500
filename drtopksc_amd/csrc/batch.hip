// batch.hip -- batched top-k over many rows of uneven length, each with its own k and key
// type (drtopk_batch_plan / drtopk_select_batched).
//
// The reference handles one vector per call and drives every phase from the host
// (radixselect.cuh:1783-1894); its MultiGPU driver loops that per vector
// (MultiGPU/main.cu:366-424).  Here a whole batch is two launches:
//
//  k_bsample  every row's delegate pass (one maximum per subrange of 2^alpha keys,
//             sampleMax, radixselect.cuh:883-918) as one flat, persistent grid over
//             fixed-size tiles of all rows, 16-byte loads, eight in flight per lane.
//             It is the only kernel that reads all keys.
//  k_brows    one workgroup per row (persistent, rows taken largest-first from a queue):
//             first top-k over the row's delegates (exact k-th largest delegate by radix
//             digits), range selection (every subrange whose delegate is greater, then the
//             lowest-ID ties up to k, radixselect.cuh:413-427 / :603-618), filtered
//             concatenation of those subranges (keys >= threshold, radixselect.cuh:1575-1607)
//             into LDS, second top-k over it and the index-ordered extraction of the
//             top-k keys and indices (MultiGPU/radixselect.cuh:1355-1383).  When the
//             concatenation does not fit in LDS it is re-read from the row's keys through the
//             selected-subrange list on every pass instead (bounded memory, same result).
//
// Why the threshold works (the same argument as the single-vector path): the k-th largest
// delegate T is <= the row's k-th largest key K; every key > T sits in a subrange whose
// delegate is > T, and when K == T the lowest-ID subranges with delegate == T hold the
// lowest-index keys equal to K that the top-k set needs.  So the filtered concatenation
// holds every key of the top-k set and K is its k-th largest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/drtopk.h"
#include "common.hpp"
#include "host_util.hpp"

namespace drtopk {
namespace {

constexpr u64 PLAN_MAGIC = 0x3142544b50545244ull;  // "DRTPKTB1"

// Plan blob: header, rows, chunk -> row map, row processing order.  Offsets are bytes
// from the start of the blob, so the same bytes work on the host and on the device.
struct BatchHeader {
    u64 magic;
    u64 nrows;
    u64 nchunks;
    u64 o_rows, o_tiles, o_order;
    u64 plan_bytes;
    u64 ws_bytes;
    u64 o_ws_deleg, o_ws_sel;  // workspace byte offsets: delegates, selected subranges
    u64 keys_len, out_len;
    u32 sample_grid, rows_grid;
    u32 pad[6];
};
static_assert(sizeof(BatchHeader) == 128, "header layout");

struct BatchRow {
    u64 key_off;    // first key in d_keys (multiple of 4)
    u64 n, k;
    u64 out_off;    // first output slot
    u64 deleg_off;  // delegates in the workspace delegate array (multiple of 4)
    u64 sel_off;    // selected subrange ids in the workspace id array, then (when k << alpha
                    // exceeds the LDS concatenation cache) room for k << alpha concatenated keys
                    // and as many row indices
    u32 S;          // subranges (0 when alpha == 0)
    u32 chunk_begin, nchunks;
    u8 alpha, f32, nb, tile_log2;
};
static_assert(sizeof(BatchRow) == 64, "row layout");

// One tile of the delegate pass in 16 bytes, so that a workgroup reads a single record per tile
// (the next tile's while it streams the current one) instead of a tile -> row map entry and then
// the row's record.
struct TileDesc {
    u64 src;  // bits 0..47: first key (index into d_keys); 48..53: alpha; 54: float32 keys
    u64 dst;  // bits 0..39: first delegate (index into the delegate array); 40..55: length - 1
};
static_assert(sizeof(TileDesc) == 16, "tile descriptor layout");

// ---------------------------------------------------------------- delegate pass
constexpr int BS_NT = 256;
constexpr int BS_UNR = 8;  // 16-byte loads in flight per lane
constexpr int MAX_TILE_LOG2 = 15;

template <bool F32>
__device__ __forceinline__ u32 ord(u32 b) {
    return F32 ? KeyF32::to_ord(b) : b;
}

// Maximum over each aligned group of g lanes (g = 1, 2, 4, .., 64, uniform over the wave),
// complete in the group's last lane, with DPP moves alone: quad permutes and the half-row / row
// mirrors give every lane of a 16-lane row its row's maximum, row broadcast 15 then carries row 0
// into row 1 and row 2 into row 3 (32-lane groups), and row broadcast 31 row 1 into rows 2 and 3
// (the whole wave).
__device__ __forceinline__ u32 group_max_last(u32 m, int g) {
    if (g >= 2) m = max(m, (u32)__builtin_amdgcn_mov_dpp((int)m, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    if (g >= 4) m = max(m, (u32)__builtin_amdgcn_mov_dpp((int)m, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    if (g >= 8) m = max(m, (u32)__builtin_amdgcn_mov_dpp((int)m, 0x141, 0xF, 0xF, false));  // row_half_mirror
    if (g >= 16) m = max(m, (u32)__builtin_amdgcn_mov_dpp((int)m, 0x140, 0xF, 0xF, false)); // row_mirror
    if (g >= 32) m = max(m, (u32)__builtin_amdgcn_update_dpp(0, (int)m, 0x142, 0xA, 0xF, false));  // row_bcast:15
    if (g >= 64) m = max(m, (u32)__builtin_amdgcn_update_dpp(0, (int)m, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return m;
}

__device__ __forceinline__ u32 max4(const u32x4& v) { return max(max(v[0], v[1]), max(v[2], v[3])); }

constexpr u32 BS_WKEYS = BS_UNR * WAVE * 4;  // keys per wave per batch (2048)
constexpr u32 BS_BATCH = BS_WKEYS * (BS_NT / WAVE);  // keys per workgroup per batch (8192)

// One tile (<= 2^15 keys) of one row: subrange maxima into dout.  Each wave owns 2048
// consecutive keys per batch (eight 16-byte loads per lane, 1 KB per load instruction), so a
// subrange of up to 2^11 keys lies inside one wave: alpha 2..8 reduce within lane groups,
// alpha 9..11 across 2^(alpha-8) of the wave's loads, with no barrier.  Longer subranges
// (alpha 12..15) combine the waves' 2048-key maxima through LDS.
template <bool F32>
__device__ __forceinline__ void bsample_tile(const u32* __restrict__ src, u64 len, int a, u32* __restrict__ dout,
                                             u32* part) {
    const int t = threadIdx.x, lane = t & (WAVE - 1), w = t / WAVE;
    const u32 nbat = (u32)((len + BS_BATCH - 1) / BS_BATCH);
    for (u32 b = 0; b < nbat; ++b) {
        const u64 wb = (u64)b * BS_BATCH + (u64)w * BS_WKEYS;  // the wave's first key
        if (wb >= len) break;  // uniform over the wave; no barrier below
        u32x4 q[BS_UNR];
        if (wb + BS_WKEYS <= len) {
#pragma unroll
            for (int u = 0; u < BS_UNR; ++u)
                q[u] = __builtin_nontemporal_load((const u32x4*)(src + wb + (u64)(u * WAVE + lane) * 4));
#pragma unroll
            for (int u = 0; u < BS_UNR; ++u)
#pragma unroll
                for (int j = 0; j < 4; ++j) q[u][j] = ord<F32>(q[u][j]);
        } else {
            // the row's last tile: 16-byte loads for every vector inside it, key by key for the
            // one vector it ends in (the keys after a row may lie outside d_keys); 0 is below
            // every ordered key
#pragma unroll
            for (int u = 0; u < BS_UNR; ++u) {
                const u64 e = wb + (u64)(u * WAVE + lane) * 4;
                if (e + 4 <= len) {
                    q[u] = *(const u32x4*)(src + e);
#pragma unroll
                    for (int j = 0; j < 4; ++j) q[u][j] = ord<F32>(q[u][j]);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j) q[u][j] = (e + j < len) ? ord<F32>(src[e + j]) : 0u;
                }
            }
        }
        if (a == 1) {
#pragma unroll
            for (int u = 0; u < BS_UNR; ++u) {
                const u64 e = wb + (u64)(u * WAVE + lane) * 4;
                if (e < len) dout[e >> 1] = max(q[u][0], q[u][1]);
                if (e + 2 < len) dout[(e >> 1) + 1] = max(q[u][2], q[u][3]);
            }
        } else if (a <= 8) {
            const int g = 1 << (a - 2);  // lanes per subrange
#pragma unroll
            for (int u = 0; u < BS_UNR; ++u) {
                const u64 e = wb + (u64)(u * WAVE + lane) * 4;
                const u32 m = group_max_last(max4(q[u]), g);
                const u64 ef = e - 4ull * (u64)(lane & (g - 1));  // the group's (subrange's) first key
                if ((lane & (g - 1)) == g - 1 && ef < len) dout[ef >> a] = m;
            }
        } else if (a <= 11) {
            const int per = 1 << (a - 8);  // loads per subrange
            u32 acc = 0;
#pragma unroll
            for (int u = 0; u < BS_UNR; ++u) {
                acc = max(acc, max4(q[u]));
                if (((u + 1) & (per - 1)) == 0) {
                    const u64 e0 = wb + (u64)(u + 1 - per) * WAVE * 4;  // the subrange's first key
                    const u32 m = group_max_last(acc, WAVE);
                    if (lane == WAVE - 1 && e0 < len) dout[e0 >> a] = m;
                    acc = 0;
                }
            }
        } else {
            u32 acc = 0;
#pragma unroll
            for (int u = 0; u < BS_UNR; ++u) acc = max(acc, max4(q[u]));
            acc = group_max_last(acc, WAVE);
            if (lane == WAVE - 1) part[wb / BS_WKEYS] = acc;  // the wave's 2048-key maximum
        }
    }
    if (a >= 12) {
        __syncthreads();
        const u32 per = 1u << (a - 11);
        const u32 nblk = (u32)((len + BS_WKEYS - 1) / BS_WKEYS);
        const u32 nsub = (u32)((len + (1ull << a) - 1) >> a);
        for (u32 sb = t; sb < nsub; sb += BS_NT) {
            u32 m = 0;
            for (u32 j = 0; j < per; ++j) {
                const u32 bk = sb * per + j;
                if (bk < nblk) m = max(m, part[bk]);
            }
            dout[sb] = m;
        }
        __syncthreads();
    }
}

// Workspace of a batched call: the error words (WsHead, as in every workspace), then the row
// queue counter of k_brows, then the delegate and selected-id regions (plan offsets).
struct BatchWs {
    WsHead head;
    u32 queue;
};

__global__ __launch_bounds__(BS_NT) void k_bsample(const unsigned char* __restrict__ plan, const u32* __restrict__ keys,
                                                   BatchWs* wsh, u32* __restrict__ deleg) {
    const BatchHeader* h = (const BatchHeader*)plan;
    const u32x4* td = (const u32x4*)(plan + h->o_tiles);
    const u64 nch = h->nchunks;
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // the call's first kernel resets the head
        ws_head_begin(&wsh->head);
        wsh->queue = 0;  // row queue of k_brows (next launch)
    }
    __shared__ u32 part[1 << (MAX_TILE_LOG2 - 11)];  // 2048-key block maxima (alpha >= 12)
    if (blockIdx.x >= nch) return;
    u32x4 cur = td[blockIdx.x];
    for (u64 c = blockIdx.x; c < nch; c += gridDim.x) {
        // the next tile's record, loaded unconditionally (clamped) so nothing waits for it here
        const u64 cn = c + gridDim.x < nch ? c + gridDim.x : c;
        const u32x4 nxt = td[cn];
        const u64 s0 = ((u64)cur.y << 32) | cur.x, d0 = ((u64)cur.w << 32) | cur.z;
        const int a = (int)((s0 >> 48) & 63);
        const bool f32 = (s0 >> 54) & 1;
        const u64 len = ((d0 >> 40) & 0xFFFF) + 1;
        const u32* src = keys + (s0 & ((1ull << 48) - 1));
        u32* dout = deleg + (d0 & ((1ull << 40) - 1));
        if (f32) bsample_tile<true>(src, len, a, dout, part);
        else bsample_tile<false>(src, len, a, dout, part);
        cur = nxt;
    }
}

// ---------------------------------------------------------------- per-row workgroup
constexpr int BR_NT = 512;
constexpr int BR_W = BR_NT / WAVE;
constexpr u32 REG_WORDS = 12288;       // 48 KB: candidate list or concatenation cache
constexpr u32 CONCAT_CAP = REG_WORDS / 2;  // keys + row indices

// Sources: a sequence of E slots read four at a time (e a multiple of 4); each slot yields
// an ordered key, its row index (or element index) and a valid bit.
struct SrcG {  // global ordered u32 array (delegates), 16-byte aligned
    const u32* p;
    u64 m;
    __device__ __forceinline__ u32 load4(u64 e, u32* x, u32* ix) const {
        u32 v = 0;
        if (e + 4 <= m) {
            const u32x4 q = *(const u32x4*)(p + e);
            x[0] = q[0]; x[1] = q[1]; x[2] = q[2]; x[3] = q[3];
            v = 0xF;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool ok = e + j < m;
                x[j] = ok ? p[e + j] : 0u;
                v |= (ok ? 1u : 0u) << j;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) ix[j] = (u32)(e + j);
        return v;
    }
    // slots [0, bulk_end(E)) can be read with load4_fast: no per-lane bounds test
    __device__ __forceinline__ u64 bulk_end(u64 E, u64 step) const { return min(E, m) / step * step; }
    __device__ __forceinline__ u32 load4_fast(u64 e, u32* x, u32* ix) const {
        const u32x4 q = *(const u32x4*)(p + e);
        x[0] = q[0]; x[1] = q[1]; x[2] = q[2]; x[3] = q[3];
#pragma unroll
        for (int j = 0; j < 4; ++j) ix[j] = (u32)(e + j);
        return 0xF;
    }
};

struct SrcL {  // keys + row indices (LDS, or a row's concatenation in the workspace), 16-byte aligned
    const u32* key;
    const u32* idx;
    u64 m;
    __device__ __forceinline__ u32 load4(u64 e, u32* x, u32* ix) const {
        u32 v = 0;
        if (e + 4 <= m) {
            const u32x4 q = *(const u32x4*)(key + e);
            const u32x4 r = *(const u32x4*)(idx + e);
            x[0] = q[0]; x[1] = q[1]; x[2] = q[2]; x[3] = q[3];
            ix[0] = r[0]; ix[1] = r[1]; ix[2] = r[2]; ix[3] = r[3];
            v = 0xF;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool ok = e + j < m;
                x[j] = ok ? key[e + j] : 0u;
                ix[j] = ok ? idx[e + j] : 0u;
                v |= (ok ? 1u : 0u) << j;
            }
        }
        return v;
    }
    __device__ __forceinline__ u64 bulk_end(u64 E, u64 step) const { return min(E, m) / step * step; }
    __device__ __forceinline__ u32 load4_fast(u64 e, u32* x, u32* ix) const {
        const u32x4 q = *(const u32x4*)(key + e);
        const u32x4 r = *(const u32x4*)(idx + e);
        x[0] = q[0]; x[1] = q[1]; x[2] = q[2]; x[3] = q[3];
        ix[0] = r[0]; ix[1] = r[1]; ix[2] = r[2]; ix[3] = r[3];
        return 0xF;
    }
};

// The row's keys >= lo, read through the selected-subrange list (alpha > 0: slot e is key
// (sel[e >> alpha] << alpha) + (e & (2^alpha - 1))) or directly (alpha == 0: slot e is key e).
struct SrcV {
    const u32* row;
    u64 n;
    const u32* sel;
    int alpha;
    int f32;
    u32 lo;
    u64 E;  // slots (alpha == 1 reads the id list slot by slot and must stop at E)
    __device__ __forceinline__ u32 cvt(u32 b) const { return f32 ? KeyF32::to_ord(b) : b; }
    __device__ __forceinline__ u32 load4(u64 e, u32* x, u32* ix) const {
        u32 v = 0;
        if (alpha != 1) {
            const u64 base = alpha == 0 ? e : (((u64)sel[e >> alpha] << alpha) | (e & ((1ull << alpha) - 1)));
            if (base + 4 <= n) {
                const u32x4 q = *(const u32x4*)(row + base);
                x[0] = q[0]; x[1] = q[1]; x[2] = q[2]; x[3] = q[3];
                v = 0xF;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const bool ok = base + j < n;
                    x[j] = ok ? row[base + j] : 0u;
                    v |= (ok ? 1u : 0u) << j;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) ix[j] = (u32)(base + j);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const u64 ee = e + j;
                const u64 base = ee < E ? (((u64)sel[ee >> 1] << 1) | (ee & 1)) : n;
                const bool ok = base < n;
                x[j] = ok ? row[base] : 0u;
                ix[j] = (u32)base;
                v |= (ok ? 1u : 0u) << j;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            x[j] = cvt(x[j]);
            if (x[j] < lo) v &= ~(1u << j);
        }
        return v;
    }
    // alpha == 1 reads slot by slot and rows under 4 keys have no in-bounds granule to clamp
    // to: those take load4 throughout
    __device__ __forceinline__ u64 bulk_end(u64 E_, u64 step) const {
        return (alpha == 1 || n < 4) ? 0 : min(E_, E) / step * step;
    }
    // One 16-byte load from an address clamped into the row, so that no lane's load depends on a
    // per-lane condition (such a load is waited for where the branches join); only a granule
    // that runs past the row's end (its last subrange) is re-read key by key.
    __device__ __forceinline__ u32 load4_fast(u64 e, u32* x, u32* ix) const {
        const u64 base = alpha == 0 ? e : (((u64)sel[e >> alpha] << alpha) | (e & ((1ull << alpha) - 1)));
        const u64 bc = base + 4 <= n ? base : ((n - 4) & ~3ull);
        const u32x4 q = *(const u32x4*)(row + bc);
        x[0] = q[0]; x[1] = q[1]; x[2] = q[2]; x[3] = q[3];
        u32 v = 0xF;
        if (bc != base) {
            v = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool ok = base + j < n;
                x[j] = ok ? row[base + j] : 0u;
                v |= (ok ? 1u : 0u) << j;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            ix[j] = (u32)(base + j);
            x[j] = cvt(x[j]);
            if (x[j] < lo) v &= ~(1u << j);
        }
        return v;
    }
};

// Granules (4 slots) per thread in each step of the bulk loops below: BR_U 16-byte loads in
// flight per lane, and one block scan per BR_NT * 4 * BR_U slots in the ordered pass.
constexpr int BR_U = 4;
constexpr u64 BR_STEP = (u64)BR_NT * 4 * BR_U;

struct RowLds {
    u32 hist[BR_W * 256];
    u64 scan[BR_W + 1];
    u32 misc[16];
};

// One granule of a digit pass: its valid keys under the prefix counted in the wave's histogram
// (and, with gather, listed as candidates).
__device__ __forceinline__ void hist_step(const u32 (&x)[4], u32 v, u32 mask, u32 prefix, int shift, int R, u32* wh,
                                          bool gather, u32* cand, u32 cap, RowLds& L) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const bool act = ((v >> j) & 1u) && (x[j] & mask) == prefix;
        wave_hist_add(wh, act, (x[j] >> shift) & (R - 1));
        if (gather && act) {
            const u32 p = atomicAdd(&L.misc[0], 1u);
            if (p < cap) cand[p] = x[j];
        }
    }
}

// k-th largest (ascending rank m-k) of the valid slots of src by radix digits of nb bits,
// most significant first.  With `gather`, once the candidates of the current prefix fit in
// cand[0..cap) the next pass copies them there and the remaining digits run on LDS.
// Returns the key; *gt = valid slots strictly greater.
template <class Src>
__device__ __forceinline__ u32 wg_kth(const Src& src, u64 E, u64 k, int nb, bool gather_ok, u32* cand, u32 cap, RowLds& L,
                      u64* gt_out) {
    const int t = threadIdx.x, w = t / WAVE, lane = t & (WAVE - 1);
    const int R = 1 << nb;
    u32 prefix = 0, mask = 0;
    u64 kk = k, gt = 0;
    bool in_cand = false, gather = gather_ok && E <= cap;
    u32 mc = 0;
    if (t == 0) L.misc[0] = 0;  // gather counter
    for (int shift = 32 - nb; shift >= 0; shift -= nb) {
        for (int i = t; i < BR_W * R; i += BR_NT) L.hist[(i / R) * 256 + (i % R)] = 0;
        __syncthreads();
        u32* wh = L.hist + w * 256;
        if (in_cand) {
            for (u32 i0 = 0; i0 < mc; i0 += BR_NT) {
                const u32 i = i0 + t;
                const u32 x = i < mc ? cand[i] : 0u;
                wave_hist_add(wh, i < mc && (x & mask) == prefix, (x >> shift) & (R - 1));
            }
        } else {
            const u64 Eb = src.bulk_end(E, BR_STEP);
            for (u64 e0 = 0; e0 < Eb; e0 += BR_STEP) {  // whole steps: every load issued before use
                u32 x[BR_U][4], ix[BR_U][4], v[BR_U];
#pragma unroll
                for (int u = 0; u < BR_U; ++u) v[u] = src.load4_fast(e0 + ((u64)u * BR_NT + t) * 4, x[u], ix[u]);
#pragma unroll
                for (int u = 0; u < BR_U; ++u) hist_step(x[u], v[u], mask, prefix, shift, R, wh, gather, cand, cap, L);
            }
            for (u64 e0 = Eb; e0 < E; e0 += (u64)BR_NT * 4) {
                const u64 e = e0 + (u64)t * 4;
                u32 x[4], ix[4];
                const u32 v = e < E ? src.load4(e, x, ix) : 0u;
                hist_step(x, v, mask, prefix, shift, R, wh, gather, cand, cap, L);
            }
        }
        __syncthreads();
        if (gather) {
            in_cand = true;
            gather = false;
            mc = min(L.misc[0], cap);
        }
        if (t < R) {
            u32 s = 0;
            for (int q = 0; q < BR_W; ++q) s += L.hist[q * 256 + t];
            L.hist[t] = s;
        }
        __syncthreads();
        if (w == 0) {
            // bins from the top: lane l holds bins R-1-l*bpl ... R-(l+1)*bpl
            const int bpl = R >= WAVE ? R / WAVE : 1;
            u32 c[4] = {0, 0, 0, 0};
            u64 s = 0;
            for (int j = 0; j < bpl; ++j) {
                const int b = R - 1 - lane * bpl - j;
                c[j] = b >= 0 ? L.hist[b] : 0u;
                s += c[j];
            }
            const u64 inc = wave_incl_scan_u64(s);
            const u64 bal = __ballot(inc >= kk);
            const int fl = bal ? __ffsll((unsigned long long)bal) - 1 : 0;
            if (lane == fl) {
                u64 above = inc - s;
                int bin = R - 1 - lane * bpl - (bpl - 1);
                u32 cnt = c[bpl - 1];
                for (int j = 0; j < bpl; ++j) {
                    if (above + c[j] >= kk) {
                        bin = R - 1 - lane * bpl - j;
                        cnt = c[j];
                        break;
                    }
                    above += c[j];
                }
                L.misc[2] = (u32)bin;
                L.misc[3] = cnt;
                L.misc[4] = (u32)above;
                L.misc[5] = (u32)(above >> 32);
            }
        }
        __syncthreads();
        const u32 bin = L.misc[2], cnt = L.misc[3];
        const u64 above = (u64)L.misc[4] | ((u64)L.misc[5] << 32);
        prefix |= bin << shift;
        mask |= (u32)(R - 1) << shift;
        kk -= above;
        gt += above;
        if (!in_cand && gather_ok && cnt <= cap && shift > 0) gather = true;
    }
    *gt_out = gt;
    return prefix;
}

// Ordered pass: calls emit(pos, key, index) for every valid slot with key > K and for the
// first `ties` valid slots with key == K (ALL: every valid slot), pos = rank among those
// emitted, in slot order.  Returns the number of valid slots > K (ALL: valid slots).
// One step of wg_ordered: thread t holds granules t*U .. t*U+U-1 of the step, so one block scan
// ranks all of them in slot order.
template <bool ALL, int U, class F>
__device__ __forceinline__ void ordered_step(const u32 (&x)[U][4], const u32 (&ix)[U][4], const u32 (&v)[U], u32 K,
                                             u64 ties, RowLds& L, u64& gt_run, u64& eq_run, F& emit) {
    u32 g = 0, q = 0;
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool ok = (v[u] >> j) & 1u;
            g += (ok && (ALL || x[u][j] > K)) ? 1u : 0u;
            q += (!ALL && ok && x[u][j] == K) ? 1u : 0u;
        }
    u64 tot;
    const u64 ex = block_excl_scan_u64<BR_NT>(((u64)q << 32) | g, L.scan, &tot);
    u64 gb = gt_run + (u32)ex, qb = eq_run + (ex >> 32);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (!((v[u] >> j) & 1u)) continue;
            if (ALL || x[u][j] > K) {
                emit(gb + min(qb, ties), x[u][j], ix[u][j]);
                ++gb;
            } else if (x[u][j] == K) {
                if (qb < ties) emit(gb + qb, x[u][j], ix[u][j]);
                ++qb;
            }
        }
    gt_run += (u32)tot;
    eq_run += tot >> 32;
}

template <bool ALL, class Src, class F>
__device__ __forceinline__ u64 wg_ordered(const Src& src, u64 E, u32 K, u64 ties, RowLds& L, F emit) {
    const int t = threadIdx.x;
    u64 gt_run = 0, eq_run = 0;
    const u64 Eb = src.bulk_end(E, BR_STEP);
    for (u64 e0 = 0; e0 < Eb; e0 += BR_STEP) {
        u32 x[BR_U][4], ix[BR_U][4], v[BR_U];
#pragma unroll
        for (int u = 0; u < BR_U; ++u) v[u] = src.load4_fast(e0 + ((u64)t * BR_U + u) * 4, x[u], ix[u]);
        ordered_step<ALL, BR_U>(x, ix, v, K, ties, L, gt_run, eq_run, emit);
    }
    for (u64 e0 = Eb; e0 < E; e0 += (u64)BR_NT * 4) {
        const u64 e = e0 + (u64)t * 4;
        u32 x[1][4], ix[1][4], v[1];
        v[0] = e < E ? src.load4(e, x[0], ix[0]) : 0u;
        ordered_step<ALL, 1>(x, ix, v, K, ties, L, gt_run, eq_run, emit);
    }
    return gt_run;
}

// One row of k_brows (its index q in the processing order).  SPILL (alpha > 0 and k larger than
// the LDS concatenation cache, so the filtered concatenation of C >= k keys cannot fit it): the
// keys that overflow the cache go to the row's workspace region after its selected ids, then the
// cached part, and the second top-k streams those C keys instead of gathering the k << alpha
// slots through the id list twice more.  Rows with k within the cache keep the LDS path (and,
// when the concatenation does overflow, the re-gather), without the extra code in their loops.
template <bool SPILL>
__device__ __forceinline__ void brow(const BatchRow* rp, u32 q, const u32* __restrict__ keys, BatchWs* wsh,
                                     const u32* __restrict__ deleg, u32* __restrict__ selws,
                                     u32* __restrict__ kth_out, u32* __restrict__ out_keys, i64* __restrict__ out_idx,
                                     RowLds& L, u32* reg) {
    const int t = threadIdx.x;
    const u64 n = rp->n, k = rp->k, out_off = rp->out_off;
    const int a = rp->alpha, nb = rp->nb, f32 = rp->f32;
    const u32* row = keys + rp->key_off;
    u32* sel = selws + rp->sel_off;

    // first top-k over the delegates and range selection
    u32 T = 0;
    u64 E = n;
    if (a > 0) {
        const SrcG dg{deleg + rp->deleg_off, rp->S};
        u64 g = 0;
        T = wg_kth(dg, rp->S, k, nb, true, reg, REG_WORDS, L, &g);
        const u64 ties = k - g;
        wg_ordered<false>(dg, rp->S, T, ties, L, [&](u64 p, u32, u32 id) {
            if (p < k) sel[p] = id;
        });
        __syncthreads();
        E = k << a;
    }
    // filtered concatenation, cached in LDS when it fits
    const SrcV vs{row, n, sel, a, f32, T, E};
    u32* ckey = reg;
    u32* cidx = reg + CONCAT_CAP;
    u32* gck = sel + ((k + 3) & ~3ull);  // SPILL: room for E keys, then E indices
    u32* gci = gck + ((E + 3) & ~3ull);
    const u64 C = wg_ordered<true>(vs, E, 0u, 0, L, [&](u64 p, u32 x, u32 ix) {
        if (p < CONCAT_CAP) {
            ckey[p] = x;
            cidx[p] = ix;
        } else if (SPILL) {  // p < C <= E
            gck[p] = x;
            gci[p] = ix;
        }
    });
    __syncthreads();
    // every key of the top-k set is >= T, so the concatenation holds at least k keys
    if (C < k) {
        if (t == 0) ws_fail(&wsh->head, ERR_CONCAT_SHORT);
        return;
    }
    // second top-k and extraction
    u64 gt = 0;
    u32 K;
    auto emit_out = [&](u64 p, u32 x, u32 ix) {
        if (p < k) {
            if (out_keys) out_keys[out_off + p] = f32 ? KeyF32::from_ord(x) : x;
            if (out_idx) out_idx[out_off + p] = (i64)ix;
        }
    };
    if (C <= CONCAT_CAP) {
        const SrcL ls{ckey, cidx, C};
        K = wg_kth(ls, C, k, nb, false, nullptr, 0, L, &gt);
        wg_ordered<false>(ls, C, K, k - gt, L, emit_out);
    } else if (SPILL) {
        for (u32 i = t; i < CONCAT_CAP; i += BR_NT) {
            gck[i] = ckey[i];
            gci[i] = cidx[i];
        }
        __syncthreads();  // the LDS cache is free again: it takes the digit passes' candidates
        const SrcL gs{gck, gci, C};
        K = wg_kth(gs, C, k, nb, true, reg, REG_WORDS, L, &gt);
        wg_ordered<false>(gs, C, K, k - gt, L, emit_out);
    } else {
        K = wg_kth(vs, E, k, nb, true, reg, REG_WORDS, L, &gt);
        wg_ordered<false>(vs, E, K, k - gt, L, emit_out);
    }
    if (t == 0 && kth_out) kth_out[q] = f32 ? KeyF32::from_ord(K) : K;
}

// The plan's rule for a row's concatenation region (it reserves the room).
__host__ __device__ __forceinline__ bool row_spills(int alpha, u64 k) { return alpha > 0 && k > CONCAT_CAP; }

// Four waves per SIMD (two workgroups per CU): the row body needs a few more registers than that
// allows unbounded, and halving the resident rows costs more than the bound does.
__global__ __launch_bounds__(BR_NT, 4) void k_brows(const unsigned char* __restrict__ plan, const u32* __restrict__ keys,
                                                    BatchWs* wsh, const u32* __restrict__ deleg,
                                                    u32* __restrict__ selws, u32* __restrict__ kth_out,
                                                    u32* __restrict__ out_keys, i64* __restrict__ out_idx) {
    const BatchHeader* h = (const BatchHeader*)plan;
    const BatchRow* rows = (const BatchRow*)(plan + h->o_rows);
    const u32* order = (const u32*)(plan + h->o_order);
    const u64 nrows = h->nrows;
    __shared__ RowLds L;
    __shared__ __align__(16) u32 reg[REG_WORDS];
    const int t = threadIdx.x;
    for (;;) {
        __syncthreads();
        if (t == 0) L.misc[15] = atomicAdd(&wsh->queue, 1u);
        __syncthreads();
        const u32 q = L.misc[15];
        if (q >= nrows) break;
        const BatchRow* rp = rows + order[q];
        if (row_spills(rp->alpha, rp->k))
            brow<true>(rp, order[q], keys, wsh, deleg, selws, kth_out, out_keys, out_idx, L, reg);
        else
            brow<false>(rp, order[q], keys, wsh, deleg, selws, kth_out, out_keys, out_idx, L, reg);
    }
}

// ---------------------------------------------------------------- tuning
#include "tuning_table.inc"

// The built-in table: cells measured by the autotuner on MI355X; every other cell keeps the
// default choice below.
void default_tuning(drtopk_tuning* t) {
    std::memset(t, DRTOPK_TUNE_DEFAULT, sizeof(*t));
    for (const auto& c : kTunedCells) {
        if (c[0] > 1 || c[1] >= 32 || c[2] >= 32) continue;
        t->alpha[c[0]][c[1]][c[2]] = c[3];
        t->digit_bits[c[0]][c[1]][c[2]] = c[4];
        t->tile_log2[c[0]][c[1]][c[2]] = c[5];
    }
}

// The built-in table, built once.
const drtopk_tuning& builtin_tuning() {
    static const drtopk_tuning t = [] {
        drtopk_tuning x;
        default_tuning(&x);
        return x;
    }();
    return t;
}

// Default choice for a cell the table leaves open: the reference's alpha policy
// (baseline+filter+beta+shuffle/main.cu:83-89), 8-bit digits, 8192-key tiles.
void choose(const drtopk_tuning& tu, u64 n, u64 k, int f32, int* alpha, int* nb, int* tile) {
    const int ln = ceil_log2(n), lk = ceil_log2(k);
    const u8 ta = tu.alpha[f32][ln][lk], tn = tu.digit_bits[f32][ln][lk], tt = tu.tile_log2[f32][ln][lk];
    int a = ta != DRTOPK_TUNE_DEFAULT ? (int)ta : std::max(0, (floor_log2(n) + 3 - lk) / 2);
    int tl = tt != DRTOPK_TUNE_DEFAULT ? (int)tt : 13;
    tl = std::min(MAX_TILE_LOG2, std::max(10, tl));
    a = std::min(a, tl);
    while (a > 0 && cdiv(n, 1ull << a) < k) --a;
    *alpha = a;
    *nb = (tn == 4 || tn == 8) ? (int)tn : 8;
    *tile = tl;
}

}  // namespace
}  // namespace drtopk

using namespace drtopk;

extern "C" {

void drtopk_tuning_default(drtopk_tuning* t) {
    if (t) default_tuning(t);
}

drtopk_status drtopk_batch_plan(const drtopk_row* rows, uint64_t nrows, uint64_t keys_len, uint64_t out_len,
                                const drtopk_tuning* tuning, void* h_plan, size_t* plan_bytes, size_t* ws_bytes) {
    clear_error();
    if (!rows || !plan_bytes) return fail(DRTOPK_ERR_INVALID_ARG, "rows and plan_bytes must not be NULL");
    if (nrows == 0 || nrows > (1ull << 24)) return fail(DRTOPK_ERR_INVALID_ARG, "nrows %llu not in [1, 2^24]", (unsigned long long)nrows);
    if (keys_len >= (1ull << 48)) return fail(DRTOPK_ERR_UNSUPPORTED, "keys_len %llu >= 2^48", (unsigned long long)keys_len);
    // a caller's table is checked on every call; the built-in one is built once and needs no check
    // (a server sizes every request with this call, so its fixed cost counts)
    const drtopk_tuning& tu = tuning ? *tuning : builtin_tuning();
    if (tuning)
        for (int f = 0; f < 2; ++f)
            for (int i = 0; i < 32; ++i)
                for (int j = 0; j < 32; ++j) {
                    const u8 d = tu.digit_bits[f][i][j], tl = tu.tile_log2[f][i][j], al = tu.alpha[f][i][j];
                    if (d != DRTOPK_TUNE_DEFAULT && d != 4 && d != 8)
                        return fail(DRTOPK_ERR_INVALID_ARG, "tuning digit_bits[%d][%d][%d] = %d not 4 or 8", f, i, j,
                                    d);
                    if (tl != DRTOPK_TUNE_DEFAULT && (tl < 10 || tl > MAX_TILE_LOG2))
                        return fail(DRTOPK_ERR_INVALID_ARG, "tuning tile_log2[%d][%d][%d] = %d not in [10, 15]", f, i,
                                    j, tl);
                    if (al != DRTOPK_TUNE_DEFAULT && al > 30)
                        return fail(DRTOPK_ERR_INVALID_ARG, "tuning alpha[%d][%d][%d] = %d not in [0, 30]", f, i, j,
                                    al);
                }
    std::vector<BatchRow> br(nrows);
    std::vector<u64> cost(nrows);
    u64 nch = 0, dl = 0, sl = 0;
    for (u64 r = 0; r < nrows; ++r) {
        const drtopk_row& in = rows[r];
        if (in.dtype != DRTOPK_U32 && in.dtype != DRTOPK_F32)
            return fail(DRTOPK_ERR_INVALID_ARG, "row %llu: bad dtype %d", (unsigned long long)r, in.dtype);
        if (in.reserved != 0) return fail(DRTOPK_ERR_INVALID_ARG, "row %llu: reserved must be 0", (unsigned long long)r);
        if (in.n == 0 || in.n > (1ull << 31) - 1)
            return fail(DRTOPK_ERR_INVALID_ARG, "row %llu: n = %llu not in [1, 2^31-1]", (unsigned long long)r,
                        (unsigned long long)in.n);
        if (in.k == 0 || in.k > in.n)
            return fail(DRTOPK_ERR_INVALID_ARG, "row %llu: k = %llu not in [1, n = %llu]", (unsigned long long)r,
                        (unsigned long long)in.k, (unsigned long long)in.n);
        if (in.offset % 4 != 0)
            return fail(DRTOPK_ERR_INVALID_ARG, "row %llu: offset %llu is not a multiple of 4 keys", (unsigned long long)r,
                        (unsigned long long)in.offset);
        if (in.offset > keys_len || in.n > keys_len - in.offset)
            return fail(DRTOPK_ERR_INVALID_ARG, "row %llu: keys [%llu, +%llu) outside keys_len %llu", (unsigned long long)r,
                        (unsigned long long)in.offset, (unsigned long long)in.n, (unsigned long long)keys_len);
        if (in.out_offset > out_len || in.k > out_len - in.out_offset)
            return fail(DRTOPK_ERR_INVALID_ARG, "row %llu: outputs [%llu, +%llu) outside out_len %llu", (unsigned long long)r,
                        (unsigned long long)in.out_offset, (unsigned long long)in.k, (unsigned long long)out_len);
        int a, nb, tl;
        choose(tu, in.n, in.k, in.dtype == DRTOPK_F32 ? 1 : 0, &a, &nb, &tl);
        BatchRow& b = br[r];
        std::memset(&b, 0, sizeof(b));
        b.key_off = in.offset;
        b.n = in.n;
        b.k = in.k;
        b.out_off = in.out_offset;
        b.alpha = (u8)a;
        b.f32 = in.dtype == DRTOPK_F32;
        b.nb = (u8)nb;
        b.tile_log2 = (u8)tl;
        if (a > 0) {
            const u64 S = cdiv(in.n, 1ull << a), ch = cdiv(in.n, 1ull << tl);
            if (nch + ch > 0xFFFFFFFFull) return fail(DRTOPK_ERR_UNSUPPORTED, "batch has more than 2^32-1 tiles");
            b.S = (u32)S;
            b.chunk_begin = (u32)nch;
            b.nchunks = (u32)ch;
            b.deleg_off = dl;
            b.sel_off = sl;
            nch += ch;
            dl += align_up(S, 4);
            sl += align_up(in.k, 4);
            if (row_spills(a, in.k)) sl += 2 * align_up(in.k << a, 4);
            cost[r] = 4 * S + (in.k << a);
        } else {
            cost[r] = 4 * in.n;
        }
    }
    if (dl >= (1ull << 40)) return fail(DRTOPK_ERR_UNSUPPORTED, "batch has 2^40 or more delegates");
    BatchHeader hd;
    std::memset(&hd, 0, sizeof(hd));
    hd.magic = PLAN_MAGIC;
    hd.nrows = nrows;
    hd.nchunks = nch;
    u64 o = sizeof(BatchHeader);
    hd.o_rows = o; o = align_up(o + sizeof(BatchRow) * nrows, 16);
    hd.o_tiles = o; o = align_up(o + sizeof(TileDesc) * nch, 16);
    hd.o_order = o; o = align_up(o + 4 * nrows, 16);
    hd.plan_bytes = o;
    u64 w = 256;
    hd.o_ws_deleg = w; w = align_up(w + 4 * std::max<u64>(dl, 4), 256);
    hd.o_ws_sel = w; w = align_up(w + 4 * std::max<u64>(sl, 4), 256);
    hd.ws_bytes = w;
    hd.keys_len = keys_len;
    hd.out_len = out_len;
    // 4096 workgroups (2.7 rounds of the 1536 resident ones): 1% faster than exactly one round
    hd.sample_grid = (u32)std::max<u64>(1, std::min<u64>(nch, 4096));
    hd.rows_grid = (u32)std::min<u64>(nrows, 512);
    if (ws_bytes) *ws_bytes = (size_t)hd.ws_bytes;
    if (!h_plan) {
        *plan_bytes = (size_t)hd.plan_bytes;
        return DRTOPK_OK;
    }
    if (*plan_bytes < hd.plan_bytes)
        return fail(DRTOPK_ERR_WORKSPACE_TOO_SMALL, "plan buffer %zu bytes < %llu", *plan_bytes,
                    (unsigned long long)hd.plan_bytes);
    *plan_bytes = (size_t)hd.plan_bytes;
    unsigned char* P = (unsigned char*)h_plan;
    std::memset(P, 0, hd.plan_bytes);
    std::memcpy(P, &hd, sizeof(hd));
    std::memcpy(P + hd.o_rows, br.data(), sizeof(BatchRow) * nrows);
    TileDesc* tds = (TileDesc*)(P + hd.o_tiles);
    for (u64 r = 0; r < nrows; ++r) {
        const BatchRow& b = br[r];
        for (u32 c = 0; c < b.nchunks; ++c) {
            const u64 start = (u64)c << b.tile_log2;
            const u64 len = std::min<u64>(1ull << b.tile_log2, b.n - start);
            TileDesc& d = tds[b.chunk_begin + c];
            d.src = (b.key_off + start) | ((u64)b.alpha << 48) | ((u64)b.f32 << 54);
            d.dst = (b.deleg_off + (start >> b.alpha)) | ((len - 1) << 40);
        }
    }
    std::vector<u32> ord(nrows);
    std::iota(ord.begin(), ord.end(), 0u);
    std::stable_sort(ord.begin(), ord.end(), [&](u32 x, u32 y) { return cost[x] > cost[y]; });
    std::memcpy(P + hd.o_order, ord.data(), 4 * nrows);
    return DRTOPK_OK;
}

drtopk_status drtopk_select_batched(const void* h_plan, const void* d_plan, const void* d_keys, void* d_workspace,
                                    size_t workspace_bytes, void* d_kth, void* d_topk_keys, int64_t* d_topk_idx,
                                    void* stream) {
    clear_error();
    if (!h_plan || !d_plan) return fail(DRTOPK_ERR_INVALID_ARG, "h_plan and d_plan must not be NULL");
    BatchHeader hd;
    std::memcpy(&hd, h_plan, sizeof(hd));
    if (hd.magic != PLAN_MAGIC) return fail(DRTOPK_ERR_INVALID_ARG, "h_plan is not a drtopk_batch_plan blob");
    if (!d_keys) return fail(DRTOPK_ERR_INVALID_ARG, "d_keys is NULL");
    if ((uintptr_t)d_keys % 16) return fail(DRTOPK_ERR_INVALID_ARG, "d_keys must be 16-byte aligned");
    if ((uintptr_t)d_plan % 16) return fail(DRTOPK_ERR_INVALID_ARG, "d_plan must be 16-byte aligned");
    if (!d_workspace || (uintptr_t)d_workspace % 256)
        return fail(DRTOPK_ERR_INVALID_ARG, "d_workspace must be a 256-byte aligned device pointer");
    if (workspace_bytes < hd.ws_bytes)
        return fail(DRTOPK_ERR_WORKSPACE_TOO_SMALL, "workspace %zu bytes < %llu", workspace_bytes,
                    (unsigned long long)hd.ws_bytes);
    if ((d_topk_keys && (uintptr_t)d_topk_keys % 4) || (d_topk_idx && (uintptr_t)d_topk_idx % 8) ||
        (d_kth && (uintptr_t)d_kth % 4))
        return fail(DRTOPK_ERR_INVALID_ARG, "misaligned output pointer");
    hipStream_t st = (hipStream_t)stream;
    unsigned char* ws = (unsigned char*)d_workspace;
    BatchWs* wsh = (BatchWs*)ws;
    u32* deleg = (u32*)(ws + hd.o_ws_deleg);
    u32* selws = (u32*)(ws + hd.o_ws_sel);
    const unsigned char* P = (const unsigned char*)d_plan;
    hipLaunchKernelGGL(k_bsample, dim3(hd.sample_grid), dim3(BS_NT), 0, st, P, (const u32*)d_keys, wsh, deleg);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_brows, dim3(hd.rows_grid), dim3(BR_NT), 0, st, P, (const u32*)d_keys, wsh,
                       (const u32*)deleg, selws, (u32*)d_kth, (u32*)d_topk_keys, (i64*)d_topk_idx);
    HIPCHK(hipGetLastError());
    return DRTOPK_OK;
}

}  // extern "C"
