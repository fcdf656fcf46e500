// sample.hip -- delegate pass: the only kernel that reads all of V.
//
// Restates sampleMax (baseline+filter+beta+shuffle/radixselect.cuh:883-918) and
// sampleMax_NoReduction_Beta2/3/4 (radixselect.cuh:967-1202): for every subrange of
// 2^alpha keys keep its maximum (beta == 1) or its top-beta multiset (beta 2..4),
// largest first, lowest index first among equal keys.
//
// Three variants, identical output, all persistent (grid-stride over work units):
//  - lane groups (alpha <= 8): a workgroup streams 8192-key chunks with 16-byte loads,
//    lane l of a wave holding keys 4l..4l+3 of a 256-key wave tile; a subrange is a group
//    of 2^alpha/4 lanes reduced with xor shuffles (2 subranges per lane at alpha == 1).
//  - wave units (alpha >= 9): a wave streams max(2^alpha, 2048) keys with 16-byte loads,
//    eight in flight per lane, and reduces each subrange over its 64 lanes.
//  - LDS staged, thread per subrange (alpha <= 5, DRTOPK_FLAG_SHUFFLE): the workgroup
//    stages 256 subranges in LDS with coalesced loads, then each thread scans its own
//    subrange with 16-byte LDS reads (the reference's "shuffle" option).
//
// Fused into the same pass: the histogram of the delegates' top digit (radix family,
// the first digit of the first top-k) or the delegates' min/max (bucket family).  The
// histogram is aggregated per wave (the most common bin of the wave is counted with one
// LDS atomic) and flushed once per workgroup into one of HIST_REPS global replicas, so
// skewed digits do not serialise on a single L2 address.
#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace drtopk {

// ---------------------------------------------------------------- top-beta lists
// E = u32: plain ordered keys (multiset of values, enough when positions are not needed).
// E = u64: composite (ordered key << 32 | ~index): larger key wins, then lower index.
// Absent = 0 (smaller than any composite of a real key, and equal to the pad value 0).
template <int BETA, class E>
struct TopList {
    E v[BETA];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int j = 0; j < BETA; ++j) v[j] = 0;
    }
    __device__ __forceinline__ void insert(u32 ord, u64 position) { insert_e(elem(ord, position)); }
    static __device__ __forceinline__ E elem(u32 ord, u64 position) {
        if (sizeof(E) == 8) return (E)(((u64)ord << 32) | (u64)(0xFFFFFFFFu - (u32)position));
        return (E)ord;
    }
    __device__ __forceinline__ void insert_e(E x) {
#pragma unroll
        for (int j = 0; j < BETA; ++j) {
            E hi = v[j] > x ? v[j] : x;
            E lo = v[j] > x ? x : v[j];
            v[j] = hi;
            x = lo;
        }
    }
    // top-BETA of two descending lists: max(a[i], b[BETA-1-i]) is bitonic and holds the
    // top BETA; a small sorting network orders it.
    __device__ __forceinline__ void merge(const E* b) {
        E c[BETA];
#pragma unroll
        for (int j = 0; j < BETA; ++j) c[j] = v[j] > b[BETA - 1 - j] ? v[j] : b[BETA - 1 - j];
#pragma unroll
        for (int p = 0; p < BETA; ++p)
#pragma unroll
            for (int q = 0; q + 1 < BETA - p; ++q) {
                E x = c[q], y = c[q + 1];
                c[q] = x > y ? x : y;
                c[q + 1] = x > y ? y : x;
            }
#pragma unroll
        for (int j = 0; j < BETA; ++j) v[j] = c[j];
    }
    __device__ __forceinline__ void shfl_merge(int off) {
        E b[BETA];
#pragma unroll
        for (int j = 0; j < BETA; ++j) {
            if (sizeof(E) == 8) b[j] = (E)shfl_xor_u64((u64)v[j], off);
            else b[j] = (E)(u32)__shfl_xor((int)(u32)v[j], off, WAVE);
        }
        merge(b);
    }
    __device__ __forceinline__ u32 val(int j) const { return sizeof(E) == 8 ? (u32)((u64)v[j] >> 32) : (u32)v[j]; }
    __device__ __forceinline__ u32 pos(int j) const {
        return (sizeof(E) == 8 && v[j] != 0) ? 0xFFFFFFFFu - (u32)(u64)v[j] : 0xFFFFFFFFu;
    }
};

// Top-2 by value with the index of the first maximum (the beta == 2 path with offsets):
// m1 = max, m2 = second largest (multiset), p1 = lowest index holding m1.  Two u32 min/max
// per key instead of the u64 composite network.
struct Top2P {
    u32 m1, m2, p1;
    __device__ __forceinline__ void clear() { m1 = 0; m2 = 0; p1 = 0xFFFFFFFFu; }
    // keys arrive in increasing index order within a lane
    __device__ __forceinline__ void insert(u32 x, u64 position) {
        const bool up = x > m1 || p1 == 0xFFFFFFFFu;
        m2 = max(min(m1, x), m2);
        if (p1 == 0xFFFFFFFFu) m2 = 0;  // first key: no second yet
        p1 = up ? (u32)position : p1;
        m1 = max(m1, x);
    }
    __device__ __forceinline__ void shfl_merge(int off) {
        const u32 bm1 = (u32)__shfl_xor((int)m1, off, WAVE);
        const u32 bm2 = (u32)__shfl_xor((int)m2, off, WAVE);
        const u32 bp1 = (u32)__shfl_xor((int)p1, off, WAVE);
        const u32 n2 = max(min(m1, bm1), max(m2, bm2));
        const bool empty_a = p1 == 0xFFFFFFFFu, empty_b = bp1 == 0xFFFFFFFFu;
        u32 np;
        if (empty_a) np = bp1;
        else if (empty_b) np = p1;
        else np = m1 > bm1 ? p1 : (bm1 > m1 ? bp1 : min(p1, bp1));
        // an empty side contributes nothing: the other side's pair stands
        const u32 nm2 = empty_a ? bm2 : (empty_b ? m2 : n2);
        m1 = max(m1, bm1);
        m2 = nm2;
        p1 = np;
    }
    __device__ __forceinline__ u32 val(int j) const { return j == 0 ? m1 : m2; }
    __device__ __forceinline__ u32 pos(int j) const { return j == 0 ? p1 : 0xFFFFFFFFu; }
};

template <int BETA, int POSM>
struct ListOf {
    using type = TopList<BETA, typename std::conditional<(POSM != 0), u64, u32>::type>;
};
template <>
struct ListOf<2, 2> {
    using type = Top2P;
};

// Leader lanes store subrange s's delegates, positions or packed offsets.
// POSM: 0 none, 1 absolute positions (u32 per delegate), 2 packed offsets of the first
// beta-1 delegates inside the subrange (8 bits each; u8 / u16 / u32 per subrange).
template <int BETA, int POSM, class LT>
__device__ __forceinline__ void store_delegates(const LT& t, u64 s, int alpha, u32* dvals, void* dpos, u8* dtop) {
    if constexpr (std::is_same<LT, Top2P>::value) {
        // both delegates with one 8-byte store (or the first and the second's top byte), the
        // first maximum's offset as one byte
        const u32 off = t.p1 == 0xFFFFFFFFu ? 0xFFu : (u32)(t.p1 - (u32)(s << alpha)) & 0xFFu;
        if (dtop) {
            // compact form: the second delegate's top byte and the offset in one 16-bit word
            dvals[s] = t.m1;
            ((uint16_t*)dtop)[s] = (uint16_t)((t.m2 >> 24) | (off << 8));
        } else {
            *(uint2*)(dvals + 2 * s) = make_uint2(t.m1, t.m2);
            ((uint8_t*)dpos)[s] = (uint8_t)off;
        }
        return;
    }
    if (BETA == 2 && POSM == 0 && dtop) {
        dvals[s] = t.val(0);
        ((uint16_t*)dtop)[s] = (uint16_t)((t.val(1) >> 24) | 0xFF00u);
        return;
    }
    u32 packed = 0;
#pragma unroll
    for (int j = 0; j < BETA; ++j) {
        dvals[s * BETA + j] = t.val(j);
        if (POSM == 1) ((u32*)dpos)[s * BETA + j] = t.pos(j);
        if (POSM == 2 && j < BETA - 1) {
            const u32 p = t.pos(j);
            const u32 off = p == 0xFFFFFFFFu ? 0xFFu : (u32)(p - (s << alpha)) & 0xFFu;
            packed |= off << (8 * j);
        }
    }
    if (POSM == 2) {
        if (BETA == 2) ((uint8_t*)dpos)[s] = (uint8_t)packed;
        else if (BETA == 3) ((uint16_t*)dpos)[s] = (uint16_t)packed;
        else if (BETA == 4) ((u32*)dpos)[s] = packed;
    }
}

template <int BETA, int FUSE, class LT>
__device__ __forceinline__ void fuse_add(const LT& t, bool leader, u32* lhist, int hshift, u32* mn,
                                         u32* mx) {
#pragma unroll
    for (int j = 0; j < BETA; ++j) {
        const u32 val = t.val(j);
        if (FUSE == 1) wave_hist_add(lhist, leader, val >> hshift);
        if (FUSE == 2 && leader) { *mn = min(*mn, val); *mx = max(*mx, val); }
    }
}

template <int FUSE>
__device__ __forceinline__ void fuse_init(u32* lhist, int bins) {
    if (FUSE == 1) {
        for (int i = threadIdx.x; i < bins; i += blockDim.x) lhist[i] = 0;
        __syncthreads();
    }
}

template <int FUSE>
__device__ __forceinline__ void fuse_flush(u32* lhist, u32* ghist, int bins, u32 mn, u32 mx, Ctrl* ctrl, u32* red) {
    if (FUSE == 1) {
        __syncthreads();
        u32* g = ghist + (size_t)(blockIdx.x & (HIST_REPS - 1)) * bins;
        for (int i = threadIdx.x; i < bins; i += blockDim.x) {
            u32 c = lhist[i];
            if (c) atomicAdd(&g[i], c);
        }
    }
    if (FUSE == 2) {
        mn = wave_min_u32(mn);
        mx = wave_max_u32(mx);
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        if (lane == 0) { red[w] = mn; red[8 + w] = mx; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 1; i < (int)(blockDim.x >> 6); ++i) { mn = min(mn, red[i]); mx = max(mx, red[8 + i]); }
            mn = min(mn, red[0]);
            mx = max(mx, red[8]);
            atomicMin(&ctrl->first.lo, mn);
            atomicMax(&ctrl->first.hi, mx);
        }
    }
}

__device__ __forceinline__ u32x4 load4(const u32x4* p) { return __builtin_nontemporal_load(p); }

// ---------------------------------------------------------------- lane groups, alpha 1..8
constexpr int GROUP_BATCHES = 8;                      // 16-byte loads in flight per lane
constexpr int GROUP_CHUNK = GROUP_BATCHES * WG * 4;   // 8192 keys per workgroup chunk

// Maximum over each aligned group of g lanes (g = 1, 2, .., 64, uniform) with DPP moves, complete
// in the group's last lane: quad permutes and the half-row / row mirrors give every lane of a
// 16-lane row its row's maximum, row broadcast 15 then carries row 0 into row 1 and row 2 into
// row 3 (32-lane groups), and row broadcast 31 row 1 into rows 2 and 3 (the whole wave).
__device__ __forceinline__ u32 group_max_last(u32 m, int g) {
    if (g >= 2) m = max(m, (u32)__builtin_amdgcn_mov_dpp((int)m, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    if (g >= 4) m = max(m, (u32)__builtin_amdgcn_mov_dpp((int)m, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    if (g >= 8) m = max(m, (u32)__builtin_amdgcn_mov_dpp((int)m, 0x141, 0xF, 0xF, false));  // row_half_mirror
    if (g >= 16) m = max(m, (u32)__builtin_amdgcn_mov_dpp((int)m, 0x140, 0xF, 0xF, false)); // row_mirror
    if (g >= 32) m = max(m, (u32)__builtin_amdgcn_update_dpp(0, (int)m, 0x142, 0xA, 0xF, false));  // row_bcast:15
    if (g >= 64) m = max(m, (u32)__builtin_amdgcn_update_dpp(0, (int)m, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return m;
}

template <class K, int BETA, int POSM, int FUSE, class LT>
__global__ __launch_bounds__(WG) void k_sample_groups(const u32* __restrict__ keys, u64 n, int alpha, int vec_ok,
                                                      u64 nchunks, u32* __restrict__ dvals, void* __restrict__ dpos,
                                                      u32* __restrict__ ghist, int hshift, int bins, Ctrl* ctrl) {
    __shared__ u32 lhist[FUSE == 1 ? 256 : 1];
    __shared__ u32 red[16];
    fuse_init<FUSE>(lhist, bins);
    u32 mn = 0xFFFFFFFFu, mx = 0u;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const u64 S = (n + (1ull << alpha) - 1) >> alpha;
    for (u64 chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const u64 c0 = chunk * GROUP_CHUNK;
        u32x4 raw[GROUP_BATCHES];
        const bool full = vec_ok && c0 + GROUP_CHUNK <= n;
        const u32x4* base4 = (const u32x4*)(keys + c0);
#pragma unroll
        for (int b = 0; b < GROUP_BATCHES; ++b) {
            const int q = b * WG + w * 64 + lane;  // uint4 index inside the chunk
            if (full) raw[b] = load4(base4 + q);
            else {
                const u64 i0 = c0 + 4ull * q;
                u32x4 r = {0, 0, 0, 0};
                if (i0 + 0 < n) r.x = keys[i0 + 0];
                if (i0 + 1 < n) r.y = keys[i0 + 1];
                if (i0 + 2 < n) r.z = keys[i0 + 2];
                if (i0 + 3 < n) r.w = keys[i0 + 3];
                raw[b] = r;
            }
        }
#pragma unroll
        for (int b = 0; b < GROUP_BATCHES; ++b) {
            const u64 i0 = c0 + 4ull * (b * WG + w * 64 + lane);
            const u32 kv[4] = {K::to_ord(raw[b].x), K::to_ord(raw[b].y), K::to_ord(raw[b].z), K::to_ord(raw[b].w)};
            if (alpha == 1) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    LT t;
                    t.clear();
#pragma unroll
                    for (int c = 0; c < 2; ++c)
                        if (i0 + 2 * h + c < n) t.insert(kv[2 * h + c], i0 + 2 * h + c);
                    const bool leader = i0 + 2 * h < n;
                    if (leader) store_delegates<BETA, POSM, LT>(t, (i0 >> 1) + h, alpha, dvals, dpos, nullptr);
                    fuse_add<BETA, FUSE, LT>(t, leader, lhist, hshift, &mn, &mx);
                }
            } else {
                LT t;
                t.clear();
                if (full) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) t.insert(kv[c], i0 + c);
                } else {
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if (i0 + c < n) t.insert(kv[c], i0 + c);
                }
                const int g = 1 << (alpha - 2);  // lanes per subrange (1..64)
                const u64 s = i0 >> alpha;
                if (BETA == 1 && POSM == 0 && full) {
                    // one maximum per subrange: DPP merge, stored by the group's last lane
                    LT r;
                    r.clear();
                    r.insert(group_max_last(t.val(0), g), 0);
                    const bool leader = (lane & (g - 1)) == g - 1;
                    if (leader) store_delegates<BETA, POSM, LT>(r, s, alpha, dvals, dpos, nullptr);
                    fuse_add<BETA, FUSE, LT>(r, leader, lhist, hshift, &mn, &mx);
                } else {
                    for (int off = g >> 1; off > 0; off >>= 1) t.shfl_merge(off);
                    const bool leader = (lane & (g - 1)) == 0 && i0 < n && s < S;
                    if (leader) store_delegates<BETA, POSM, LT>(t, s, alpha, dvals, dpos, nullptr);
                    fuse_add<BETA, FUSE, LT>(t, leader, lhist, hshift, &mn, &mx);
                }
            }
        }
    }
    fuse_flush<FUSE>(lhist, ghist, bins, mn, mx, ctrl, red);
}

// A whole wave's subrange s (every lane holds part of it): merge and store.  One maximum without
// positions merges with DPP moves and is stored by lane 63; other lists merge by lane shuffles.
template <int BETA, int POSM, int FUSE, class LT>
__device__ __forceinline__ void wave_finish(LT& t, u64 s, int alpha, u32* dvals, void* dpos, u32* lhist, int hshift,
                                            u32* mn, u32* mx) {
    const int lane = threadIdx.x & (WAVE - 1);
    if constexpr (BETA == 1 && POSM == 0) {
        LT r;
        r.clear();
        r.insert(group_max_last(t.val(0), WAVE), 0);
        if (lane == WAVE - 1) store_delegates<BETA, POSM, LT>(r, s, alpha, dvals, dpos, nullptr);
        fuse_add<BETA, FUSE, LT>(r, lane == WAVE - 1, lhist, hshift, mn, mx);
    } else {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t.shfl_merge(off);
        if (lane == 0) store_delegates<BETA, POSM, LT>(t, s, alpha, dvals, dpos, nullptr);
        fuse_add<BETA, FUSE, LT>(t, lane == 0, lhist, hshift, mn, mx);
    }
}

// ---------------------------------------------------------------- wave units, alpha >= 9
// A workgroup takes `rounds` consecutive groups of 4 units (one per wave) per step, so that with
// rounds 4 it streams 16 consecutive units (32 K keys at alpha <= 11) before striding by the grid.
template <class K, int BETA, int POSM, int FUSE, class LT>
__global__ __launch_bounds__(WG) void k_sample_waves(const u32* __restrict__ keys, u64 n, int alpha, int vec_ok,
                                                     u64 nunits, int rounds, u32* __restrict__ dvals,
                                                     void* __restrict__ dpos, u32* __restrict__ ghist, int hshift,
                                                     int bins, Ctrl* ctrl) {
    __shared__ u32 lhist[FUSE == 1 ? 256 : 1];
    __shared__ u32 red[16];
    fuse_init<FUSE>(lhist, bins);
    u32 mn = 0xFFFFFFFFu, mx = 0u;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const u64 L = 1ull << alpha;
    const u64 S = (n + L - 1) >> alpha;
    const int ps = (int)(L >> 8);            // 16-byte loads per lane per subrange (>= 2)
    const u64 U = L > 2048 ? L : 2048;       // keys per wave unit
    const int spu = (int)(U >> alpha);       // subranges per unit
    const u64 per = (u64)rounds * (WG / WAVE);  // units per workgroup step
    const u64 nsteps = (nunits + per - 1) / per;
    for (u64 step = blockIdx.x; step < nsteps; step += gridDim.x)
    for (int rd = 0; rd < rounds; ++rd) {
        const u64 unit = step * per + (u64)rd * (WG / WAVE) + w;
        if (unit >= nunits) break;  // uniform over the wave; no barrier in the loop
        const u64 u0 = unit * U;
        const bool full = vec_ok && u0 + U <= n;
        if (full && spu > 1) {
            // 8 loads per lane cover spu subranges of ps loads each
            u32x4 raw[8];
            const u32x4* p = (const u32x4*)(keys + u0) + lane;
#pragma unroll
            for (int j = 0; j < 8; ++j) raw[j] = load4(p + j * 64);
            for (int r = 0; r < spu; ++r) {
                LT t;
                t.clear();
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (j / ps != r) continue;
                    const u64 i0 = u0 + 4ull * (j * 64 + lane);
                    t.insert(K::to_ord(raw[j].x), i0);
                    t.insert(K::to_ord(raw[j].y), i0 + 1);
                    t.insert(K::to_ord(raw[j].z), i0 + 2);
                    t.insert(K::to_ord(raw[j].w), i0 + 3);
                }
                const u64 s = (u0 >> alpha) + r;
                wave_finish<BETA, POSM, FUSE, LT>(t, s, alpha, dvals, dpos, lhist, hshift, &mn, &mx);
            }
        } else if (full) {
            // one subrange per unit, ps >= 8 loads per lane in batches of 8
            LT t;
            t.clear();
            const u32x4* p = (const u32x4*)(keys + u0) + lane;
            for (int j0 = 0; j0 < ps; j0 += 8) {
                u32x4 raw[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) raw[u] = load4(p + (j0 + u) * 64);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const u64 i0 = u0 + 4ull * ((j0 + u) * 64 + lane);
                    if (BETA == 1 && POSM == 0) {
                        const u32 m = max(max(K::to_ord(raw[u].x), K::to_ord(raw[u].y)),
                                          max(K::to_ord(raw[u].z), K::to_ord(raw[u].w)));
                        t.insert(m, 0);
                    } else {
                        t.insert(K::to_ord(raw[u].x), i0);
                        t.insert(K::to_ord(raw[u].y), i0 + 1);
                        t.insert(K::to_ord(raw[u].z), i0 + 2);
                        t.insert(K::to_ord(raw[u].w), i0 + 3);
                    }
                }
            }
            wave_finish<BETA, POSM, FUSE, LT>(t, u0 >> alpha, alpha, dvals, dpos, lhist, hshift, &mn, &mx);
        } else {
            // tail unit: scalar loads, one subrange at a time
            for (int r = 0; r < spu; ++r) {
                const u64 sb = u0 + (u64)r * L;
                if (sb >= n) break;
                LT t;
                t.clear();
                for (u64 i = sb + lane; i < sb + L && i < n; i += 64) t.insert(K::to_ord(keys[i]), i);
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) t.shfl_merge(off);
                const u64 s = sb >> alpha;
                const bool leader = lane == 0 && s < S;
                if (leader) store_delegates<BETA, POSM, LT>(t, s, alpha, dvals, dpos, nullptr);
                fuse_add<BETA, FUSE, LT>(t, leader, lhist, hshift, &mn, &mx);
            }
        }
    }
    fuse_flush<FUSE>(lhist, ghist, bins, mn, mx, ctrl, red);
}

// ---------------------------------------------------------------- LDS staged, alpha 1..5
// A tile is 8192 keys (2^(13-alpha) subranges, 2^(5-alpha) per thread).  The next tile's
// eight 16-byte loads per lane are issued before the current tile is reduced from LDS, so a
// workgroup always has 32 KB of reads in flight.
constexpr int STAGE_Q = 8;                     // uint4 per thread per tile
constexpr int STAGE_KEYS = STAGE_Q * WG * 4;   // 8192

template <class K>
__device__ __forceinline__ void stage_fetch(const u32* __restrict__ keys, u64 n, u64 k0, bool full, u32x4 (&r)[STAGE_Q]) {
    if (full) {
        const u32x4* p = (const u32x4*)(keys + k0);
#pragma unroll
        for (int u = 0; u < STAGE_Q; ++u) r[u] = load4(p + u * WG + threadIdx.x);
    } else {
#pragma unroll
        for (int u = 0; u < STAGE_Q; ++u) {
            const u64 i = k0 + 4ull * (u * WG + threadIdx.x);
            u32x4 v = {0, 0, 0, 0};
            if (i + 0 < n) v.x = keys[i + 0];
            if (i + 1 < n) v.y = keys[i + 1];
            if (i + 2 < n) v.z = keys[i + 2];
            if (i + 3 < n) v.w = keys[i + 3];
            r[u] = v;
        }
    }
}

// Subrange j of a full tile (every key valid), 2^A keys, A >= 2, read from LDS.
template <int A, class LT>
__device__ __forceinline__ void staged_reduce_full(const u32x4* tile, int j, u64 kb, LT& t) {
    constexpr int QPS = 1 << (A - 2);
    if constexpr (std::is_same<LT, Top2P>::value) {
        u32 m1 = 0, m2 = 0, p1 = 0;
#pragma unroll
        for (int i = 0; i < QPS; ++i) {
            const int q = j * QPS + i;
            const u32x4 v = tile[q + (q >> 3)];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const u32 x = v[c];
                if (i == 0 && c == 0) {
                    m1 = x;
                    continue;
                }
                m2 = max(m2, min(m1, x));
                p1 = x > m1 ? (u32)(4 * i + c) : p1;
                m1 = max(m1, x);
            }
        }
        t.m1 = m1;
        t.m2 = m2;
        t.p1 = (u32)kb + p1;
    } else {
#pragma unroll
        for (int i = 0; i < QPS; ++i) {
            const int q = j * QPS + i;
            const u32x4 v = tile[q + (q >> 3)];
#pragma unroll
            for (int c = 0; c < 4; ++c) t.insert(v[c], kb + 4 * i + c);
        }
    }
}

template <int A, int BETA, int POSM, int FUSE, class LT>
__device__ __forceinline__ void staged_tile_full(const u32x4* tile, u64 k0, u32* dvals, void* dpos, u8* dtop,
                                                 u32* lhist, int hshift, u32* mn, u32* mx) {
    constexpr int SPT = STAGE_KEYS >> A;
#pragma unroll
    for (int j0 = 0; j0 < SPT; j0 += WG) {
        const int j = j0 + threadIdx.x;
        const u64 s = (k0 >> A) + j;
        LT t;
        t.clear();
        staged_reduce_full<A, LT>(tile, j, k0 + ((u64)j << A), t);
        store_delegates<BETA, POSM, LT>(t, s, A, dvals, dpos, dtop);
        fuse_add<BETA, FUSE, LT>(t, true, lhist, hshift, mn, mx);
    }
}

// At most 128 VGPRs (4 waves per SIMD): the LDS tile allows 4 workgroups per CU and the grid is
// sized for exactly that many (at 130 VGPRs one workgroup in four waited for a second round).
template <class K, int BETA, int POSM, int FUSE, class LT>
__global__ __launch_bounds__(WG, 4) void k_sample_staged(const u32* __restrict__ keys, u64 n, int alpha, int vec_ok,
                                                      u64 t_begin, u64 ntiles, u32* __restrict__ dvals,
                                                      void* __restrict__ dpos, u32* __restrict__ ghist, int hshift,
                                                      int bins, Ctrl* ctrl, u8* __restrict__ dtop) {
    // one pad uint4 after every 8 uint4 spreads the threads' subrange reads over the LDS banks
    __shared__ u32x4 tile[STAGE_Q * WG + STAGE_Q * WG / 8];
    __shared__ u32 lhist[FUSE == 1 ? 256 : 1];
    __shared__ u32 red[16];
    fuse_init<FUSE>(lhist, bins);
    u32 mn = 0xFFFFFFFFu, mx = 0u;
    const u64 L = 1ull << alpha;
    const u64 S = (n + L - 1) >> alpha;
    const int spt = STAGE_KEYS >> alpha;  // subranges per tile
    const int qps = alpha >= 2 ? (int)(L >> 2) : 1;  // uint4 per subrange
    u32x4 r[STAGE_Q];
    u64 tt = t_begin + blockIdx.x;
    if (tt < ntiles) {
        const u64 k0 = tt * STAGE_KEYS;
        stage_fetch<K>(keys, n, k0, vec_ok && k0 + STAGE_KEYS <= n, r);
    }
    for (; tt < ntiles; tt += gridDim.x) {
        const u64 k0 = tt * STAGE_KEYS;
#pragma unroll
        for (int u = 0; u < STAGE_Q; ++u) {
            const int q = u * WG + threadIdx.x;
            u32x4 v = r[u];
            v.x = K::to_ord(v.x); v.y = K::to_ord(v.y); v.z = K::to_ord(v.z); v.w = K::to_ord(v.w);
            tile[q + (q >> 3)] = v;
        }
        const u64 nt = tt + gridDim.x;
        if (nt < ntiles) {
            const u64 k1 = nt * STAGE_KEYS;
            stage_fetch<K>(keys, n, k1, vec_ok && k1 + STAGE_KEYS <= n, r);
        }
        __syncthreads();
        const u64 s0 = k0 >> alpha;
        if (vec_ok && k0 + STAGE_KEYS <= n && alpha >= 2) {  // uniform: a full tile
            switch (alpha) {
                case 2: staged_tile_full<2, BETA, POSM, FUSE, LT>(tile, k0, dvals, dpos, dtop, lhist, hshift, &mn, &mx); break;
                case 3: staged_tile_full<3, BETA, POSM, FUSE, LT>(tile, k0, dvals, dpos, dtop, lhist, hshift, &mn, &mx); break;
                case 4: staged_tile_full<4, BETA, POSM, FUSE, LT>(tile, k0, dvals, dpos, dtop, lhist, hshift, &mn, &mx); break;
                default: staged_tile_full<5, BETA, POSM, FUSE, LT>(tile, k0, dvals, dpos, dtop, lhist, hshift, &mn, &mx); break;
            }
            __syncthreads();
            continue;
        }
        for (int j = threadIdx.x; j < spt; j += WG) {  // uniform trip count: spt is a multiple of WG
            const u64 s = s0 + j;
            const u64 kb = k0 + ((u64)j << alpha);
            LT t;
            t.clear();
            if (alpha >= 2) {
                for (int i = 0; i < qps; ++i) {
                    const int q = j * qps + i;
                    const u32x4 v = tile[q + (q >> 3)];
                    const u64 i0 = kb + 4ull * i;
                    if (i0 + 3 < n) {
                        t.insert(v.x, i0); t.insert(v.y, i0 + 1);
                        t.insert(v.z, i0 + 2); t.insert(v.w, i0 + 3);
                    } else {
                        if (i0 + 0 < n) t.insert(v.x, i0);
                        if (i0 + 1 < n) t.insert(v.y, i0 + 1);
                        if (i0 + 2 < n) t.insert(v.z, i0 + 2);
                    }
                }
            } else {
                // alpha == 1: two keys per subrange
                const int q = j >> 1;
                const u32x4 v = tile[q + (q >> 3)];
                const u32 a = (j & 1) == 0 ? v.x : v.z, b = (j & 1) == 0 ? v.y : v.w;
                if (kb < n) t.insert(a, kb);
                if (kb + 1 < n) t.insert(b, kb + 1);
            }
            const bool leader = s < S;
            if (leader) store_delegates<BETA, POSM, LT>(t, s, alpha, dvals, dpos, dtop);
            fuse_add<BETA, FUSE, LT>(t, leader, lhist, hshift, &mn, &mx);
        }
        __syncthreads();
    }
    fuse_flush<FUSE>(lhist, ghist, bins, mn, mx, ctrl, red);
}

// ---------------------------------------------------------------- beta 2, lane groups with DPP
// Beta 2, full 8192-key chunks, alpha 2..5 a compile-time constant, no LDS staging: lane l of a
// wave holds keys 4l..4l+3 of a 256-key wave tile (eight 16-byte loads in flight per lane), takes
// their top two (with the lowest index of the maximum) and merges them across the 2^(A-2) lanes
// of its subrange with DPP lane moves: quad_perm swaps within a quad (xor 1, xor 2) and the row
// half mirror (lane i <-> 7 - i of each 8) for 8-lane groups.  Every lane of a group ends with
// the subrange's result; the group's first lane stores it.
__device__ __forceinline__ u32 dpp_x1(u32 v) { return (u32)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false); }
__device__ __forceinline__ u32 dpp_x2(u32 v) { return (u32)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false); }
__device__ __forceinline__ u32 dpp_hm(u32 v) { return (u32)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false); }

template <int STEP>
__device__ __forceinline__ u32 dpp_partner(u32 v) {
    if constexpr (STEP == 0) return dpp_x1(v);
    else if constexpr (STEP == 1) return dpp_x2(v);
    else return dpp_hm(v);
}

// top two (multiset) of two lanes' top twos, and the lowest index holding the maximum
template <int STEP, bool P>
__device__ __forceinline__ void top2_merge(u32& m1, u32& m2, u32& p1) {
    const u32 b1 = dpp_partner<STEP>(m1), b2 = dpp_partner<STEP>(m2);
    m2 = max(min(m1, b1), max(m2, b2));
    if (P) {
        const u32 bp = dpp_partner<STEP>(p1);
        p1 = m1 > b1 ? p1 : (b1 > m1 ? bp : min(p1, bp));
    }
    m1 = max(m1, b1);
}

template <class K, int POSM, int FUSE, int A>
__global__ __launch_bounds__(WG) void k_sample_dpp2(const u32* __restrict__ keys, u64 nchunks, u32* __restrict__ dvals,
                                                    void* __restrict__ dpos, u32* __restrict__ ghist, int hshift,
                                                    int bins, Ctrl* ctrl, u8* __restrict__ dtop) {
    constexpr int G = 1 << (A - 2);  // lanes per subrange
    constexpr bool P = POSM == 2;
    __shared__ u32 lhist[FUSE == 1 ? 256 : 1];
    __shared__ u32 red[16];
    fuse_init<FUSE>(lhist, bins);
    u32 mn = 0xFFFFFFFFu, mx = 0u;
    const int lane = threadIdx.x & (WAVE - 1);
    const bool leader = (lane & (G - 1)) == 0;
    for (u64 chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const u64 c0 = chunk * GROUP_CHUNK;
        const u32x4* base4 = (const u32x4*)(keys + c0);
        u32x4 raw[GROUP_BATCHES];
#pragma unroll
        for (int b = 0; b < GROUP_BATCHES; ++b) raw[b] = load4(base4 + b * WG + threadIdx.x);
#pragma unroll
        for (int b = 0; b < GROUP_BATCHES; ++b) {
            const u32 i0 = (u32)(c0 + 4ull * (b * WG + threadIdx.x));  // n < 2^31
            const u32 x0 = K::to_ord(raw[b].x), x1 = K::to_ord(raw[b].y), x2 = K::to_ord(raw[b].z),
                      x3 = K::to_ord(raw[b].w);
            // top two of the lane's four keys, index of the first maximum
            u32 m1 = x0, m2 = 0, p1 = i0;
            m2 = max(m2, min(m1, x1)); p1 = x1 > m1 ? i0 + 1 : p1; m1 = max(m1, x1);
            m2 = max(m2, min(m1, x2)); p1 = x2 > m1 ? i0 + 2 : p1; m1 = max(m1, x2);
            m2 = max(m2, min(m1, x3)); p1 = x3 > m1 ? i0 + 3 : p1; m1 = max(m1, x3);
            if constexpr (G >= 2) top2_merge<0, P>(m1, m2, p1);
            if constexpr (G >= 4) top2_merge<1, P>(m1, m2, p1);
            if constexpr (G >= 8) top2_merge<2, P>(m1, m2, p1);
            const u32 s = i0 >> A;
            if (leader) {
                const u32 off = p1 - (s << A);
                if (dtop) {
                    dvals[s] = m1;
                    ((uint16_t*)dtop)[s] = (uint16_t)((m2 >> 24) | ((P ? off : 0xFFu) << 8));
                } else {
                    *(uint2*)(dvals + 2 * (u64)s) = make_uint2(m1, m2);
                    if (P) ((uint8_t*)dpos)[s] = (uint8_t)off;
                }
            }
            if (FUSE == 1) {
                wave_hist_add(lhist, leader, m1 >> hshift);
                wave_hist_add(lhist, leader, m2 >> hshift);
            }
            if (FUSE == 2 && leader) {
                mn = min(mn, m2);
                mx = max(mx, m1);
            }
        }
    }
    fuse_flush<FUSE>(lhist, ghist, bins, mn, mx, ctrl, red);
}

// ---------------------------------------------------------------- dispatch
constexpr u64 SAMPLE_GRID = 8192;  // persistent workgroups (32 per CU)
// Streaming order of the wave-unit sampler (alpha >= 9): a workgroup takes `rounds` consecutive
// groups of its 4 waves' units per step and the steps stride by the grid.  Over 2^30 uint32 keys
// 4 rounds (32 K keys per step at alpha <= 11) with 4096 workgroups ran the alpha-11 pass in
// 0.627 ms against 0.660 ms with one round and 8192 (2 rounds: 0.662, 8 rounds: 0.624; 2048 or
// 8192 workgroups at 4 rounds: 0.633 / 0.632).  The lane-group, DPP and staged samplers (alpha
// <= 8) gained nothing from the same order and keep one chunk per step.  Below 2^27 keys one
// round keeps every resident workgroup busy.
constexpr u64 STREAM_GRID = 4096;
static int stream_rounds(u64 nchunks) { return nchunks >= 16384 ? 4 : 1; }

template <class K, int BETA, int POSM, int FUSE>
static hipError_t launch_sample_t(const SampleArgs& a, hipStream_t st) {
    using LT = typename ListOf<BETA, POSM>::type;
    const u64 L = 1ull << a.alpha;
    const u64 S = (a.n + L - 1) >> a.alpha;
    const int vec_ok = ((uintptr_t)a.keys & 15) == 0;
    // the compact beta-2 form (dtop) is written by the staged and the DPP lane-group kernels
    const bool dpp2_ok = BETA == 2 && POSM != 1 && FUSE != 0 && a.alpha >= 2 && a.alpha <= 5 && vec_ok &&
                         a.n >= (u64)GROUP_CHUNK;
    if ((a.staged && a.alpha <= 5) || (a.dtop && !dpp2_ok)) {
        // exactly the resident workgroups (LDS allows 4 per CU): every one keeps its pipeline full
        const u64 ntiles = (a.n + STAGE_KEYS - 1) / STAGE_KEYS;
        const unsigned grid = (unsigned)std::min<u64>(ntiles, 4ull * device_cus());
        hipLaunchKernelGGL((k_sample_staged<K, BETA, POSM, FUSE, LT>), dim3(grid), dim3(WG), 0, st, a.keys, a.n,
                           a.alpha, vec_ok, (u64)0, ntiles, a.dvals, a.dpos, a.ghist, a.hshift, a.bins, a.ctrl,
                           BETA == 2 ? a.dtop : nullptr);
    } else if (dpp2_ok) {
        // lane groups with DPP merges over the full chunks, the staged kernel over the tail
        const u64 nfull = a.n / GROUP_CHUNK;
        const unsigned g2 = (unsigned)std::min<u64>(nfull, SAMPLE_GRID);
#define D2_LAUNCH(AA)                                                                                        \
    hipLaunchKernelGGL((k_sample_dpp2<K, POSM, FUSE, AA>), dim3(g2), dim3(WG), 0, st, a.keys, nfull, a.dvals, \
                       a.dpos, a.ghist, a.hshift, a.bins, a.ctrl, a.dtop)
        switch (a.alpha) {
            case 2: D2_LAUNCH(2); break;
            case 3: D2_LAUNCH(3); break;
            case 4: D2_LAUNCH(4); break;
            default: D2_LAUNCH(5); break;
        }
#undef D2_LAUNCH
        static_assert(GROUP_CHUNK == STAGE_KEYS, "the tail is resumed in staged tiles of the same size");
        const u64 ntiles = (a.n + STAGE_KEYS - 1) / STAGE_KEYS;
        if (nfull < ntiles)
            hipLaunchKernelGGL((k_sample_staged<K, BETA, POSM, FUSE, LT>), dim3(1), dim3(WG), 0, st, a.keys, a.n,
                               a.alpha, vec_ok, nfull, ntiles, a.dvals, a.dpos, a.ghist, a.hshift, a.bins, a.ctrl,
                               BETA == 2 ? a.dtop : nullptr);
    } else if (a.alpha <= 8) {
        const u64 nchunks = (a.n + GROUP_CHUNK - 1) / GROUP_CHUNK;
        const unsigned grid = (unsigned)std::min<u64>(nchunks, SAMPLE_GRID);
        hipLaunchKernelGGL((k_sample_groups<K, BETA, POSM, FUSE, LT>), dim3(grid), dim3(WG), 0, st, a.keys, a.n,
                           a.alpha, vec_ok, nchunks, a.dvals, a.dpos, a.ghist, a.hshift, a.bins, a.ctrl);
    } else {
        const u64 U = L > 2048 ? L : 2048;
        const u64 nunits = (a.n + U - 1) / U;
        const int rounds = stream_rounds((nunits + 3) / 4);  // 4 wave units per workgroup round
        const unsigned grid = (unsigned)std::min<u64>((nunits + 4 * rounds - 1) / (4 * rounds),
                                                      rounds > 1 ? STREAM_GRID : SAMPLE_GRID);
        hipLaunchKernelGGL((k_sample_waves<K, BETA, POSM, FUSE, LT>), dim3(grid), dim3(WG), 0, st, a.keys, a.n,
                           a.alpha, vec_ok, nunits, rounds, a.dvals, a.dpos, a.ghist, a.hshift, a.bins, a.ctrl);
    }
    return hipGetLastError();
}

template <class K, int BETA>
static hipError_t launch_sample_b(const SampleArgs& a, hipStream_t st) {
    switch (a.posmode * 3 + a.fuse) {
        case 0: return launch_sample_t<K, BETA, 0, 0>(a, st);
        case 1: return launch_sample_t<K, BETA, 0, 1>(a, st);
        case 2: return launch_sample_t<K, BETA, 0, 2>(a, st);
        case 3: return launch_sample_t<K, BETA, 1, 0>(a, st);
        case 4: return launch_sample_t<K, BETA, 1, 1>(a, st);
        case 5: return launch_sample_t<K, BETA, 1, 2>(a, st);
        case 6: return launch_sample_t<K, BETA, 2, 0>(a, st);
        case 7: return launch_sample_t<K, BETA, 2, 1>(a, st);
        case 8: return launch_sample_t<K, BETA, 2, 2>(a, st);
    }
    return hipErrorInvalidValue;
}

template <class K>
static hipError_t launch_sample_k(const SampleArgs& a, hipStream_t st) {
    switch (a.beta) {
        case 1: return launch_sample_b<K, 1>(a, st);
        case 2: return launch_sample_b<K, 2>(a, st);
        case 3: return launch_sample_b<K, 3>(a, st);
        case 4: return launch_sample_b<K, 4>(a, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_sample(const SampleArgs& a, hipStream_t st) {
    if (a.posmode == 2 && a.beta < 2) return hipErrorInvalidValue;
    if (a.dtop && !(a.beta == 2 && a.alpha <= 5)) return hipErrorInvalidValue;
    return a.f32 ? launch_sample_k<KeyF32>(a, st) : launch_sample_k<KeyU32>(a, st);
}

}  // namespace drtopk
