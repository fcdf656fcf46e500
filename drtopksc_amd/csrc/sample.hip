// sample.hip -- delegate pass: the only kernel that reads all of V.
//
// Restates sampleMax (baseline+filter+beta+shuffle/radixselect.cuh:883-918) and
// sampleMax_NoReduction_Beta2/3/4 (radixselect.cuh:967-1202): for every subrange of
// 2^alpha keys keep its maximum (beta == 1) or its top-beta multiset (beta 2..4),
// largest first, lowest index first among equal keys.
//
// Three variants, identical output:
//  - lane groups (alpha <= 8): a workgroup streams 8192 keys with 16-byte loads, lane l
//    of a wave holding keys 4l..4l+3 of a 256-key wave tile; a subrange is a group of
//    2^alpha/4 lanes reduced with xor shuffles (2 subranges per lane at alpha == 1).
//  - wave per subrange (alpha >= 9): a wave strides one subrange with 16-byte loads and
//    reduces over its 64 lanes.
//  - LDS staged, thread per subrange (alpha <= 5, DRTOPK_FLAG_SHUFFLE): the workgroup
//    stages 256 subranges in LDS with coalesced loads, then each thread scans its own
//    subrange (the reference's "shuffle" option).
//
// Fused into the same pass: the histogram of the delegates' top digit (radix family,
// the first digit of the first top-k) or the delegates' min/max (bucket family).
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace drtopk {

// ---------------------------------------------------------------- top-beta lists
// Element: u64 composite (ordered key << 32 | ~index) when positions are tracked, else the
// ordered key (u32 widened).  Larger element = larger key, then lower index.  Absent = 0.
template <int BETA>
struct TopList {
    u64 v[BETA];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int j = 0; j < BETA; ++j) v[j] = 0;
    }
    __device__ __forceinline__ void insert(u64 x) {
        // keep v sorted descending; drop the smallest
#pragma unroll
        for (int j = 0; j < BETA; ++j) {
            u64 hi = v[j] > x ? v[j] : x;
            u64 lo = v[j] > x ? x : v[j];
            v[j] = hi;
            x = lo;
        }
    }
    // top-BETA of the union of two descending lists: max(a[i], b[BETA-1-i]) is a bitonic
    // sequence holding the top BETA; a small sorting network orders it.
    __device__ __forceinline__ void merge(const u64* b) {
        u64 c[BETA];
#pragma unroll
        for (int j = 0; j < BETA; ++j) c[j] = v[j] > b[BETA - 1 - j] ? v[j] : b[BETA - 1 - j];
#pragma unroll
        for (int p = 0; p < BETA; ++p)
#pragma unroll
            for (int q = 0; q + 1 < BETA - p; ++q) {
                u64 x = c[q], y = c[q + 1];
                c[q] = x > y ? x : y;
                c[q + 1] = x > y ? y : x;
            }
#pragma unroll
        for (int j = 0; j < BETA; ++j) v[j] = c[j];
    }
    __device__ __forceinline__ void shfl_merge(int off) {
        u64 b[BETA];
#pragma unroll
        for (int j = 0; j < BETA; ++j) b[j] = shfl_xor_u64(v[j], off);
        merge(b);
    }
};

template <bool COMPOSITE>
__device__ __forceinline__ u64 elem(u32 ord, u64 pos) {
    if (COMPOSITE) return ((u64)ord << 32) | (u64)(0xFFFFFFFFu - (u32)pos);
    return (u64)ord;
}

// Write subrange s's delegates (+ positions / packed offsets) and feed the fused reduction.
// POSM: 0 none, 1 absolute positions (u32 per delegate), 2 packed offsets of the first
// beta-1 delegates inside the subrange (8 bits each; u8 / u16 / u32 per subrange).
// FUSE: 0 none, 1 radix top-digit histogram into lhist, 2 min/max into *mn/*mx.
template <int BETA, int POSM, int FUSE, bool COMPOSITE>
__device__ __forceinline__ void emit(const TopList<BETA>& t, u64 s, int alpha, u32* dvals, void* dpos,
                                     u32* lhist, int hshift, u32* mn, u32* mx) {
    u32 packed = 0;
#pragma unroll
    for (int j = 0; j < BETA; ++j) {
        u32 val = COMPOSITE ? (u32)(t.v[j] >> 32) : (u32)t.v[j];
        dvals[s * BETA + j] = val;
        if (POSM != 0) {
            u32 pos = t.v[j] == 0 ? 0xFFFFFFFFu : 0xFFFFFFFFu - (u32)t.v[j];
            if (POSM == 1) ((u32*)dpos)[s * BETA + j] = pos;
            if (POSM == 2 && j < BETA - 1) {
                u32 off = pos == 0xFFFFFFFFu ? 0xFFu : (u32)(pos - (s << alpha)) & 0xFFu;
                packed |= off << (8 * j);
            }
        }
        if (FUSE == 1) atomicAdd(&lhist[val >> hshift], 1u);
        if (FUSE == 2) { *mn = min(*mn, val); *mx = max(*mx, val); }
    }
    if (POSM == 2) {
        if (BETA == 2) ((uint8_t*)dpos)[s] = (uint8_t)packed;
        else if (BETA == 3) ((uint16_t*)dpos)[s] = (uint16_t)packed;
        else if (BETA == 4) ((u32*)dpos)[s] = packed;
    }
}

template <int FUSE>
__device__ __forceinline__ void fuse_flush(u32* lhist, u32* ghist, int bins, u32 mn, u32 mx, Ctrl* ctrl,
                                           u32* red) {
    if (FUSE == 1) {
        __syncthreads();
        for (int i = threadIdx.x; i < bins; i += blockDim.x) {
            u32 c = lhist[i];
            if (c) atomicAdd(&ghist[i], c);
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
            mn = min(mn, red[0]); mx = max(mx, red[8]);
            atomicMin(&ctrl->first.lo, mn);
            atomicMax(&ctrl->first.hi, mx);
        }
    }
}

__device__ __forceinline__ u32x4 load4(const u32x4* p) { return __builtin_nontemporal_load(p); }

// ---------------------------------------------------------------- lane groups, alpha 1..8
constexpr int GROUP_BATCHES = 8;                      // 16-byte loads in flight per lane
constexpr int GROUP_CHUNK = GROUP_BATCHES * WG * 4;   // 8192 keys per workgroup

template <class K, int BETA, int POSM, int FUSE>
__global__ __launch_bounds__(WG) void k_sample_groups(const u32* __restrict__ keys, u64 n, int alpha, int vec_ok,
                                                      u32* __restrict__ dvals, void* __restrict__ dpos,
                                                      u32* __restrict__ ghist, int hshift, int bins, Ctrl* ctrl) {
    constexpr bool COMP = (BETA > 1) || (POSM != 0);
    __shared__ u32 lhist[FUSE == 1 ? 256 : 1];
    __shared__ u32 red[16];
    if (FUSE == 1) {
        for (int i = threadIdx.x; i < bins; i += WG) lhist[i] = 0;
        __syncthreads();
    }
    u32 mn = 0xFFFFFFFFu, mx = 0u;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const u64 c0 = (u64)blockIdx.x * GROUP_CHUNK;
    const u64 S = (n + (1ull << alpha) - 1) >> alpha;
    u32x4 raw[GROUP_BATCHES];
    const bool full = vec_ok && c0 + GROUP_CHUNK <= n;
    const u32x4* base4 = (const u32x4*)(keys + c0);
#pragma unroll
    for (int b = 0; b < GROUP_BATCHES; ++b) {
        const int q = b * WG + w * 64 + lane;  // uint4 index inside the chunk
        if (full) raw[b] = load4(base4 + q);
        else {
            u64 i0 = c0 + 4ull * q;
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
        u32 kv[4] = {raw[b].x, raw[b].y, raw[b].z, raw[b].w};
        if (alpha == 1) {
            // two subranges of two keys in this lane
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                TopList<BETA> t;
                t.clear();
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    u64 i = i0 + 2 * h + c;
                    if (i < n) t.insert(elem<COMP>(K::to_ord(kv[2 * h + c]), i));
                }
                u64 s = (i0 >> 1) + h;
                if (i0 + 2 * h < n)
                    emit<BETA, POSM, FUSE, COMP>(t, s, alpha, dvals, dpos, lhist, hshift, &mn, &mx);
            }
        } else {
            TopList<BETA> t;
            t.clear();
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (i0 + c < n) t.insert(elem<COMP>(K::to_ord(kv[c]), i0 + c));
            const int g = 1 << (alpha - 2);  // lanes per subrange (1..64)
            for (int off = g >> 1; off > 0; off >>= 1) t.shfl_merge(off);
            const u64 s = i0 >> alpha;
            if ((lane & (g - 1)) == 0 && s < S && i0 < n)
                emit<BETA, POSM, FUSE, COMP>(t, s, alpha, dvals, dpos, lhist, hshift, &mn, &mx);
        }
    }
    fuse_flush<FUSE>(lhist, ghist, bins, mn, mx, ctrl, red);
}

// ---------------------------------------------------------------- wave per subrange, alpha >= 9
template <class K, int BETA, int POSM, int FUSE>
__global__ __launch_bounds__(WG) void k_sample_waves(const u32* __restrict__ keys, u64 n, int alpha, int vec_ok, int spw,
                                                     u32* __restrict__ dvals, void* __restrict__ dpos,
                                                     u32* __restrict__ ghist, int hshift, int bins, Ctrl* ctrl) {
    constexpr bool COMP = (BETA > 1) || (POSM != 0);
    __shared__ u32 lhist[FUSE == 1 ? 256 : 1];
    __shared__ u32 red[16];
    if (FUSE == 1) {
        for (int i = threadIdx.x; i < bins; i += WG) lhist[i] = 0;
        __syncthreads();
    }
    u32 mn = 0xFFFFFFFFu, mx = 0u;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const u64 L = 1ull << alpha;
    const u64 S = (n + L - 1) >> alpha;
    const int per_lane = (int)(L >> 8);  // 16-byte loads per lane per subrange (>= 2)
    for (int r = 0; r < spw; ++r) {
        const u64 s = ((u64)blockIdx.x * 4 + w) * spw + r;
        if (s >= S) break;
        const u64 sb = s << alpha;
        TopList<BETA> t;
        t.clear();
        const bool full = vec_ok && sb + L <= n;
        if (full) {
            const u32x4* p = (const u32x4*)(keys + sb) + lane;
            for (int j0 = 0; j0 < per_lane; j0 += 8) {
                u32x4 raw[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (j0 + u < per_lane) raw[u] = load4(p + (j0 + u) * 64);
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (j0 + u < per_lane) {
                        u64 i0 = sb + 4ull * ((j0 + u) * 64 + lane);
                        if (BETA == 1 && !COMP) {
                            u32 m = max(max(K::to_ord(raw[u].x), K::to_ord(raw[u].y)),
                                        max(K::to_ord(raw[u].z), K::to_ord(raw[u].w)));
                            t.v[0] = t.v[0] > m ? t.v[0] : m;
                        } else {
                            t.insert(elem<COMP>(K::to_ord(raw[u].x), i0));
                            t.insert(elem<COMP>(K::to_ord(raw[u].y), i0 + 1));
                            t.insert(elem<COMP>(K::to_ord(raw[u].z), i0 + 2));
                            t.insert(elem<COMP>(K::to_ord(raw[u].w), i0 + 3));
                        }
                    }
            }
        } else {
            for (u64 i = sb + lane; i < sb + L && i < n; i += 64) t.insert(elem<COMP>(K::to_ord(keys[i]), i));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t.shfl_merge(off);
        if (lane == 0) emit<BETA, POSM, FUSE, COMP>(t, s, alpha, dvals, dpos, lhist, hshift, &mn, &mx);
    }
    fuse_flush<FUSE>(lhist, ghist, bins, mn, mx, ctrl, red);
}

// ---------------------------------------------------------------- LDS staged, alpha 1..5
template <class K, int BETA, int POSM, int FUSE>
__global__ __launch_bounds__(WG) void k_sample_staged(const u32* __restrict__ keys, u64 n, int alpha,
                                                      u32* __restrict__ dvals, void* __restrict__ dpos,
                                                      u32* __restrict__ ghist, int hshift, int bins, Ctrl* ctrl) {
    constexpr bool COMP = (BETA > 1) || (POSM != 0);
    // 256 subranges of up to 32 keys; one pad word after every 32 keys breaks bank conflicts
    __shared__ u32 tile[WG * 32 + WG];
    __shared__ u32 lhist[FUSE == 1 ? 256 : 1];
    __shared__ u32 red[16];
    if (FUSE == 1) {
        for (int i = threadIdx.x; i < bins; i += WG) lhist[i] = 0;
    }
    u32 mn = 0xFFFFFFFFu, mx = 0u;
    const u64 L = 1ull << alpha;
    const u64 S = (n + L - 1) >> alpha;
    const u64 s0 = (u64)blockIdx.x * WG;
    const u64 k0 = s0 << alpha;
    const int tile_keys = WG << alpha;
    // coalesced staging: key j of the tile goes to tile[j + j/32]
    for (int j = threadIdx.x; j < tile_keys; j += WG) {
        u64 i = k0 + j;
        u32 v = i < n ? K::to_ord(keys[i]) : 0u;
        tile[j + (j >> 5)] = v;
    }
    __syncthreads();
    const u64 s = s0 + threadIdx.x;
    if (s < S) {
        TopList<BETA> t;
        t.clear();
        const int jb = threadIdx.x << alpha;
        for (int c = 0; c < (int)L; ++c) {
            u64 i = k0 + jb + c;
            if (i < n) {
                int j = jb + c;
                t.insert(elem<COMP>(tile[j + (j >> 5)], i));
            }
        }
        emit<BETA, POSM, FUSE, COMP>(t, s, alpha, dvals, dpos, lhist, hshift, &mn, &mx);
    }
    fuse_flush<FUSE>(lhist, ghist, bins, mn, mx, ctrl, red);
}

// ---------------------------------------------------------------- dispatch
template <class K, int BETA, int POSM, int FUSE>
static hipError_t launch_sample_t(const SampleArgs& a, hipStream_t st) {
    const u64 L = 1ull << a.alpha;
    const u64 S = (a.n + L - 1) >> a.alpha;
    const int vec_ok = ((uintptr_t)a.keys & 15) == 0;
    if (a.staged && a.alpha <= 5) {
        unsigned grid = (unsigned)((S + WG - 1) / WG);
        hipLaunchKernelGGL((k_sample_staged<K, BETA, POSM, FUSE>), dim3(grid), dim3(WG), 0, st, a.keys, a.n,
                           a.alpha, a.dvals, a.dpos, a.ghist, a.hshift, a.bins, a.ctrl);
    } else if (a.alpha <= 8) {
        unsigned grid = (unsigned)((a.n + GROUP_CHUNK - 1) / GROUP_CHUNK);
        hipLaunchKernelGGL((k_sample_groups<K, BETA, POSM, FUSE>), dim3(grid), dim3(WG), 0, st, a.keys, a.n,
                           a.alpha, vec_ok, a.dvals, a.dpos, a.ghist, a.hshift, a.bins, a.ctrl);
    } else {
        int spw = a.alpha >= 11 ? 1 : (2048 >> a.alpha);
        u64 waves = (S + spw - 1) / spw;
        unsigned grid = (unsigned)((waves + 3) / 4);
        hipLaunchKernelGGL((k_sample_waves<K, BETA, POSM, FUSE>), dim3(grid), dim3(WG), 0, st, a.keys, a.n,
                           a.alpha, vec_ok, spw, a.dvals, a.dpos, a.ghist, a.hshift, a.bins, a.ctrl);
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
    return a.f32 ? launch_sample_k<KeyF32>(a, st) : launch_sample_k<KeyU32>(a, st);
}

}  // namespace drtopk
