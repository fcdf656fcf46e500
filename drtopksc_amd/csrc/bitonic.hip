// bitonic.hip -- the bitonic top-k family of Dr.Top-k on gfx950.
//
// Restates the SIGMOD-style network of bitonic/bitonic_com.cuh: sort into k-runs of
// alternating direction (Combined_Sigmod_Merge*, :101-228 / :231-386), merge pairs of runs
// by elementwise max (:299-321), rebuild each run with a bitonic merge (:332-356), repeat
// until k elements remain.  Every element is a 64-bit composite (ordered key << 32 | ~index),
// so every comparison is strict: the top-k is by key, lowest index first among equal keys,
// and the result does not depend on how the network is scheduled.
//
// Scheduling on gfx950 (160 KB of LDS per CU):
//  - k <= T/2: one workgroup reduces an LDS tile of T composites to its own top-k (sort into
//    k-runs, then log2(T/k) merge+rebuild rounds entirely in LDS); tiles are reduced level by
//    level until one tile remains.  T = 8192 (64 KB, two workgroups per CU) for k <= 2048 and
//    T = 16384 (128 KB, one 1024-thread workgroup per CU) above, so k up to 8192 stays in LDS
//    and each level shrinks the data by T/k >= 2.  The reference instead halves once per launch
//    (bitonic_com.cuh:648-657).
//  - k > 8192 (the reference's largerK global-scratch variant): the network runs over a global
//    buffer; strides >= 16384 are global passes, all smaller strides of a stage are fused into
//    one LDS pass per 16384-composite tile.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace drtopk {

constexpr u32 BT_SMALL = 8192;   // LDS tile of composites for k <= 2048 (64 KB)
constexpr int BTT_SMALL = 512;
constexpr u32 BT = 16384;        // LDS tile for larger k and the global network (128 KB)
constexpr int BTT = 1024;

__device__ __forceinline__ void cmpx(u64* A, u32 i, u32 inc, bool up) {
    const u64 a = A[i], b = A[i + inc];
    if (up == (a > b)) {
        A[i] = b;
        A[i + inc] = a;
    }
}

// Steps inc = inc0 .. 1 of stage `len` on a tile of T elements whose first global index is
// `base`; the direction of a pair comes from bit 2*len of its global index.
template <int NT>
__device__ __forceinline__ void lds_steps(u64* A, u32 T, u64 base, u64 len, u64 inc0) {
    for (u64 inc = inc0; inc > 0; inc >>= 1) {
        for (u32 t = threadIdx.x; t < T / 2; t += NT) {
            const u32 low = t & (u32)(inc - 1), i = 2 * t - low;
            cmpx(A, i, (u32)inc, ((base + i) & (2 * len)) == 0);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- tile reduction (k <= TILE/2)
// Reduces tile b of A (T <= TILE elements) to its top-k, sorted ascending, at out[b*k, b*k + k).
// presorted: A holds runs of k each sorted ascending (a previous level's output); odd runs are
// loaded reversed, which gives the alternating k-runs the merge needs without sorting again.
template <u32 TILE, int NT>
__global__ __launch_bounds__(NT) void k_bt_tile_reduce(const u64* __restrict__ A, u32 T, u32 k, int presorted,
                                                       u64* __restrict__ out) {
    __shared__ u64 S[TILE];
    const u64 base = (u64)blockIdx.x * T;
    if (presorted) {
        for (u32 i = threadIdx.x; i < T; i += NT) {
            const u32 j = i & (k - 1), r0 = i - j;  // position in its run, first index of the run
            S[i] = A[base + ((i & k) ? r0 + (k - 1 - j) : i)];
        }
        __syncthreads();
    } else {
        for (u32 i = threadIdx.x; i < T; i += NT) S[i] = A[base + i];
        __syncthreads();
        // sort into alternating k-runs (tile-local directions)
        for (u64 len = 1; len < k; len <<= 1) lds_steps<NT>(S, T, 0, len, len);
    }
    u32 cur = T;
    while (cur > k) {
        // merge: element t of the halved tile = max of the pair (i, i+k) of runs 2r, 2r+1
        u64 m[TILE / 2 / NT];
#pragma unroll
        for (u32 j = 0; j < TILE / 2 / NT; ++j) {
            const u32 t = threadIdx.x + j * NT;
            if (t < cur / 2) {
                const u32 i = 2 * t - (t & (k - 1));
                m[j] = S[i] > S[i + k] ? S[i] : S[i + k];
            }
        }
        __syncthreads();
#pragma unroll
        for (u32 j = 0; j < TILE / 2 / NT; ++j) {
            const u32 t = threadIdx.x + j * NT;
            if (t < cur / 2) S[t] = m[j];
        }
        __syncthreads();
        cur >>= 1;
        // rebuild every run of k into sorted order (alternating)
        lds_steps<NT>(S, cur, 0, k >> 1, k >> 1);
    }
    for (u32 i = threadIdx.x; i < k; i += NT) out[(u64)blockIdx.x * k + i] = S[i];
}

// ---------------------------------------------------------------- global network (k > BT/2)
// Stages len < BT of the run sort, all in LDS (runs of up to BT).
__global__ __launch_bounds__(BTT) void k_bt_sort_lds(u64* __restrict__ A, u64 lenmax) {
    __shared__ u64 S[BT];
    const u64 base = (u64)blockIdx.x * BT;
    for (u32 i = threadIdx.x; i < BT; i += BTT) S[i] = A[base + i];
    __syncthreads();
    for (u64 len = 1; len <= lenmax; len <<= 1) lds_steps<BTT>(S, BT, base, len, len);
    for (u32 i = threadIdx.x; i < BT; i += BTT) A[base + i] = S[i];
}

// The strides < BT of stage `len` (inc = BT/2 .. 1), one LDS pass per tile.
__global__ __launch_bounds__(BTT) void k_bt_stage_lds(u64* __restrict__ A, u64 len) {
    __shared__ u64 S[BT];
    const u64 base = (u64)blockIdx.x * BT;
    for (u32 i = threadIdx.x; i < BT; i += BTT) S[i] = A[base + i];
    __syncthreads();
    lds_steps<BTT>(S, BT, base, len, BT / 2);
    for (u32 i = threadIdx.x; i < BT; i += BTT) A[base + i] = S[i];
}

// One stride >= BT of stage `len`, over all n/2 pairs.
__global__ __launch_bounds__(WG) void k_bt_step(u64* __restrict__ A, u64 n, u64 len, u64 inc) {
    const u64 stride = (u64)gridDim.x * WG;
    for (u64 t = (u64)blockIdx.x * WG + threadIdx.x; t < n / 2; t += stride) {
        const u64 low = t & (inc - 1), i = 2 * t - low;
        const bool up = (i & (2 * len)) == 0;
        const u64 a = A[i], b = A[i + inc];
        if (up == (a > b)) {
            A[i] = b;
            A[i + inc] = a;
        }
    }
}

// Merge: B[t] = max(A[i], A[i+k]) for the pair of runs holding t.
__global__ __launch_bounds__(WG) void k_bt_merge(const u64* __restrict__ A, u64 n, u64 k, u64* __restrict__ B) {
    const u64 stride = (u64)gridDim.x * WG;
    for (u64 t = (u64)blockIdx.x * WG + threadIdx.x; t < n / 2; t += stride) {
        const u64 i = 2 * t - (t & (k - 1));
        const u64 a = A[i], b = A[i + k];
        B[t] = a > b ? a : b;
    }
}

static unsigned grid_for(u64 items) { return (unsigned)std::max<u64>(1, std::min<u64>((items + WG - 1) / WG, 8192)); }

// Tile reduction levels with tiles of TILE composites: n -> k.
// Levels after the first read the sorted runs the previous one wrote.
template <u32 TILE, int NT>
u64* tile_levels(u64* A, u64* B, u64 n, u64 k, hipStream_t st) {
    int presorted = 0;
    while (n > TILE) {
        hipLaunchKernelGGL((k_bt_tile_reduce<TILE, NT>), dim3((unsigned)(n / TILE)), dim3(NT), 0, st, A, TILE,
                           (u32)k, presorted, B);
        n = (n / TILE) * k;
        std::swap(A, B);
        presorted = 1;
    }
    // last tile: n <= TILE elements -> k (also sorts when n == k)
    const u32 T = (u32)std::max<u64>(n, 1);
    hipLaunchKernelGGL((k_bt_tile_reduce<TILE, NT>), dim3(1), dim3(NT), 0, st, A, T, (u32)k, presorted, B);
    return B;
}

// Top-k (ascending, at A[0..k)) of the n composites in A (n a power of two, n >= k, k a power
// of two); B is scratch of n elements.  Returns the buffer holding the result.
u64* bitonic_topk(u64* A, u64* B, u64 n, u64 k, hipStream_t st) {
    if (k <= BT_SMALL / 4) return tile_levels<BT_SMALL, BTT_SMALL>(A, B, n, k, st);
    if (k <= BT / 2) return tile_levels<BT, BTT>(A, B, n, k, st);
    // global network
    hipLaunchKernelGGL(k_bt_sort_lds, dim3((unsigned)(n / BT)), dim3(BTT), 0, st, A, (u64)(BT / 2));
    for (u64 len = BT; len < k; len <<= 1) {
        for (u64 inc = len; inc >= BT; inc >>= 1)
            hipLaunchKernelGGL(k_bt_step, dim3(grid_for(n / 2)), dim3(WG), 0, st, A, n, len, inc);
        hipLaunchKernelGGL(k_bt_stage_lds, dim3((unsigned)(n / BT)), dim3(BTT), 0, st, A, len);
    }
    while (n > k) {
        hipLaunchKernelGGL(k_bt_merge, dim3(grid_for(n / 2)), dim3(WG), 0, st, A, n, k, B);
        n >>= 1;
        std::swap(A, B);
        const u64 len = k >> 1;
        for (u64 inc = len; inc >= BT; inc >>= 1)
            hipLaunchKernelGGL(k_bt_step, dim3(grid_for(n / 2)), dim3(WG), 0, st, A, n, len, inc);
        hipLaunchKernelGGL(k_bt_stage_lds, dim3((unsigned)(n / BT)), dim3(BTT), 0, st, A, len);
    }
    return A;
}

// ---------------------------------------------------------------- composites
// Delegates: (max << 32) | ~s, zero padding up to npad.
__global__ __launch_bounds__(WG) void k_bt_comp_delegates(const u32* __restrict__ dvals, u64 S, u64* __restrict__ A,
                                                         u64 npad) {
    const u64 stride = (u64)gridDim.x * WG;
    for (u64 i = (u64)blockIdx.x * WG + threadIdx.x; i < npad; i += stride)
        A[i] = i < S ? (((u64)dvals[i] << 32) | (u64)(0xFFFFFFFFu - (u32)i)) : 0ull;
}

// Keys of the subranges named by the top-k delegate composites D[0..k) (whole subranges,
// CreateConcatenateRange bitonic_com.cuh:72-98), or of V itself when D is null.
template <class K>
__global__ __launch_bounds__(WG) void k_bt_comp_keys(const u32* __restrict__ keys, u64 n, int alpha,
                                                     const u64* __restrict__ D, u64 nv, u64* __restrict__ A, u64 npad,
                                                     Ctrl* ctrl) {
    __shared__ u64 red[WG / WAVE];
    const u64 L = 1ull << alpha;
    const u64 stride = (u64)gridDim.x * WG;
    u64 valid = 0;
    for (u64 v = (u64)blockIdx.x * WG + threadIdx.x; v < npad; v += stride) {
        u64 c = 0;
        if (v < nv) {
            u64 g = v;
            if (D) {
                const u32 sid = 0xFFFFFFFFu - (u32)D[v >> alpha];
                g = ((u64)sid << alpha) + (v & (L - 1));
            }
            if (g < n) {
                c = ((u64)K::to_ord(keys[g]) << 32) | (u64)(0xFFFFFFFFu - (u32)g);
                ++valid;
            }
        }
        A[v] = c;
    }
    valid = wave_sum_u64(valid);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = valid;
    __syncthreads();
    if (threadIdx.x == 0 && ctrl) {
        u64 s = 0;
        for (int i = 0; i < WG / WAVE; ++i) s += red[i];
        if (s) atomicAdd((unsigned long long*)&ctrl->concat_size, (unsigned long long)s);
    }
}

// Threshold (k-th delegate) and the first-stage statistics.
__global__ void k_bt_first_done(const u64* __restrict__ D, u64 k, Ctrl* ctrl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctrl->first.result = (u32)(D[0] >> 32);
        ctrl->first.done = 1;
        ctrl->selected = k;
    }
}

// Final stage: k-th key, counts, and (index << 32 | key) pairs for the index-order sort.
__global__ __launch_bounds__(WG) void k_bt_final(const u64* __restrict__ R, u64 k, u64* __restrict__ P, u64 npad,
                                                 Ctrl* ctrl) {
    __shared__ u64 red[WG / WAVE];
    const u32 K0 = (u32)(R[0] >> 32);
    const u64 stride = (u64)gridDim.x * WG;
    u64 gt = 0;
    for (u64 i = (u64)blockIdx.x * WG + threadIdx.x; i < npad; i += stride) {
        if (i < k) {
            const u64 c = R[i];
            const u32 key = (u32)(c >> 32), idx = 0xFFFFFFFFu - (u32)c;
            gt += key > K0;
            P[i] = ((u64)idx << 32) | key;
        } else {
            P[i] = ~0ull;  // sorts after every real pair
        }
    }
    gt = wave_sum_u64(gt);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = gt;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 s = 0;
        for (int i = 0; i < WG / WAVE; ++i) s += red[i];
        if (s) atomicAdd((unsigned long long*)&ctrl->count_gt, (unsigned long long)s);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl->kth = K0;
        ctrl->second.result = K0;
        ctrl->second.done = 1;
    }
}

// Index-ordered outputs from the sorted pairs.
template <class K>
__global__ __launch_bounds__(WG) void k_bt_emit(const u64* __restrict__ P, u64 k, u32* __restrict__ out_keys,
                                                i64* __restrict__ out_idx, u32* __restrict__ d_kth, Ctrl* ctrl) {
    const u64 stride = (u64)gridDim.x * WG;
    for (u64 i = (u64)blockIdx.x * WG + threadIdx.x; i < k; i += stride) {
        const u64 p = P[i];
        if (out_keys) out_keys[i] = K::from_ord((u32)p);
        if (out_idx) out_idx[i] = (i64)(p >> 32);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl->ties = k - ctrl->count_gt;
        if (d_kth) *d_kth = K::from_ord(ctrl->kth);
    }
}

hipError_t launch_bitonic(const BitonicArgs& a, hipStream_t st) {
    const u64 k = a.k;
    u64* A = a.bufA;
    u64* B = a.bufB;
    const u64* top = nullptr;  // top-k delegate composites
    u64 nv, npad;
    if (a.alpha > 0) {
        // first top-k over the delegates (bitonic_firstTopk, bitonic_com.cuh:610-666)
        npad = a.npad1;
        hipLaunchKernelGGL(k_bt_comp_delegates, dim3(grid_for(npad)), dim3(WG), 0, st, a.dvals, a.S, A, npad);
        top = bitonic_topk(A, B, npad, k, st);
        hipLaunchKernelGGL(k_bt_first_done, dim3(1), dim3(1), 0, st, top, k, a.ctrl);
        // keep the k delegate composites while the second stage reuses the buffers
        u64* D = a.dsel;
        hipError_t e = hipMemcpyAsync(D, top, k * sizeof(u64), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
        top = D;
        nv = k << a.alpha;
    } else {
        nv = a.n;
    }
    // second top-k over the concatenation (bitonic, bitonic_com.cuh:669-728) or over V
    npad = a.npad2;
    if (a.f32)
        hipLaunchKernelGGL((k_bt_comp_keys<KeyF32>), dim3(grid_for(npad)), dim3(WG), 0, st, a.keys, a.n, a.alpha, top,
                           nv, A, npad, a.ctrl);
    else
        hipLaunchKernelGGL((k_bt_comp_keys<KeyU32>), dim3(grid_for(npad)), dim3(WG), 0, st, a.keys, a.n, a.alpha, top,
                           nv, A, npad, a.ctrl);
    const u64* R = bitonic_topk(A, B, npad, k, st);
    // outputs in index order: sort (index << 32 | key) ascending over the k survivors
    u64* P = (R == A) ? B : A;
    const u64 kp = k;  // k is a power of two
    hipLaunchKernelGGL(k_bt_final, dim3(grid_for(kp)), dim3(WG), 0, st, R, k, P, kp, a.ctrl);
    u64* Q = (P == A) ? B : A;
    const u64* sorted = bitonic_topk(P, Q, kp, kp, st);  // one run of k: a full ascending sort
    if (a.f32)
        hipLaunchKernelGGL((k_bt_emit<KeyF32>), dim3(grid_for(k)), dim3(WG), 0, st, sorted, k, (u32*)a.out_keys,
                           a.out_idx, (u32*)a.d_kth, a.ctrl);
    else
        hipLaunchKernelGGL((k_bt_emit<KeyU32>), dim3(grid_for(k)), dim3(WG), 0, st, sorted, k, (u32*)a.out_keys,
                           a.out_idx, (u32*)a.d_kth, a.ctrl);
    return hipGetLastError();
}

}  // namespace drtopk
