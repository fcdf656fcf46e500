// common.hpp -- shared device/host definitions of the MI355X Dr.Top-k library.
//
// Keys are handled in the "ordered" uint32 domain: uint32 keys as they are, float32
// keys by the sign flip  b ^ (b>>31 ? 0xFFFFFFFF : 0x80000000), which is monotone in
// the float total order (-0 < +0).  Every kernel converts at load and converts back
// only when it writes user-visible keys.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/drtopk.h"

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;
typedef u32 u32x4 __attribute__((ext_vector_type(4)));

namespace drtopk {

constexpr int WAVE = 64;       // CDNA wavefront width
constexpr int WG = 256;        // default workgroup: 4 waves
// Histograms of <= 256 bins are kept as HIST_REPS replicas in global memory: workgroup b
// flushes into replica b % HIST_REPS and the select step sums them, so a skewed digit
// does not serialise every flush on one L2 address.
constexpr int HIST_REPS = 16;
__host__ __device__ __forceinline__ int hist_reps(int bins) { return bins <= 256 ? HIST_REPS : 1; }

// ---------------------------------------------------------------- key transforms
struct KeyU32 {
    static __host__ __device__ __forceinline__ u32 to_ord(u32 b) { return b; }
    static __host__ __device__ __forceinline__ u32 from_ord(u32 o) { return o; }
};
struct KeyF32 {
    static __host__ __device__ __forceinline__ u32 to_ord(u32 b) { return b ^ ((b >> 31) ? 0xFFFFFFFFu : 0x80000000u); }
    static __host__ __device__ __forceinline__ u32 from_ord(u32 o) { return o ^ ((o >> 31) ? 0x80000000u : 0xFFFFFFFFu); }
};

// Prefix of length (32 - shleft) bits; a prefix of length 0 is 0.  Written out because
// v >> 32 is v on AMD (the shift count is taken mod 32).
__host__ __device__ __forceinline__ u32 prefix_of(u32 v, int shleft) {
    return shleft >= 32 ? 0u : (v >> shleft) << shleft;
}

// ---------------------------------------------------------------- select state
// One selection (radix digits or bucket levels) over an array of m keys, finding the
// key of top rank k (ascending rank m-k).  Lives in device memory so that every
// digit decision is made on the device, with no host round trip.
struct SelState {
    u32 flag;        // radix: accumulated prefix of the target
    int digit;       // radix: digit to process next (-1 when all are done)
    u32 lo, hi;      // bucket: current key range [lo, hi] (min/max reduced atomically first)
    u64 m;           // keys in the array
    u64 k;           // top rank searched: ascending rank m-k
    u64 below;       // keys strictly below the current prefix / range
    u64 kcount;      // keys in the current k-bucket
    u32 done;        // 1 once the selection is finished
    u32 passes;      // histogram passes run
    u32 result;      // threshold / k-th key (ordered bits) once done
    u32 need_find;   // bucket: the k-bucket holds one key that a find pass must fetch
    u32 nb;          // radix digit width; bucket: number of buckets
    u32 stop_digit;  // radix: finished once digit < stop_digit (3 on the beta path)
    u32 stop_on_one; // finished once the k-bucket holds a single key
    u32 max_passes;  // bucket: finished after this many levels (result = k-bucket lower bound)
    u32 pad0;
    u64 dense_limit; // radix: once digit < stop_digit, one more digit while more than this many
                     // keys lie at or above the prefix (0: never)
};

// First 16 bytes of every workspace (single-vector and batched): the device-side error words.
// `error` holds the bits of the call that last used the workspace; `sticky` accumulates the
// bits of every call since the last drtopk_workspace_check, which reads and clears it.  The
// first kernel of a call sets `magic` and clears `sticky` when `magic` does not hold WS_MAGIC
// yet (a workspace fresh from the allocator).
constexpr u64 WS_MAGIC = 0x4b504f5452445753ull;  // "SWDRTOPK"
struct WsHead {
    u64 magic;
    u32 error;
    u32 sticky;
};
static_assert(sizeof(WsHead) == 16, "workspace head layout");

// Resets a workspace head at the start of a call (one thread).
__device__ __forceinline__ void ws_head_begin(WsHead* h) {
    const u32 sticky = h->magic == WS_MAGIC ? h->sticky : 0u;
    h->magic = WS_MAGIC;
    h->error = 0;
    h->sticky = sticky;
}
// Records a failed device-side check in both words.
__device__ __forceinline__ void ws_fail(WsHead* h, u32 code) {
    atomicOr(&h->error, code);
    atomicOr(&h->sticky, code);
}

// Pipeline control block (one per call, at the start of the workspace).
struct Ctrl {
    WsHead head;          // error words (must stay first)
    SelState first;       // first top-k over delegates
    SelState second;      // second top-k over the concatenation (or over V)
    u64 selected;         // selected subranges
    u64 lonely_sub;       // lonely subranges (beta path)
    u64 lonely_el;        // lonely delegates appended (as the reference counts them)
    u64 lonely_dropped;   // of those, below the threshold and not written (FILTER)
    u64 concat_size;      // keys in the concatenation
    u64 count_gt;         // keys > k-th largest
    u64 ties;             // keys == k-th largest in the top-k set
    u32 kth;              // ordered bits of the k-th largest
    u32 pad[15];
};

enum : u32 {
    ERR_CONCAT_SHORT = DRTOPK_DEVICE_ERR_CONCAT_SHORT,  // concatenation smaller than k (not for valid input)
    ERR_CAPACITY = DRTOPK_DEVICE_ERR_CAPACITY           // a buffer bound was exceeded
};
__device__ __forceinline__ void ctrl_fail(Ctrl* c, u32 code) { ws_fail(&c->head, code); }

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ u32 wave_max_u32(u32 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (u32)__shfl_xor((int)v, o, WAVE));
    return v;
}
__device__ __forceinline__ u32 wave_min_u32(u32 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, (u32)__shfl_xor((int)v, o, WAVE));
    return v;
}
__device__ __forceinline__ u64 shfl_xor_u64(u64 v, int o) {
    u32 lo = (u32)__shfl_xor((int)(u32)v, o, WAVE);
    u32 hi = (u32)__shfl_xor((int)(u32)(v >> 32), o, WAVE);
    return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 shfl_up_u64(u64 v, int o) {
    u32 lo = (u32)__shfl_up((int)(u32)v, o, WAVE);
    u32 hi = (u32)__shfl_up((int)(u32)(v >> 32), o, WAVE);
    return ((u64)hi << 32) | lo;
}
// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ u64 wave_incl_scan_u64(u64 v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        u64 t = shfl_up_u64(v, o);
        if (lane >= o) v += t;
    }
    return v;
}
// 32-bit inclusive prefix sum over the 64 lanes of a wave with DPP: row_shr 1/2/4/8 scans each
// row of 16 lanes, row_bcast:15 adds row 0's (row 2's) total to row 1 (row 3), row_bcast:31
// adds lane 31's running total to rows 2 and 3.  Source lanes outside the row read the 0 `old`.
__device__ __forceinline__ u32 wave_incl_scan_u32(u32 v) {
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

// u64 sum over the 64 lanes of a wave (result in every lane).
__device__ __forceinline__ u64 wave_sum_u64(u64 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += shfl_xor_u64(v, o);
    return v;
}

// Exclusive prefix of a per-thread u64 count across the workgroup (blockDim.x == NT,
// NT a multiple of 64).  Returns the exclusive prefix; *total gets the block sum.
// `scratch` needs NT/64 + 1 u64 slots of LDS.  Contains __syncthreads.
template <int NT>
__device__ __forceinline__ u64 block_excl_scan_u64(u64 v, u64* scratch, u64* total) {
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    u64 inc = wave_incl_scan_u64(v);
    if (lane == WAVE - 1) scratch[w] = inc;
    __syncthreads();
    u64 off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / WAVE; ++i) {
        u64 s = scratch[i];
        off += (i < w) ? s : 0;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// u32 form of block_excl_scan_u64 (the block sum must fit in 32 bits); `scratch` needs NT/64
// u32 slots of LDS.  Contains __syncthreads.
template <int NT>
__device__ __forceinline__ u32 block_excl_scan_u32(u32 v, u32* scratch, u32* total) {
    const int w = threadIdx.x / WAVE;
    const u32 inc = wave_incl_scan_u32(v);
    if ((threadIdx.x & (WAVE - 1)) == WAVE - 1) scratch[w] = inc;
    __syncthreads();
    u32 off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / WAVE; ++i) {
        const u32 s = scratch[i];
        off += (i < w) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// Whole-wave histogram update: lanes with `active` add bin `b`.  The bins of the first PEEL
// active lanes are peeled with one LDS atomic each (all lanes sharing that bin at once), the
// rest go one by one; skewed keys then cost one atomic per wave, not one per lane.
template <int PEEL = 2>
__device__ __forceinline__ void wave_hist_add(u32* lhist, bool active, u32 b) {
    const int lane = threadIdx.x & (WAVE - 1);
    bool rem = active;
#pragma unroll
    for (int it = 0; it < PEEL; ++it) {
        const u64 act = __ballot(rem);
        if (!act) return;
        const int src = __ffsll((unsigned long long)act) - 1;
        const u32 pb = (u32)__shfl((int)b, src, WAVE);
        const bool mine = rem && b == pb;
        const u64 same = __ballot(mine);
        if (lane == src) atomicAdd(&lhist[pb], (u32)__popcll(same));
        if (mine) rem = false;
    }
    if (rem) atomicAdd(&lhist[b], 1u);
}

// ---------------------------------------------------------------- beta concatenation counts
// The beta-path concatenation leaves one record per workgroup: (selected subranges, lonely
// subranges, lonely delegates appended, lonely delegates dropped).  Thousands of workgroups
// adding them to four Ctrl words would queue on one L2 line at the kernel's end (40 us at 2048
// workgroups measured); the pass that totals the segment counts sums the records instead.
// Every thread of an NT-thread block calls this; thread 0 writes the Ctrl fields.
template <int NT>
__device__ __forceinline__ void beta_part_total(const u64* __restrict__ part, u64 npart, Ctrl* ctrl,
                                                u64 (*lds)[NT / WAVE]) {
    // one workgroup reads every record: unrolled so that a thread's loads are all in flight
    // together (a loop that waits for each record costs a memory latency per record)
    u64 s[4] = {0, 0, 0, 0};
    const ulonglong2* p2 = (const ulonglong2*)part;  // record i: p2[2i], p2[2i + 1] (32-byte aligned)
#pragma unroll 8
    for (u64 i = threadIdx.x; i < npart; i += NT) {
        const ulonglong2 a = p2[2 * i], b = p2[2 * i + 1];
        s[0] += a.x;
        s[1] += a.y;
        s[2] += b.x;
        s[3] += b.y;
    }
    const int w = threadIdx.x / WAVE;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        s[q] = wave_sum_u64(s[q]);
        if ((threadIdx.x & (WAVE - 1)) == 0) lds[q][w] = s[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 t[4] = {0, 0, 0, 0};
        for (int i = 0; i < NT / WAVE; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) t[q] += lds[q][i];
        ctrl->selected = t[0];
        ctrl->lonely_sub = t[1];
        ctrl->lonely_el = t[2];
        ctrl->lonely_dropped = t[3];
    }
}

// ---------------------------------------------------------------- packed counts
// Pack two 31-bit counts for the look-back payload.
__device__ __forceinline__ u64 pack2(u64 hi, u64 lo) { return (hi << 31) | lo; }
__device__ __forceinline__ u64 unpack_hi(u64 p) { return (p >> 31) & 0x7FFFFFFFull; }
__device__ __forceinline__ u64 unpack_lo(u64 p) { return p & 0x7FFFFFFFull; }

}  // namespace drtopk
