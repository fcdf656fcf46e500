// common.hpp -- shared device/host definitions of the MI355X Dr.Top-k library.
//
// Keys are handled in the "ordered" uint32 domain: uint32 keys as they are, float32
// keys by the sign flip  b ^ (b>>31 ? 0xFFFFFFFF : 0x80000000), which is monotone in
// the float total order (-0 < +0).  Every kernel converts at load and converts back
// only when it writes user-visible keys.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;
typedef u32 u32x4 __attribute__((ext_vector_type(4)));

namespace drtopk {

constexpr int WAVE = 64;       // CDNA wavefront width
constexpr int WG = 256;        // default workgroup: 4 waves

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
};

// Pipeline control block (one per call, at the start of the workspace).
struct Ctrl {
    SelState first;       // first top-k over delegates
    SelState second;      // second top-k over the concatenation (or over V)
    u64 selected;         // selected subranges
    u64 lonely_sub;       // lonely subranges (beta path)
    u64 lonely_el;        // lonely delegates appended
    u64 concat_size;      // keys in the concatenation
    u64 count_gt;         // keys > k-th largest
    u64 ties;             // keys == k-th largest in the top-k set
    u32 kth;              // ordered bits of the k-th largest
    u32 error;            // nonzero: a device-side check failed
    u32 tickets[8];       // dynamic tile tickets of the look-back kernels
    u32 pad[6];
};

enum : u32 {
    ERR_LOOKBACK_SPIN = 1u,   // a look-back wait ran out of its spin budget
    ERR_CONCAT_SHORT = 2u,    // concatenation smaller than k (cannot happen for valid input)
    ERR_CAPACITY = 4u         // a buffer bound was exceeded
};

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
__device__ __forceinline__ u32 wave_incl_scan_u32(u32 v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        u32 t = (u32)__shfl_up((int)v, o, WAVE);
        if (lane >= o) v += t;
    }
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

// ---------------------------------------------------------------- decoupled look-back
// Tiles are taken in ticket order (an atomic counter), so a tile only ever waits for
// tiles that were taken before it, whose workgroups are already running: the wait
// always ends.  As a second guard every wait has a spin budget; running out sets
// Ctrl::error and continues, so no kernel can hang.
// status word: bits 62..63 = 0 not ready, 1 aggregate only, 2 inclusive prefix;
// bits 0..61 = payload (a count, or two 31-bit counts packed hi<<31|lo).
constexpr u64 LB_AGG = 1ull << 62, LB_INC = 2ull << 62, LB_PAYLOAD = (1ull << 62) - 1;
constexpr u32 LB_SPIN_LIMIT = 1u << 22;

__device__ __forceinline__ u64 lb_load(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(u64* p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by ONE thread of the tile.  Returns the exclusive prefix of `aggregate`.
__device__ __forceinline__ u64 lookback(u64* status, u32 tile, u64 aggregate, u32* err) {
    if (tile == 0) {
        lb_store(&status[0], LB_INC | aggregate);
        return 0;
    }
    lb_store(&status[tile], LB_AGG | aggregate);
    u64 excl = 0;
    long t = (long)tile - 1;
    u32 spins = 0;
    while (t >= 0) {
        u64 s = lb_load(&status[t]);
        while ((s >> 62) == 0) {
            if (++spins > LB_SPIN_LIMIT) {
                atomicOr(err, ERR_LOOKBACK_SPIN);
                s = LB_INC;  // give up: treat as an inclusive zero
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            s = lb_load(&status[t]);
        }
        excl += s & LB_PAYLOAD;
        if ((s >> 62) == 2) break;
        --t;
    }
    lb_store(&status[tile], LB_INC | (excl + aggregate));
    return excl;
}

// Ticket for the next tile (thread 0 draws, the block reads it from LDS).
__device__ __forceinline__ u32 next_tile(u32* ticket, u32* lds_slot) {
    if (threadIdx.x == 0) *lds_slot = atomicAdd(ticket, 1u);
    __syncthreads();
    u32 t = *lds_slot;
    __syncthreads();
    return t;
}

// Pack two 31-bit counts for the look-back payload.
__device__ __forceinline__ u64 pack2(u64 hi, u64 lo) { return (hi << 31) | lo; }
__device__ __forceinline__ u64 unpack_hi(u64 p) { return (p >> 31) & 0x7FFFFFFFull; }
__device__ __forceinline__ u64 unpack_lo(u64 p) { return p & 0x7FFFFFFFull; }

}  // namespace drtopk
