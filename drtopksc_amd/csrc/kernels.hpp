// kernels.hpp -- host-side launchers of the device kernels (one per .hip file group).
#pragma once

#include "common.hpp"

namespace drtopk {

struct SampleArgs {
    const u32* keys;   // raw keys (uint32 or float32 bits)
    u64 n;
    int alpha, beta;
    int f32;           // keys are float32
    int staged;        // LDS-staged variant (alpha <= 5)
    int posmode;       // 0 none, 1 absolute positions, 2 packed offsets (beta >= 2)
    int fuse;          // 0 none, 1 radix top-digit histogram, 2 min/max
    u32* dvals;        // [beta*S] ordered delegates
    void* dpos;        // positions / packed offsets
    u32* ghist;        // fused histogram (fuse == 1)
    int hshift, bins;  // fused histogram: bin = delegate >> hshift, bins bins
    Ctrl* ctrl;
};
hipError_t launch_sample(const SampleArgs& a, hipStream_t st);

// Selection machinery -------------------------------------------------------------------
// mode 0: radix digits (prefix match + digit); mode 1: bucket levels (range + linear map)
hipError_t launch_init(Ctrl* ctrl, const Ctrl& init, u32* zero_begin, u64 zero_words, hipStream_t st);
hipError_t launch_hist(int mode, const u32* arr, SelState* sel, u32* ghist, int bins, unsigned grid,
                       hipStream_t st);
hipError_t launch_hist_raw(int mode, int f32, const u32* keys, SelState* sel, u32* ghist, int bins,
                           unsigned grid, hipStream_t st);
hipError_t launch_select(int mode, u32* ghist, int bins, SelState* sel, Ctrl* ctrl, hipStream_t st);
hipError_t launch_bucket_begin(SelState* sel, Ctrl* ctrl, int from_ctrl_minmax, hipStream_t st);
hipError_t launch_minmax(const u32* arr, SelState* sel, Ctrl* ctrl, unsigned grid, hipStream_t st);
hipError_t launch_minmax_raw(int f32, const u32* keys, SelState* sel, Ctrl* ctrl, unsigned grid,
                             hipStream_t st);
hipError_t launch_find(const u32* arr, int raw_f32, SelState* sel, unsigned grid, hipStream_t st);

struct RangeArgs {
    const u32* dvals;
    u64 S;             // subranges
    u64 k;
    int mode;          // first-top-k mode (0 radix prefix compare, 1 full compare)
    SelState* sel;     // first top-k state
    Ctrl* ctrl;
    u32* sel_ids;      // [k] selected subrange ids, ascending
    u64* status;       // look-back status, one per tile
    u32* ticket;
};
hipError_t launch_range_default(const RangeArgs& a, hipStream_t st);

struct GatherArgs {
    const u32* keys;   // raw keys
    u64 n;
    int alpha;
    int f32;
    int filter;
    const u32* sel_ids;
    u64 nsel;          // == k on the default path
    Ctrl* ctrl;
    u32* ckeys;        // concatenation (ordered keys)
    u32* cidx;         // concatenation indices or nullptr
    u64 cap;           // capacity of ckeys/cidx
    u64* status;
    u32* ticket;
    u32* ghist2;       // optional fused top-digit histogram of the concatenation (second top-k)
    int hshift2, bins2;
};
hipError_t launch_gather_default(const GatherArgs& a, hipStream_t st);

struct BetaGatherArgs {
    const u32* keys;
    u64 n;
    int alpha, beta, f32, filter;
    const u32* dvals;
    const void* doff;  // packed offsets or nullptr
    Ctrl* ctrl;
    u32* ckeys;
    u32* cidx;
    u64 cap;
    u64* status;
    u32* ticket;
    u32* ghist2;
    int hshift2, bins2;
};
hipError_t launch_range_gather_beta(const BetaGatherArgs& a, hipStream_t st);

struct ExtractArgs {
    const u32* ckeys;  // ordered keys (or raw keys when direct)
    const u32* cidx;   // indices or nullptr (direct: position)
    int direct;        // 1: extract straight from V (plain path), keys are raw
    int f32;
    u64 cap;           // bound on the array length (for the grid)
    u64 k;
    SelState* sel;     // second top-k state (k-th largest)
    Ctrl* ctrl;
    void* out_keys;
    i64* out_idx;
    u64* status;
    u32* ticket;
};
hipError_t launch_extract(const ExtractArgs& a, hipStream_t st);
hipError_t launch_finalize(SelState* sel, Ctrl* ctrl, int f32, void* d_kth, u64 k, hipStream_t st);
hipError_t launch_to_user(const u32* ord, u64 m, int f32, void* out, hipStream_t st);

}  // namespace drtopk
