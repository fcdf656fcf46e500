// kernels.hpp -- host-side launchers of the device kernels (one per .hip file group).
#pragma once

#include "common.hpp"

namespace drtopk {

// A segmented array: segment s holds len(s) elements at [s*segcap, s*segcap + len(s)).
// len(s) = cnt[s], or min(segcap, n - s*segcap) when cnt is null (a flat array in tiles).
struct Seg {
    const u32* keys;
    const u32* idx;
    const u32* cnt;
    u64 segcap;
    u64 nseg;
    u64 n;
};

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
    // beta 2: when set, dvals[s] holds the first delegate and the 16-bit word s of dtop the top
    // byte of the second (all a beta-path threshold on the top 8-bit digit needs of it) and,
    // above it, the first's offset in the subrange (0xFF without positions); dpos unused
    u8* dtop = nullptr;
};
hipError_t launch_sample(const SampleArgs& a, hipStream_t st);

// Compute units of the current device (cached per device).
unsigned device_cus();

// Selection machinery -------------------------------------------------------------------
// mode 0: radix digits (prefix match + digit); mode 1: bucket levels (range + linear map)
hipError_t launch_init(Ctrl* ctrl, const Ctrl& init, u32* zero_begin, u64 zero_words, hipStream_t st);
hipError_t launch_hist(int mode, const u32* arr, SelState* sel, u32* ghist, int bins, unsigned grid,
                       hipStream_t st);
hipError_t launch_hist_raw(int mode, int f32, const u32* keys, SelState* sel, u32* ghist, int bins,
                           unsigned grid, hipStream_t st);
hipError_t launch_hist_seg(int mode, const Seg& seg, SelState* sel, u32* ghist, int bins, unsigned grid,
                           hipStream_t st);
// segcnt / nseg: first pass over a segmented concatenation; the kernel then also sets its size
// (concat_size, sel->m and the too-short error) from the segment counts, and with `part` first
// sums the beta concatenation's npart per-workgroup records (beta_part_total).
hipError_t launch_select(int mode, u32* ghist, int bins, SelState* sel, Ctrl* ctrl, hipStream_t st,
                         const u32* segcnt = nullptr, u64 nseg = 0, const u64* part = nullptr, u64 npart = 0);
hipError_t launch_bucket_begin(SelState* sel, Ctrl* ctrl, int from_ctrl_minmax, hipStream_t st);
hipError_t launch_minmax(const u32* arr, SelState* sel, Ctrl* ctrl, unsigned grid, hipStream_t st);
hipError_t launch_minmax_raw(int f32, const u32* keys, SelState* sel, Ctrl* ctrl, unsigned grid,
                             hipStream_t st);
hipError_t launch_minmax_seg(const Seg& seg, SelState* sel, unsigned grid, hipStream_t st);
hipError_t launch_find(const u32* arr, int raw_f32, SelState* sel, unsigned grid, hipStream_t st);
hipError_t launch_find_seg(const Seg& seg, SelState* sel, unsigned grid, hipStream_t st);

// compact.hip -------------------------------------------------------------------------
// Ordered compactions run as count pass -> single-workgroup scan -> write pass, with no
// waiting between workgroups.
hipError_t launch_scan(const u64* in, u64 nseg, u64* out_excl, u64* total, hipStream_t st);

// FindTopKRanges + FillWithKElementsID: exactly k subrange ids, ascending.
struct RangeArgs {
    const u32* dvals;
    u64 S;
    u64 k;
    int fullcmp;       // 1: compare full keys (bucket), 0: radix prefix of the first top-k
    SelState* sel;     // first top-k state
    Ctrl* ctrl;
    u32* sel_ids;      // [k]
    u64* cnt;          // per-tile packed (greater, equal) counts
    u64* off;          // their exclusive prefix
};
hipError_t launch_range(const RangeArgs& a, hipStream_t st);
__host__ __device__ inline u64 range_tiles(u64 S) { return (S + 4095) / 4096; }

// Default-path concatenation, staged per 4096-key tile of the selected subranges.
struct GatherArgs {
    const u32* keys;
    u64 n;
    int alpha, f32, filter;
    const u32* sel_ids;
    u64 nsel;
    Ctrl* ctrl;
    u32* skeys;        // segment t at [t*GATHER_TILE, ...)
    u32* sidx;         // or nullptr
    u32* cnt;          // [nseg]
    u32* ghist2;       // fused top-digit histogram of the concatenation, or nullptr
    int hshift2, bins2;
    u64 seg_limit = 0; // entries a segment may hold (0: its capacity); beyond it the error is raised
};
hipError_t launch_gather(const GatherArgs& a, hipStream_t st);
u64 gather_segcap(u64 V);   // segment capacity / count for V = nsel << alpha virtual keys
u64 gather_nseg(u64 V);

// Beta-path range selection + concatenation, staged per 1024-subrange tile.
struct BetaGatherArgs {
    const u32* keys;
    u64 n;
    int alpha, beta, f32, filter;  // filter: bit 0 keep keys >= threshold, bit 1 drop lonely delegates below it
    const u32* dvals;
    const void* doff;
    Ctrl* ctrl;
    u32* skeys;        // segment t at [t*segcap, ...)
    u32* sidx;
    u32* cnt;
    u32* ghist2;
    int hshift2, bins2;
    const u8* dtop = nullptr;  // beta 2: delegates in the sampler's dtop form (dvals[s] first, 16-bit dtop words)
    u64 seg_limit = 0;         // entries a segment may hold (0: its capacity)
    u64* part = nullptr;       // [beta_grid] per-workgroup records (selected, lonely, lonely delegates, dropped)
};
constexpr int BETA_SUBS_PER_THREAD = 4;  // 4 measured faster than 8 (fewer VGPRs, more resident workgroups)
constexpr u64 BETA_TILE_SUBS = 256 * BETA_SUBS_PER_THREAD;  // subranges per chunk of the beta gather
hipError_t launch_beta_gather(const BetaGatherArgs& a, hipStream_t st);
u64 beta_segcap(u64 S, int alpha);
u64 beta_nseg(u64 S, int alpha);
u64 beta_grid(u64 S, int alpha);  // workgroups of the beta gather = records it writes

// concatenation size = sum of segment counts -> ctrl->concat_size, second.m (with `part`: the
// beta concatenation's per-workgroup records summed into the Ctrl counts first)
hipError_t launch_seg_total(const u32* cnt, u64 nseg, Ctrl* ctrl, hipStream_t st, const u64* part = nullptr,
                            u64 npart = 0);

// findTopk + FillTopk: keys above the k-th largest, then the lowest-index ties, index order.
struct ExtractArgs {
    Seg seg;           // concatenation, or V itself (cnt == nullptr, raw keys)
    int direct;        // 1: seg is V (raw keys, index = position)
    int f32;
    u64 k;
    SelState* sel;     // second top-k state
    Ctrl* ctrl;
    void* out_keys;
    i64* out_idx;
    u64* cnt;          // [nseg] packed (greater, equal)
    u64* off;
    bool finalize = false;  // also do k_finalize's work (k-th key out, greater / tie counts)
    void* d_kth = nullptr;
};
constexpr u64 DIRECT_TILE = 16384;
hipError_t launch_extract(const ExtractArgs& a, hipStream_t st);
hipError_t launch_finalize(SelState* sel, Ctrl* ctrl, int f32, void* d_kth, u64 k, hipStream_t st);

// bitonic.hip ---------------------------------------------------------------------------
// Bitonic family: first top-k over the delegates, whole-subrange concatenation, second
// top-k, index-ordered outputs.  alpha == 0: one top-k over V.
struct BitonicArgs {
    const u32* keys;
    u64 n;
    int alpha, f32;
    u64 k, S;
    const u32* dvals;  // delegates (alpha > 0)
    Ctrl* ctrl;
    u64* bufA;         // max(npad1, npad2) composites each
    u64* bufB;
    u64* dsel;         // k composites
    u64 npad1, npad2;  // next powers of two of S and of k << alpha (or n)
    void* out_keys;
    i64* out_idx;
    void* d_kth;
};
hipError_t launch_bitonic(const BitonicArgs& a, hipStream_t st);
hipError_t launch_to_user(const u32* ord, u64 m, int f32, void* out, hipStream_t st);

}  // namespace drtopk
