// drtopk.hpp -- C++ host interface over the C ABI (drtopk.h), mirroring the reference's
// header-only entry points by name and argument meaning so that a caller of the
// reference can switch with a one-line change per call site:
//
//   reference (CUDA, baseline+filter+beta+shuffle/radixselect.cuh)      here
//   radix_select(vec_d, n, k, num_bucket, TopKElement, NBits, CurDigit)  drtopk::radix_select(...)      :1473
//   sample_radix_select(vec_d, n, k, num_bucket, TopKElement, NBits,     drtopk::sample_radix_select(...) :1783
//       CurDigit, NSubranges, SubRangesize, alpha, t_sample, t_first,
//       t_second, t_concat, <buffers>, beta, defaultContribution, ...)
//   sample_bucket_select(vec_d, n, k, num_bucket, TopKElement, ...)     drtopk::sample_bucket_select(...) (bucket_select/bucketselect_combined.cuh:1280)
//   bucket_select(vec_d, n, k, num_bucket, TopKElement)                  drtopk::bucket_select(...)       (:473)
//   bitonic_firstTopk + bitonic (bitonic/bitonic_com.cuh:610, :669)      drtopk::bitonic_topk(...)
//
// Differences, all deliberate:
//  - The reference's caller-allocated buffers (Max_d, SubrangeId_d, SelectedSubrangeId_d,
//    ConcatenatedRange_d, counters; main.cu:205-250) become one drtopk::Workspace, sized by the
//    caller before the calls (Workspace::reserve_for); the calls themselves never allocate and
//    return DRTOPK_ERR_WORKSPACE_TOO_SMALL when it is too small.
//  - Errors are returned as drtopk_status (the reference prints and exit()s, radixselect.cuh:20-28).
//  - The reference reports the k-th element only; the optional TopKElements_d / TopKIndices_d
//    receive the top-k keys and indices (index order, lowest index wins ties).
//  - bucket_select's k is the ascending rank n-k, as in the reference (bucketselect_combined.cuh:1355);
//    every other entry point takes the top rank k.
//  - Phase times are device times of the stream (HIP events), in seconds like wtime().
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "drtopk.h"

namespace drtopk {

// Device scratch owned by the caller.  reserve / reserve_for allocate (hipMalloc); the select
// calls below only use what has been reserved.
class Workspace {
   public:
    Workspace() = default;
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;
    ~Workspace() {
        if (ptr_) (void)hipFree(ptr_);
    }
    drtopk_status reserve(size_t bytes) {
        if (bytes <= size_) return DRTOPK_OK;
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        size_ = 0;
        if (hipMalloc(&ptr_, bytes) != hipSuccess) return DRTOPK_ERR_HIP;
        size_ = bytes;
        return DRTOPK_OK;
    }
    // Reserve enough for any call of the family `algo` on n keys of type data_t with this k:
    // every alpha (auto, 0 .. log2 n), beta (1 .. 4), digit width (1 .. 16) and bucket count up
    // to 4096, plus the slot the k-th key is returned through.
    template <typename data_t>
    drtopk_status reserve_for(drtopk_algo algo, uint64_t n, uint64_t k);
    void* data() const { return ptr_; }
    size_t size() const { return size_; }

   private:
    void* ptr_ = nullptr;
    size_t size_ = 0;
};

template <typename data_t>
struct dtype_of;
template <>
struct dtype_of<unsigned int> { static constexpr drtopk_dtype value = DRTOPK_U32; };
template <>
struct dtype_of<float> { static constexpr drtopk_dtype value = DRTOPK_F32; };

template <typename data_t>
drtopk_status Workspace::reserve_for(drtopk_algo algo, uint64_t n, uint64_t k) {
    size_t need = 0;
    bool any = false;
    for (int nb = 1; nb <= 16; nb *= 2)
        for (int beta = 1; beta <= 4; ++beta)
            for (int alpha = -1; alpha <= 31; ++alpha) {
                drtopk_params p;
                drtopk_params_default(&p);
                p.alpha = alpha;
                p.beta = beta;
                p.digit_bits = nb;
                p.num_buckets = 4096;
                size_t b = 0;
                if (drtopk_workspace_size(algo, dtype_of<data_t>::value, n, k, &p, &b) == DRTOPK_OK) {
                    need = b > need ? b : need;
                    any = true;
                }
            }
    if (!any) return DRTOPK_ERR_INVALID_ARG;
    return reserve((need + 255) / 256 * 256 + 256);
}

namespace detail {
template <typename data_t>
inline drtopk_status run(drtopk_algo algo, const data_t* vec_d, uint64_t n, uint64_t k, const drtopk_params& p,
                         data_t& TopKElement, Workspace& ws, data_t* keys_d, int64_t* idx_d, drtopk_stats* st,
                         hipStream_t stream) {
    size_t bytes = 0;
    drtopk_status s = drtopk_workspace_size(algo, dtype_of<data_t>::value, n, k, &p, &bytes);
    if (s != DRTOPK_OK) return s;
    const size_t kth_off = (bytes + 255) / 256 * 256;  // the k-th key lives after the workspace
    if (!ws.data() || ws.size() < kth_off + 256) return DRTOPK_ERR_WORKSPACE_TOO_SMALL;
    data_t* kth_d = (data_t*)((char*)ws.data() + kth_off);
    drtopk_stats local;
    s = drtopk_select(algo, dtype_of<data_t>::value, vec_d, n, k, &p, ws.data(), kth_off, kth_d, keys_d, idx_d,
                      st ? st : &local, stream);
    if (s == DRTOPK_OK && hipMemcpyAsync(&TopKElement, kth_d, sizeof(data_t), hipMemcpyDeviceToHost, stream) != hipSuccess)
        s = DRTOPK_ERR_HIP;
    if (hipStreamSynchronize(stream) != hipSuccess && s == DRTOPK_OK) s = DRTOPK_ERR_HIP;
    return s;
}
inline int bits_to_digit(int NBits) { return NBits > 0 ? 32 / NBits - 1 : -1; }
}  // namespace detail

// Plain MSD radix select over V (the alpha == 0 path, radixselect.cuh:1473-1543).
template <typename data_t, typename index_t>
drtopk_status radix_select(const data_t* vec_d, index_t num_element, index_t k, index_t num_bucket, data_t& TopKElement,
                           int NBitsperDigit, int CurrentDigit, Workspace& ws, data_t* TopKElements_d = nullptr,
                           int64_t* TopKIndices_d = nullptr, hipStream_t stream = nullptr) {
    if ((index_t)1 << NBitsperDigit != num_bucket || CurrentDigit != detail::bits_to_digit(NBitsperDigit))
        return DRTOPK_ERR_INVALID_ARG;
    drtopk_params p;
    drtopk_params_default(&p);
    p.alpha = 0;
    p.digit_bits = NBitsperDigit;
    return detail::run(DRTOPK_ALGO_RADIX, vec_d, (uint64_t)num_element, (uint64_t)k, p, TopKElement, ws,
                       TopKElements_d, TopKIndices_d, nullptr, stream);
}

// Dr.Top-k radix select (radixselect.cuh:1783-1894).  defaultContribution selects the
// one-maximum-per-subrange path; otherwise the beta path with `beta` delegates per subrange
// (the reference sets defaultContribution = alpha > 5, main.cu:87-89).  `flags` picks the
// filter / beta / shuffle options of the experiment folders (DRTOPK_FLAGS_DEFAULT = final).
template <typename data_t, typename index_t>
drtopk_status sample_radix_select(const data_t* vec_d, index_t num_element, index_t k, index_t num_bucket,
                                  data_t& TopKElement, int NBits, int CurrentDigit, index_t NSubranges,
                                  index_t SubRangesize, index_t alpha, double& timeforMaxsample,
                                  double& timeforFirstTopK, double& timeForsecondTopK, double& timeforConcatenation,
                                  int beta, bool defaultContribution, Workspace& ws,
                                  unsigned flags = DRTOPK_FLAGS_DEFAULT, data_t* TopKElements_d = nullptr,
                                  int64_t* TopKIndices_d = nullptr, hipStream_t stream = nullptr) {
    if ((index_t)1 << NBits != num_bucket || CurrentDigit != detail::bits_to_digit(NBits)) return DRTOPK_ERR_INVALID_ARG;
    if (alpha < 1 || ((index_t)1 << alpha) != SubRangesize ||
        NSubranges != (num_element + SubRangesize - 1) / SubRangesize)
        return DRTOPK_ERR_INVALID_ARG;
    drtopk_params p;
    drtopk_params_default(&p);
    p.alpha = (int32_t)alpha;
    p.digit_bits = NBits;
    p.beta = defaultContribution ? 1 : beta;
    p.flags = (flags & ~DRTOPK_FLAG_BETA) | (defaultContribution ? 0u : DRTOPK_FLAG_BETA) | DRTOPK_FLAG_TIMING;
    drtopk_stats st;
    drtopk_status s = detail::run(DRTOPK_ALGO_RADIX, vec_d, (uint64_t)num_element, (uint64_t)k, p, TopKElement, ws,
                                  TopKElements_d, TopKIndices_d, &st, stream);
    timeforMaxsample = st.phase_ms[0] * 1e-3;
    timeForsecondTopK = (st.phase_ms[3] + st.phase_ms[4]) * 1e-3;
    timeforFirstTopK = st.phase_ms[1] * 1e-3;
    timeforConcatenation = st.phase_ms[2] * 1e-3;
    return s;
}

// Plain bucket select over V (bucketselect_combined.cuh:473-568): k is the ascending rank.
template <typename data_t, typename index_t>
drtopk_status bucket_select(const data_t* vec_d, index_t num_element, index_t k, index_t num_bucket,
                            data_t& TopKElement, Workspace& ws, hipStream_t stream = nullptr) {
    if (k < 0 || k >= num_element) return DRTOPK_ERR_INVALID_ARG;
    drtopk_params p;
    drtopk_params_default(&p);
    p.alpha = 0;
    p.num_buckets = (int32_t)num_bucket;
    return detail::run(DRTOPK_ALGO_BUCKET, vec_d, (uint64_t)num_element, (uint64_t)(num_element - k), p, TopKElement,
                       ws, (data_t*)nullptr, nullptr, nullptr, stream);
}

// Dr.Top-k bucket select (bucketselect_combined.cuh:1280-1358); alpha / beta as in bucket_select/main.cu:66-90.
template <typename data_t, typename index_t>
drtopk_status sample_bucket_select(const data_t* vec_d, index_t num_element, index_t k, index_t num_bucket,
                                   data_t& TopKElement, index_t alpha, int beta, bool defaultContribution,
                                   Workspace& ws, data_t* TopKElements_d = nullptr, int64_t* TopKIndices_d = nullptr,
                                   hipStream_t stream = nullptr) {
    drtopk_params p;
    drtopk_params_default(&p);
    p.alpha = (int32_t)alpha;
    p.num_buckets = (int32_t)num_bucket;
    p.beta = defaultContribution ? 1 : beta;
    p.flags = DRTOPK_FLAG_FILTER | DRTOPK_FLAG_SHUFFLE | (defaultContribution ? 0u : DRTOPK_FLAG_BETA);
    return detail::run(DRTOPK_ALGO_BUCKET, vec_d, (uint64_t)num_element, (uint64_t)k, p, TopKElement, ws,
                       TopKElements_d, TopKIndices_d, nullptr, stream);
}

// Dr.Top-k bitonic (bitonic_firstTopk + bitonic, bitonic_com.cuh:610-728, driven as in
// sampleBitonic.cu:283-308); alpha == 0 is the plain bitonic top-k (NormalBitonic = 1).
// k must be a power of two (sampleBitonic.cu:198-202).
template <typename data_t, typename index_t>
drtopk_status bitonic_topk(const data_t* vec_d, index_t num_element, index_t k, index_t alpha, data_t& TopKElement,
                           Workspace& ws, data_t* TopKElements_d = nullptr, int64_t* TopKIndices_d = nullptr,
                           hipStream_t stream = nullptr) {
    drtopk_params p;
    drtopk_params_default(&p);
    p.alpha = (int32_t)alpha;
    return detail::run(DRTOPK_ALGO_BITONIC, vec_d, (uint64_t)num_element, (uint64_t)k, p, TopKElement, ws,
                       TopKElements_d, TopKIndices_d, nullptr, stream);
}

}  // namespace drtopk
