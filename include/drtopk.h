/*
 * drtopk.h -- C ABI of the MI355X-native Dr.Top-k library (libdrtopk.so).
 *
 * Delegate-centric top-k selection (Dr.Top-k) over one device-resident key
 * vector: a delegate pass keeps the maximum (or top-beta) of every subrange of
 * 2^alpha keys, a first top-k over the delegates yields a threshold and the
 * qualifying subranges, their keys are concatenated, and a second top-k over
 * the concatenation gives the k-th largest key, the top-k multiset and the
 * top-k indices.
 *
 * Every entry point below replaces one host-side template of the reference,
 * cited next to it.  The reference passes up to 37 positional parameters and
 * caller-allocated buffers; here the same parameters (vec, |V|, k, alpha,
 * beta, digit width, bucket count, filter/beta/shuffle options) travel in
 * drtopk_params, all scratch lives in one caller-owned workspace whose size
 * drtopk_workspace_size() returns, and no call allocates device memory.
 *
 * Conventions
 *  - Keys: DRTOPK_U32 (uint32) or DRTOPK_F32 (IEEE float32).  Float keys are
 *    ordered by their sign-flipped bit pattern: b ^ (b>>31 ? 0xFFFFFFFF :
 *    0x80000000).  So -0.0 < +0.0, and NaNs order by their bits (positive NaN
 *    above +inf, negative NaN below -inf).
 *  - k-th largest = the key at ascending rank n-k (0-based), exactly the
 *    reference's check vec1[n-k]==TopKElement.
 *  - Top-k set: every key strictly greater than the k-th largest, plus the
 *    lowest-index keys equal to it (lowest-index-wins tie rule), written in
 *    ascending index order: topk_idx[j] < topk_idx[j+1] and
 *    topk_keys[j] == keys[topk_idx[j]].
 *  - All calls are stream-ordered and asynchronous unless a host stats pointer
 *    is passed (then the call synchronises the stream before returning).
 *  - Errors are returned as drtopk_status; drtopk_last_error() gives a
 *    thread-local message.  No call exits the process.
 */
#ifndef DRTOPK_H
#define DRTOPK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRTOPK_VERSION_MAJOR 0
#define DRTOPK_VERSION_MINOR 1

typedef enum drtopk_status {
    DRTOPK_OK = 0,
    DRTOPK_ERR_INVALID_ARG = 1,        /* bad pointer, k==0, k>n, alpha/beta/digit width out of range */
    DRTOPK_ERR_WORKSPACE_TOO_SMALL = 2,
    DRTOPK_ERR_UNSUPPORTED = 3,        /* valid but not implemented (e.g. n >= 2^32) */
    DRTOPK_ERR_HIP = 4,                /* a HIP runtime call failed */
    DRTOPK_ERR_INTERNAL = 5,           /* a device-side consistency check failed */
    DRTOPK_ERR_BUDGET = 6              /* serving: request larger than the memory budget */
} drtopk_status;

typedef enum drtopk_algo {
    /* Radix select family: sample_radix_select (baseline+filter+beta+shuffle/radixselect.cuh:1783)
     * and, with alpha == 0, radix_select (radixselect.cuh:1473). */
    DRTOPK_ALGO_RADIX = 0,
    /* Bucket select family: sample_bucket_select (bucket_select/bucketselect_combined.cuh:1280)
     * and, with alpha == 0, bucket_select (bucketselect_combined.cuh:473). */
    DRTOPK_ALGO_BUCKET = 1,
    /* Bitonic family: bitonic_firstTopk + bitonic (bitonic/bitonic_com.cuh:610, :669), driven as in
     * bitonic/sampleBitonic.cu:283-308; alpha == 0 is the plain bitonic top-k (NormalBitonic=1). */
    DRTOPK_ALGO_BITONIC = 2
} drtopk_algo;

typedef enum drtopk_dtype {
    DRTOPK_U32 = 0,
    DRTOPK_F32 = 1
} drtopk_dtype;

/* Option bits (reference experiment folders: baseline, baseline+filter, baseline+filter+beta,
 * baseline+filter+shuffle, baseline+filter+beta+shuffle). */
#define DRTOPK_FLAG_FILTER   0x1u  /* concatenate only keys >= first-top-k threshold (radixselect.cuh:1575-1639) */
#define DRTOPK_FLAG_BETA     0x2u  /* alpha<=5: top-beta delegates + lonely rule (radixselect.cuh:472-600) */
#define DRTOPK_FLAG_SHUFFLE  0x4u  /* alpha<=5: LDS-staged thread-per-subrange sampling (radixselect.cuh:967-1041) */
#define DRTOPK_FLAG_TIMING   0x100u /* record per-phase device times into drtopk_stats (synchronises) */
#define DRTOPK_FLAGS_DEFAULT (DRTOPK_FLAG_FILTER | DRTOPK_FLAG_BETA | DRTOPK_FLAG_SHUFFLE)

#define DRTOPK_ALPHA_AUTO (-1)
/* Radix family: alpha and beta measured on MI355X for this (floor log2 n, ceil log2 k), from the
 * library's built-in table (drtopksc_amd/autotune.py --single); a cell without an entry, or
 * another family, uses DRTOPK_ALPHA_AUTO.  No reference counterpart: the reference tuned alpha
 * by hand per run (alphaManualTuning/main.cu:84). */
#define DRTOPK_ALPHA_TUNED (-2)

typedef struct drtopk_params {
    int32_t alpha;        /* log2 of the subrange size; DRTOPK_ALPHA_AUTO = reference policy, DRTOPK_ALPHA_TUNED =
                           * measured table; 0 = no delegates */
    int32_t beta;         /* delegates per subrange on the beta path, 1..4 (reference: argv beta, default 2) */
    int32_t digit_bits;   /* radix digit width, one of 1,2,4,8,16 (reference NBits, default 8) */
    int32_t num_buckets;  /* bucket family: buckets per refinement level, 4..4096 (reference NBucket, 256) */
    uint32_t flags;       /* DRTOPK_FLAG_* */
    uint32_t reserved[3]; /* must be zero */
} drtopk_params;

/* Intermediate results, for parity checks and the bench; ordered-key bits throughout. */
typedef struct drtopk_stats {
    int32_t algo;
    int32_t alpha;            /* resolved alpha (0 = no delegates) */
    int32_t beta;             /* resolved delegates per subrange (1 on the default path) */
    int32_t path;             /* 0 = plain (no delegates), 1 = default (one max per subrange), 2 = beta path */
    int32_t digit_bits;
    int32_t num_buckets;
    int32_t first_passes;     /* histogram passes of the first top-k (the fused one included) */
    int32_t second_passes;    /* histogram passes of the second top-k */
    uint64_t num_subranges;
    uint64_t num_delegates;
    uint32_t threshold;       /* first top-k threshold (radix flag / bucket threshold / k-th delegate) */
    uint32_t kth_bits;        /* ordered-key bits of the k-th largest key */
    uint64_t selected_subranges;
    uint64_t lonely_subranges;
    uint64_t lonely_elements; /* beta delegates per lonely subrange, as the reference counts them */
    uint64_t concat_size;
    uint64_t count_greater;   /* keys strictly greater than the k-th largest */
    uint64_t ties_taken;      /* keys equal to the k-th largest in the top-k set */
    float phase_ms[5];        /* sample, first top-k, concatenate, second top-k, extract (FLAG_TIMING) */
    uint32_t reserved[3];
} drtopk_stats;

/* Fill *p with the reference's final configuration: alpha auto, beta 2, 8-bit digits, 256 buckets,
 * FILTER|BETA|SHUFFLE (baseline+filter+beta+shuffle/main.cu:83-89, :140-141; Makefile test line). */
void drtopk_params_default(drtopk_params* p);

/* Bytes of device workspace drtopk_select needs for this problem (an upper bound that does not depend
 * on the key values). */
drtopk_status drtopk_workspace_size(drtopk_algo algo, drtopk_dtype dtype, uint64_t n, uint64_t k,
                                    const drtopk_params* p, size_t* bytes);

/* Top-k of d_keys[0..n).  Replaces sample_radix_select / radix_select, sample_bucket_select /
 * bucket_select and bitonic_firstTopk + bitonic (see drtopk_algo).  Keys are read only.
 *   d_kth        device pointer, one key (uint32 or float per dtype), or NULL
 *   d_topk_keys  device pointer, k keys, or NULL
 *   d_topk_idx   device pointer, k int64 indices, or NULL
 *   h_stats      host pointer or NULL; when set the call synchronises the stream and fills it
 *   stream       a hipStream_t, or NULL for the null stream */
drtopk_status drtopk_select(drtopk_algo algo, drtopk_dtype dtype, const void* d_keys, uint64_t n,
                            uint64_t k, const drtopk_params* p, void* d_workspace,
                            size_t workspace_bytes, void* d_kth, void* d_topk_keys,
                            int64_t* d_topk_idx, drtopk_stats* h_stats, void* stream);

/* drtopk_select without a workspace argument, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Scratch is a file-scope device array of
 * DRTOPK_STATIC_SCRATCH_BYTES held by every GPU the library runs on until the process ends; its
 * address is looked up once per device on the first call there.  The call allocates nothing
 * (no hipMalloc / hipFree / event creation); the caller still passes keys and outputs.
 *   - A problem whose drtopk_workspace_size exceeds DRTOPK_STATIC_SCRATCH_BYTES is refused with
 *     DRTOPK_ERR_WORKSPACE_TOO_SMALL (at 64 MiB: every k for n <= 2^22, and k <= 4096 up to
 *     n = 2^31-1 on the radix and bucket families).
 *   - DRTOPK_FLAG_TIMING is refused (DRTOPK_ERR_INVALID_ARG).
 *   - Every call on one GPU shares that GPU's scratch, which keeps the last call's contents:
 *     calls that overlap on one GPU (other streams, other host threads) race and must be
 *     serialised by the caller.  Device-side checks are reported through h_stats only.
 * Other arguments and results are exactly drtopk_select's. */
#define DRTOPK_STATIC_SCRATCH_BYTES (64ull << 20)
drtopk_status drtopk_select_static(drtopk_algo algo, drtopk_dtype dtype, const void* d_keys, uint64_t n,
                                   uint64_t k, const drtopk_params* p, void* d_kth, void* d_topk_keys,
                                   int64_t* d_topk_idx, drtopk_stats* h_stats, void* stream);

/* Delegate pass alone.  Replaces sampleMax (radixselect.cuh:883-918; bitonic/sampleBitonic.cu:150-185)
 * for beta == 1 and sampleMax_NoReduction_Beta2/3/4 (radixselect.cuh:967-1202) for beta 2..4.
 * Subrange s = keys[s*2^alpha, min((s+1)*2^alpha, n)); S = ceil(n / 2^alpha).
 *   d_vals  [beta*S] delegates of subrange s at [beta*s, beta*s+beta), largest first (a multiset:
 *           an equal maximum appears twice); a subrange with fewer than beta keys pads with
 *           ordered key 0 (uint32 0; float bit pattern 0xFFFFFFFF) and position 0xFFFFFFFF.
 *   d_pos   [beta*S] absolute index of each delegate (ties: lowest index first), or NULL
 *   flags   DRTOPK_FLAG_SHUFFLE selects the LDS-staged variant; results are identical. */
drtopk_status drtopk_sample_delegates(drtopk_dtype dtype, const void* d_keys, uint64_t n,
                                      int32_t alpha, int32_t beta, uint32_t flags, void* d_vals,
                                      uint32_t* d_pos, void* stream);

/* Device-side checks.  Kernels verify invariants that valid input always satisfies (the
 * concatenation holds at least k keys; no buffer bound is exceeded).  A failed check sets bits
 * in the workspace; drtopk_select reports them itself only when h_stats is passed.
 * drtopk_workspace_check synchronises the stream, reads the bits of every call made on the
 * workspace (drtopk_select's or drtopk_select_batched's) since the previous check, clears them,
 * stores them in *h_error (may be NULL) and returns DRTOPK_ERR_INTERNAL when any is set,
 * DRTOPK_ERR_INVALID_ARG when no call has used the workspace yet.  No reference counterpart:
 * the reference exits the process on a failed CUDA call (radixselect.cuh:20-28) and checks
 * nothing on the device. */
#define DRTOPK_DEVICE_ERR_CONCAT_SHORT 0x2u /* the concatenation held fewer than k keys */
#define DRTOPK_DEVICE_ERR_CAPACITY     0x4u /* a buffer bound was exceeded */
drtopk_status drtopk_workspace_check(const void* d_workspace, void* stream, uint32_t* h_error);

/* Testing only: every later drtopk_select call in the process (any thread) lets a segment of its
 * concatenation hold at most cap entries (0 restores the full capacity).  A segment that would
 * exceed the bound raises DRTOPK_DEVICE_ERR_CAPACITY instead of being written, which exercises
 * the device-side checks and their reporting (the result of such a call is not valid). */
drtopk_status drtopk_testing_set_segment_cap(uint64_t cap);

/* ------------------------------------------------------------------------------------------
 * Batched top-k: many independent rows of uneven length, each with its own k and key type, in
 * two kernel launches.  The reference has no batched entry point; the nearest caller is its
 * MultiGPU driver, which runs sample -> concat -> second top-k -> findTopk/FillTopk once per
 * vector (MultiGPU/main.cu:366-424).  Per row the result is exactly drtopk_select's: k-th
 * largest, top-k keys and indices (row-local, ascending, lowest index wins ties).
 *
 * Use: drtopk_batch_plan() on the host (validates the rows, picks alpha / digit width / tile per
 * row from a tuning table, lays out the workspace) -> copy the plan blob to device memory once ->
 * drtopk_select_batched() as often as wanted (stream-ordered, no host synchronisation, no
 * allocation, no memset: it can be captured in a HIP graph).
 * ------------------------------------------------------------------------------------------ */
typedef struct drtopk_row {
    uint64_t offset;      /* first key of the row in d_keys, in elements; a multiple of 4 */
    uint64_t n;           /* keys in the row, 1 <= n <= 2^31-1 */
    uint64_t k;           /* top rank, 1 <= k <= n */
    uint64_t out_offset;  /* first of the row's k slots in d_topk_keys / d_topk_idx */
    int32_t dtype;        /* drtopk_dtype of this row's keys */
    int32_t reserved;     /* must be zero */
} drtopk_row;

#define DRTOPK_TUNE_DEFAULT 0xFFu
/* Per-(key type, ceil log2 n, ceil log2 k) bucket choices, as written by the project's autotuner
 * (drtopksc_amd/autotune.py); the first index is the row's drtopk_dtype (0 uint32, 1 float32),
 * whose keys are distributed differently (float32: sign and exponent in the top digits).
 * DRTOPK_TUNE_DEFAULT in a cell keeps the library's built-in choice.  alpha 0 = no delegates (the
 * row is selected directly); the planner lowers an alpha that leaves fewer than k subranges or
 * exceeds the tile.  digit_bits 4 or 8; tile_log2 10..15 is log2 of the keys one workgroup of
 * the delegate pass streams per work unit. */
typedef struct drtopk_tuning {
    uint8_t alpha[2][32][32];
    uint8_t digit_bits[2][32][32];
    uint8_t tile_log2[2][32][32];
} drtopk_tuning;

/* Plan a batch.  Host only.  Call with h_plan == NULL to get *plan_bytes; then again with a host
 * buffer of that size.  keys_len / out_len are the element counts of d_keys / the output arrays
 * (every row must fit).  tuning may be NULL (built-in table).  *ws_bytes receives the workspace
 * size drtopk_select_batched needs. */
drtopk_status drtopk_batch_plan(const drtopk_row* rows, uint64_t nrows, uint64_t keys_len,
                                uint64_t out_len, const drtopk_tuning* tuning, void* h_plan,
                                size_t* plan_bytes, size_t* ws_bytes);

/* Run a planned batch.  h_plan is the host plan, d_plan a device copy of the same bytes.
 *   d_keys       device keys of all rows (16-byte aligned); each row's keys are read as its dtype
 *   d_kth        device array of nrows 32-bit keys (row r's k-th largest, in its dtype), or NULL
 *   d_topk_keys  device array of out_len 32-bit keys, or NULL
 *   d_topk_idx   device array of out_len int64 row-local indices, or NULL
 * Calls that share a workspace must not overlap. */
drtopk_status drtopk_select_batched(const void* h_plan, const void* d_plan, const void* d_keys,
                                    void* d_workspace, size_t workspace_bytes, void* d_kth,
                                    void* d_topk_keys, int64_t* d_topk_idx, void* stream);

/* Built-in tuning table (what tuning == NULL uses). */
void drtopk_tuning_default(drtopk_tuning* t);

/* ------------------------------------------------------------------------------------------
 * Serving: a request loop over a device-resident key store with continuous batching.  The
 * reference has no serving layer (its drivers run one vector per process, main.cu:56-323).
 *
 * A server owns a worker thread, max_inflight HIP streams and one device arena of
 * budget_bytes that holds every launch's scratch and the requests' outputs.  Requests name a
 * row of the key store (offset, n, k, dtype).  Whenever a stream slot is free the worker
 * launches everything queued: requests with n >= solo_n or k > solo_k as single-vector calls
 * (drtopk_select with the reference configuration and DRTOPK_ALPHA_TUNED), the rest as one
 * batched launch (drtopk_select_batched).  A batch takes small requests queued behind large ones
 * only while the arena's largest free block still holds the outputs and scratch of every large
 * request it passes, so a client that waits and releases in submission order never stalls one.
 * Requests that arrive
 * while launches are running join the next launch as soon as a slot and arena space free up, so
 * launches of different streams overlap on the device.  Admission: a request waits in the queue
 * until its outputs and scratch fit the arena; one that could never fit is refused with
 * DRTOPK_ERR_BUDGET.  Results are exactly drtopk_select's.
 * ------------------------------------------------------------------------------------------ */
typedef struct drtopk_server drtopk_server;

typedef struct drtopk_server_config {
    const void* d_store;      /* device key store (16-byte aligned); requests index it */
    uint64_t store_len;       /* its length in 32-bit elements */
    uint64_t budget_bytes;    /* device arena for scratch and outputs (0: 64 GiB, capped at 90% of free memory) */
    uint32_t max_inflight;    /* launches on the device at once, one stream and worker thread each (0: 8) */
    uint32_t max_batch_rows;  /* rows per batched launch (0: 4096) */
    uint64_t solo_n;          /* n at or above which a request runs as a single-vector call (0: 2^25) */
    uint64_t solo_k;          /* k above which a request runs as a single-vector call (0: 4096) */
    const drtopk_tuning* tuning; /* batched launches' tuning table (NULL: built-in) */
    uint32_t reserved[4];     /* must be zero */
} drtopk_server_config;

typedef struct drtopk_query {
    uint64_t offset;          /* first key in the store (elements, a multiple of 4) */
    uint64_t n, k;
    int32_t dtype;            /* drtopk_dtype */
    int32_t reserved;         /* must be zero */
} drtopk_query;

typedef struct drtopk_result {
    int32_t status;              /* drtopk_status of the request */
    uint32_t kth_bits;           /* k-th largest key, 32-bit pattern of the request's dtype */
    const void* d_topk_keys;     /* k keys (ascending index, lowest index wins ties), device, until release */
    const int64_t* d_topk_idx;   /* their row-local indices, device, until release */
    uint64_t t_submit_ns;        /* steady-clock times: submitted, launched, seen complete */
    uint64_t t_launch_ns;
    uint64_t t_done_ns;
    uint32_t launch_rows;        /* requests in the launch that served it (single-vector call: 1) */
    uint32_t solo;               /* 1: served by a single-vector call */
} drtopk_result;

void drtopk_server_config_default(drtopk_server_config* cfg);
drtopk_status drtopk_server_create(const drtopk_server_config* cfg, drtopk_server** srv);
/* Queue a request (thread-safe, does not block); *ticket identifies it. */
drtopk_status drtopk_server_submit(drtopk_server* srv, const drtopk_query* q, uint64_t* ticket);
/* Wait until the request is done (timeout_ns < 0: no limit); fills *res.  DRTOPK_ERR_BUDGET is
 * returned when the wait timed out. */
drtopk_status drtopk_server_wait(drtopk_server* srv, uint64_t ticket, int64_t timeout_ns, drtopk_result* res);
/* Free the request's outputs (after wait). */
drtopk_status drtopk_server_release(drtopk_server* srv, uint64_t ticket);
/* Drain the queue, stop the worker, free the arena and streams. */
drtopk_status drtopk_server_destroy(drtopk_server* srv);

/* Load generator: submit q[i] at t0 + arrival_ns[i] (busy-wait pacing on the calling thread),
 * then wait for and release every request; results[i] keeps the timestamps and kth.  *t0_ns
 * receives t0. */
drtopk_status drtopk_server_replay(drtopk_server* srv, const drtopk_query* q, const uint64_t* arrival_ns,
                                   uint64_t count, drtopk_result* results, uint64_t* t0_ns);
/* The one-at-a-time baseline: the same trace served first-come first-served by single-vector
 * drtopk_select calls (the server's parameters for them) on one stream, each synchronised before
 * the next starts. */
drtopk_status drtopk_replay_serial(const void* d_store, uint64_t store_len, const drtopk_query* q,
                                   const uint64_t* arrival_ns, uint64_t count, drtopk_result* results,
                                   uint64_t* t0_ns);

/* ------------------------------------------------------------------------------------------
 * Raw binary key files: the reference's real-dataset format, a bare sequence of 32-bit
 * little-endian integers with no header, each taken as one key of the same 32 bits
 * (baseline+filter+beta+shuffle/RealDatasets/BinaryDataset/main.cu:109-115 reads them with
 * bin_in.read((char*)&value, sizeof(int)) into its unsigned key array).
 * ------------------------------------------------------------------------------------------ */
/* Number of whole 32-bit values in the file. */
drtopk_status drtopk_file_count(const char* path, uint64_t* count);
/* Values [first, first + count) of the file into dst: host memory, or device memory when
 * dst_on_device is nonzero (staged through pinned host buffers on `stream`, which is
 * synchronised before returning).  A file holding fewer values is DRTOPK_ERR_INVALID_ARG (the
 * reference would keep storing the last value it read). */
drtopk_status drtopk_file_read(const char* path, uint64_t first, uint64_t count, void* dst, int dst_on_device,
                               void* stream);

/* Thread-local message for the last error on this thread ("" if none). */
const char* drtopk_last_error(void);

/* "major.minor gfx950" */
const char* drtopk_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DRTOPK_H */
