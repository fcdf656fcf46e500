// hpp_mirror.cpp -- GPU check of the C++ mirror (include/drtopk.hpp): each reference-named entry
// point, called as the reference's drivers call theirs, must give exactly drtopk_select's result
// (k-th key, top-k keys and indices) on the same keys, and none may allocate: with an empty
// Workspace every one returns DRTOPK_ERR_WORKSPACE_TOO_SMALL.  Built by drtopksc_amd/build.py,
// run by tests/test_gpu_hpp.py.  Exit status 0 when every check passes.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "drtopk.hpp"

static int g_fail = 0, g_checks = 0;

// A failed HIP runtime call ends the program (status 2): the checks below need their buffers.
#define HC(x)                                                                       \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::printf("HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            std::exit(2);                                                           \
        }                                                                           \
    } while (0)

#define CHECK(cond, ...)                              \
    do {                                              \
        ++g_checks;                                   \
        if (!(cond)) {                                \
            ++g_fail;                                 \
            std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                 \
            std::printf("\n");                        \
        }                                             \
    } while (0)

template <typename T>
struct Direct {
    T kth;
    std::vector<T> keys;
    std::vector<int64_t> idx;
};

// drtopk_select through the C ABI with its own workspace and outputs.
template <typename T>
static Direct<T> direct(drtopk_algo algo, const T* d_keys, uint64_t n, uint64_t k, const drtopk_params& p) {
    Direct<T> r;
    size_t b = 0;
    drtopk_workspace_size(algo, drtopk::dtype_of<T>::value, n, k, &p, &b);
    void *ws, *kth, *keys, *idx;
    HC(hipMalloc(&ws, b));
    HC(hipMalloc(&kth, 256));
    HC(hipMalloc(&keys, k * sizeof(T)));
    HC(hipMalloc(&idx, k * 8));
    drtopk_status s = drtopk_select(algo, drtopk::dtype_of<T>::value, d_keys, n, k, &p, ws, b, kth, keys,
                                    (int64_t*)idx, nullptr, nullptr);
    CHECK(s == DRTOPK_OK, "direct select: %s", drtopk_last_error());
    r.keys.resize(k);
    r.idx.resize(k);
    HC(hipMemcpy(&r.kth, kth, sizeof(T), hipMemcpyDeviceToHost));
    HC(hipMemcpy(r.keys.data(), keys, k * sizeof(T), hipMemcpyDeviceToHost));
    HC(hipMemcpy(r.idx.data(), idx, k * 8, hipMemcpyDeviceToHost));
    HC(hipFree(ws)); HC(hipFree(kth)); HC(hipFree(keys)); HC(hipFree(idx));
    return r;
}

template <typename T>
static void compare(const char* name, const Direct<T>& want, T kth, const T* d_keys, const int64_t* d_idx, uint64_t k) {
    CHECK(std::memcmp(&kth, &want.kth, sizeof(T)) == 0, "%s: k-th key differs", name);
    if (d_keys) {
        std::vector<T> g(k);
        HC(hipMemcpy(g.data(), d_keys, k * sizeof(T), hipMemcpyDeviceToHost));
        CHECK(std::memcmp(g.data(), want.keys.data(), k * sizeof(T)) == 0, "%s: top-k keys differ", name);
    }
    if (d_idx) {
        std::vector<int64_t> g(k);
        HC(hipMemcpy(g.data(), d_idx, k * 8, hipMemcpyDeviceToHost));
        CHECK(std::memcmp(g.data(), want.idx.data(), k * 8) == 0, "%s: top-k indices differ", name);
    }
}

int main() {
    const int n = (1 << 20) + 3;
    std::mt19937 g(17);
    std::vector<unsigned> hu(n);
    std::vector<float> hf(n);
    for (int i = 0; i < n; ++i) hu[i] = g();
    for (int i = 0; i < n; ++i) hf[i] = (float)((double)(g() >> 8) * (1.0 / 8388608.0) - 1.0);
    unsigned* du;
    float* df;
    HC(hipMalloc(&du, n * 4ull));
    HC(hipMalloc(&df, n * 4ull));
    HC(hipMemcpy(du, hu.data(), n * 4ull, hipMemcpyHostToDevice));
    HC(hipMemcpy(df, hf.data(), n * 4ull, hipMemcpyHostToDevice));
    const int kmax = 1 << 17;
    void *okeys, *oidx;
    HC(hipMalloc(&okeys, kmax * 4ull));
    HC(hipMalloc(&oidx, kmax * 8ull));

    // nothing reserved: every mirror refuses instead of allocating
    {
        drtopk::Workspace empty;
        unsigned K = 0;
        double t0, t1, t2, t3;
        CHECK((drtopk::radix_select<unsigned, int>(du, n, 100, 256, K, 8, 3, empty)) == DRTOPK_ERR_WORKSPACE_TOO_SMALL,
              "radix_select without a workspace");
        CHECK((drtopk::sample_radix_select<unsigned, int>(du, n, 100, 256, K, 8, 3, (n + 1023) / 1024, 1024, 10, t0, t1,
                                                          t2, t3, 2, true, empty)) == DRTOPK_ERR_WORKSPACE_TOO_SMALL,
              "sample_radix_select without a workspace");
        CHECK((drtopk::bucket_select<unsigned, int>(du, n, n - 100, 256, K, empty)) == DRTOPK_ERR_WORKSPACE_TOO_SMALL,
              "bucket_select without a workspace");
        CHECK(empty.data() == nullptr && empty.size() == 0, "a mirror call allocated");
    }

    for (int k : {1, 128, 4096, 1 << 17}) {
        drtopk::Workspace ws;
        CHECK(ws.reserve_for<unsigned>(DRTOPK_ALGO_RADIX, n, k) == DRTOPK_OK, "reserve_for radix");
        CHECK(ws.reserve_for<unsigned>(DRTOPK_ALGO_BUCKET, n, k) == DRTOPK_OK, "reserve_for bucket");
        CHECK(ws.reserve_for<float>(DRTOPK_ALGO_RADIX, n, k) == DRTOPK_OK, "reserve_for f32");
        const size_t reserved = ws.size();
        drtopk_params p;
        char name[96];

        // sample_radix_select: the reference's driver (main.cu:83-89, :256-263)
        {
            const int L2 = 20;  // floor(log2 n)
            int alpha = (L2 + 3 - (int)std::ceil(std::log2((double)k))) / 2;
            const bool def = alpha > 5;
            const int S = (n + (1 << alpha) - 1) >> alpha;
            unsigned K = 0;
            double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
            drtopk_status s = drtopk::sample_radix_select<unsigned, int>(du, n, k, 256, K, 8, 3, S, 1 << alpha, alpha,
                                                                           t0, t1, t2, t3, 2, def, ws,
                                                                           DRTOPK_FLAGS_DEFAULT, (unsigned*)okeys,
                                                                           (int64_t*)oidx);
            std::snprintf(name, sizeof name, "sample_radix_select k=%d alpha=%d", k, alpha);
            CHECK(s == DRTOPK_OK, "%s: %s", name, drtopk_last_error());
            drtopk_params_default(&p);
            p.alpha = alpha;
            p.beta = def ? 1 : 2;
            if (def) p.flags &= ~DRTOPK_FLAG_BETA;
            compare(name, direct<unsigned>(DRTOPK_ALGO_RADIX, du, n, k, p), K, (unsigned*)okeys, (int64_t*)oidx, k);
            CHECK(t0 > 0 && t2 >= 0, "%s: phase times", name);
        }
        // radix_select (alpha 0), float keys
        {
            float K = 0;
            drtopk_status s = drtopk::radix_select<float, int>(df, n, k, 256, K, 8, 3, ws, (float*)okeys, (int64_t*)oidx);
            std::snprintf(name, sizeof name, "radix_select f32 k=%d", k);
            CHECK(s == DRTOPK_OK, "%s: %s", name, drtopk_last_error());
            drtopk_params_default(&p);
            p.alpha = 0;
            compare(name, direct<float>(DRTOPK_ALGO_RADIX, df, n, k, p), K, (float*)okeys, (int64_t*)oidx, k);
        }
        // bucket_select: ascending rank n-k (bucketselect_combined.cuh:1355)
        {
            unsigned K = 0;
            drtopk_status s = drtopk::bucket_select<unsigned, int>(du, n, n - k, 256, K, ws);
            std::snprintf(name, sizeof name, "bucket_select k=%d", k);
            CHECK(s == DRTOPK_OK, "%s: %s", name, drtopk_last_error());
            drtopk_params_default(&p);
            p.alpha = 0;
            compare(name, direct<unsigned>(DRTOPK_ALGO_BUCKET, du, n, k, p), K, (unsigned*)nullptr, nullptr, k);
        }
        // sample_bucket_select: alpha per bucket_select/main.cu:66-67
        {
            int alpha = (20 + 3 - (int)std::ceil(std::log2((double)k))) / 2;
            if (alpha < 5) alpha++;
            const bool def = alpha > 5;
            unsigned K = 0;
            drtopk_status s = drtopk::sample_bucket_select<unsigned, int>(du, n, k, 256, K, alpha, 2, def, ws,
                                                                            (unsigned*)okeys, (int64_t*)oidx);
            std::snprintf(name, sizeof name, "sample_bucket_select k=%d alpha=%d", k, alpha);
            CHECK(s == DRTOPK_OK, "%s: %s", name, drtopk_last_error());
            drtopk_params_default(&p);
            p.alpha = alpha;
            p.num_buckets = 256;
            p.beta = def ? 1 : 2;
            p.flags = DRTOPK_FLAG_FILTER | DRTOPK_FLAG_SHUFFLE | (def ? 0u : DRTOPK_FLAG_BETA);
            compare(name, direct<unsigned>(DRTOPK_ALGO_BUCKET, du, n, k, p), K, (unsigned*)okeys, (int64_t*)oidx, k);
        }
        // bitonic_topk (power-of-two k, auto alpha)
        if (k <= 4096) {
            drtopk::Workspace wb;
            CHECK(wb.reserve_for<unsigned>(DRTOPK_ALGO_BITONIC, n, k) == DRTOPK_OK, "reserve_for bitonic");
            unsigned K = 0;
            drtopk_status s = drtopk::bitonic_topk<unsigned, int>(du, n, k, -1, K, wb, (unsigned*)okeys, (int64_t*)oidx);
            std::snprintf(name, sizeof name, "bitonic_topk k=%d", k);
            CHECK(s == DRTOPK_OK, "%s: %s", name, drtopk_last_error());
            drtopk_params_default(&p);
            compare(name, direct<unsigned>(DRTOPK_ALGO_BITONIC, du, n, k, p), K, (unsigned*)okeys, (int64_t*)oidx, k);
        }
        CHECK(ws.size() == reserved, "k=%d: a mirror call grew the workspace", k);
    }
    HC(hipFree(du)); HC(hipFree(df)); HC(hipFree(okeys)); HC(hipFree(oidx));
    std::printf("hpp mirror: %d checks, %d failed\n", g_checks, g_fail);
    return g_fail ? 1 : 0;
}
