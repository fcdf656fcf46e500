// io.hip -- raw binary key files (drtopk_file_count / drtopk_file_read).
//
// The reference's real-dataset drivers read a bare sequence of 32-bit integers with no header,
// one sizeof(int) at a time, and store each as an unsigned key (the same 32 bits)
// (baseline+filter+beta+shuffle/RealDatasets/BinaryDataset/main.cu:109-115, `#ifdef readBinary`).
// Here the file is read in large pieces, into host memory directly or into device memory
// through a pinned staging buffer whose pieces are copied while the next one is read.  A file
// shorter than the request is an error (the reference keeps storing the last value it read).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <sys/stat.h>

#include "../../include/drtopk.h"
#include "common.hpp"
#include "host_util.hpp"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the file's little-endian words are read as they lie");

using namespace drtopk;

namespace {

constexpr size_t PIECE = 32u << 20;  // bytes per read / staged copy

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

drtopk_status open_at(const char* path, uint64_t first, uint64_t count, File& fh) {
    if (!path) return fail(DRTOPK_ERR_INVALID_ARG, "path is NULL");
    struct stat sb;
    if (stat(path, &sb) != 0) return fail(DRTOPK_ERR_INVALID_ARG, "cannot stat %s: %s", path, std::strerror(errno));
    const uint64_t have = (uint64_t)sb.st_size / 4;
    if (first > have || count > have - first)
        return fail(DRTOPK_ERR_INVALID_ARG, "%s holds %llu 32-bit values, asked for [%llu, %llu)", path,
                    (unsigned long long)have, (unsigned long long)first, (unsigned long long)(first + count));
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return fail(DRTOPK_ERR_INVALID_ARG, "cannot open %s: %s", path, std::strerror(errno));
    if (fseeko(fh.f, (off_t)(first * 4), SEEK_SET) != 0)
        return fail(DRTOPK_ERR_INVALID_ARG, "cannot seek in %s: %s", path, std::strerror(errno));
    return DRTOPK_OK;
}

drtopk_status read_exact(File& fh, void* dst, size_t bytes, const char* path) {
    if (std::fread(dst, 1, bytes, fh.f) != bytes)
        return fail(DRTOPK_ERR_INVALID_ARG, "short read from %s", path);
    return DRTOPK_OK;
}

}  // namespace

extern "C" {

drtopk_status drtopk_file_count(const char* path, uint64_t* count) {
    clear_error();
    if (!path || !count) return fail(DRTOPK_ERR_INVALID_ARG, "path and count must not be NULL");
    struct stat sb;
    if (stat(path, &sb) != 0) return fail(DRTOPK_ERR_INVALID_ARG, "cannot stat %s: %s", path, std::strerror(errno));
    *count = (uint64_t)sb.st_size / 4;
    return DRTOPK_OK;
}

drtopk_status drtopk_file_read(const char* path, uint64_t first, uint64_t count, void* dst, int dst_on_device,
                               void* stream) {
    clear_error();
    if (!dst && count) return fail(DRTOPK_ERR_INVALID_ARG, "dst is NULL");
    File fh;
    drtopk_status s = open_at(path, first, count, fh);
    if (s != DRTOPK_OK || count == 0) return s;
    const uint64_t bytes = count * 4;
    if (!dst_on_device) return read_exact(fh, dst, (size_t)bytes, path);
    // device destination: two pinned pieces, one being read while the other is copied
    hipStream_t st = (hipStream_t)stream;
    unsigned char* stage[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(st);  // no staged copy may still read a piece being freed
        for (int i = 0; i < 2; ++i) {
            if (done[i]) (void)hipEventDestroy(done[i]);
            if (stage[i]) (void)hipHostFree(stage[i]);
        }
    };
    const size_t piece = (size_t)std::min<uint64_t>(PIECE, bytes);
    for (int i = 0; i < 2; ++i) {
        if (hipHostMalloc((void**)&stage[i], piece, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess) {
            cleanup();
            return fail(DRTOPK_ERR_HIP, "cannot allocate the pinned staging buffers");
        }
    }
    unsigned char* d = (unsigned char*)dst;
    bool used[2] = {false, false};
    int b = 0;
    for (uint64_t off = 0; off < bytes; off += piece, b ^= 1) {
        const size_t len = (size_t)std::min<uint64_t>(piece, bytes - off);
        // the copy that last used this piece must have finished before it is overwritten
        if (used[b] && hipEventSynchronize(done[b]) != hipSuccess) {
            cleanup();
            return fail(DRTOPK_ERR_HIP, "staging copy failed");
        }
        if ((s = read_exact(fh, stage[b], len, path)) != DRTOPK_OK) {
            cleanup();
            return s;
        }
        if (hipMemcpyAsync(d + off, stage[b], len, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipEventRecord(done[b], st) != hipSuccess) {
            cleanup();
            return fail(DRTOPK_ERR_HIP, "host-to-device copy failed");
        }
        used[b] = true;
    }
    const hipError_t e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) return fail(DRTOPK_ERR_HIP, "host-to-device copy failed: %s", hipGetErrorString(e));
    return DRTOPK_OK;
}

}  // extern "C"
