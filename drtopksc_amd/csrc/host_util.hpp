// host_util.hpp -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <stdint.h>

#include "../../include/drtopk.h"

namespace drtopk {

// Record a thread-local message for drtopk_last_error() and return s.
drtopk_status fail(drtopk_status s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
// Forget the last message (start of every entry point).
void clear_error();

inline int floor_log2(uint64_t x) { int l = -1; while (x) { x >>= 1; ++l; } return l; }
inline int ceil_log2(uint64_t x) { int l = floor_log2(x); return ((1ull << l) == x) ? l : l + 1; }
inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }
inline uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

}  // namespace drtopk

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return ::drtopk::fail(DRTOPK_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)
