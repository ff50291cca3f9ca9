// api_util.hpp -- error reporting shared by the host runtimes behind the
// C ABI: a per-thread message read by gpu_bpf_last_error().
#pragma once

#include <hip/hip_runtime.h>

namespace bpf {
void set_err(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_err();
}  // namespace bpf

#define HIP_TRY(expr)                                                          \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            bpf::set_err("%s failed: %s", #expr, hipGetErrorString(e_));      \
            return -1;                                                         \
        }                                                                      \
    } while (0)
