// hip_check.hpp -- error propagation and scoped device buffers shared by the
// library's sources (host code).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

// Evaluates a HIP call; on failure stores "<call> failed: <reason>" in the
// enclosing function's std::string *err and returns -1 from it.
#define HIP_TRY(expr)                                                         \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            if (err) *err = std::string(#expr " failed: ") + hipGetErrorString(e_); \
            return -1;                                                        \
        }                                                                     \
    } while (0)

namespace calypso {

// A device allocation released when it goes out of scope (unless released
// to an owner first).
struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void *release() { void *q = p; p = nullptr; return q; }
};

}  // namespace calypso
