#pragma once
// Host-side error type of the library and the HIP call check: every failure
// becomes a zamr::Error carrying a ZAMR_E* code (include/zamr.h) and a message.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace zamr {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

}  // namespace zamr

#define ZAMR_HIP(call)                                                                                  \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::zamr::Error(4, std::string(#call) + ": " + hipGetErrorString(e_));                 \
    } while (0)
