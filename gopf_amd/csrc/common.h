// common.h -- error type shared by the host-side plan code.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace gpf {

// A failed call; code is one of the GPF_E* status codes of include/gpf.h.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) throw ::gpf::Error(GPF_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace gpf
