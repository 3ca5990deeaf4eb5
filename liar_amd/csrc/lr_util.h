// lr_util.h -- error reporting shared by the library's translation units.
#pragma once

#include <string>

#include "../../include/liar_amd.h"

namespace lr {

// Sets the calling thread's last-error message and returns `code`.
int fail(int code, const std::string& msg);

}  // namespace lr

#define LR_HIP_CHECK(expr)                                                                     \
    do {                                                                                       \
        hipError_t lr_e_ = (expr);                                                             \
        if (lr_e_ != hipSuccess)                                                               \
            return ::lr::fail(LR_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(lr_e_)); \
    } while (0)
