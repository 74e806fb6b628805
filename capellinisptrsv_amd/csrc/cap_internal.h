// Internal declarations shared by the library's translation units.
#pragma once
#include <cstdio>

#define CAP_HIP_TRY(expr)                                                                         \
    do {                                                                                          \
        hipError_t cap_e_ = (expr);                                                               \
        if (cap_e_ != hipSuccess) {                                                               \
            std::fprintf(stderr, "capellini: %s failed: %s (%s:%d)\n", #expr,                     \
                         hipGetErrorString(cap_e_), __FILE__, __LINE__);                          \
            return CAP_ERR_HIP;                                                                   \
        }                                                                                         \
    } while (0)
