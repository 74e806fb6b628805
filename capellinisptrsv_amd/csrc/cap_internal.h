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

// returns the CAP_* code of expr when it is not CAP_OK
#define CAP_TRY(expr)                          \
    do {                                       \
        int cap_rc_ = (expr);                  \
        if (cap_rc_ != CAP_OK) return cap_rc_; \
    } while (0)
