// ctx.h -- the context object behind muca_ctx (include/muca.h) and the error
// helpers shared by the C-ABI translation units.  Internal to libmuca.
#pragma once
#include "../../include/muca.h"

#include <hip/hip_runtime.h>

#include <string>

#include "kernels.h"

// the calling thread's last error message (muca_last_error)
extern thread_local std::string g_muca_err;

inline int fail(int code, const std::string &msg)
{
    g_muca_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(MUCA_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct muca_ctx {
    muca_params p{};
    uint32_t N = 0;
    uint64_t W = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int8_t *lat = nullptr;
    int32_t *E = nullptr, *D = nullptr, *M = nullptr;
    float *lnw32 = nullptr;
    double *lnw64 = nullptr;
    uint32_t *order = nullptr, *naver = nullptr;
    uint64_t *hist = nullptr;
    uint64_t *joint = nullptr;
    uint64_t *moms = nullptr;   // MUCA_OBS_MOMENTS: blocks * kMomWords * (N+1)
    uint64_t joint_stride = 0, block_walkers = 0;
    double *mom = nullptr, *deltas = nullptr, *rw_out = nullptr;
    uint32_t *rw_flag = nullptr;
    uint32_t rw_cap = 0;
    bool have_weights = false;
    uint32_t hist_preload = 0;   // test hook, see muca_set_test_hist_preload
    muca::Launch cfg{};
    uint32_t cus = 0;
};
