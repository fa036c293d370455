// acclog.h -- the acceptance logarithm: correctly rounded float log(u) for
// u = u01fixedpt<float>(r) = (1 | (r >> 8)) * 2^-24.
//
// The reference compares the CUDA device logf of this uniform against the
// weight difference (muca_delta_multigpu.cu:781,816; produce2_multigpu.cu:446,481).
// That library function is closed, so the engine defines the acceptance log
// as the correctly rounded one.  It is evaluated in double with a
// table-driven reduction (Tang, ACM TOMS 16(4) 1990; table in
// acclog_table.h), carried as hi + lo, and rounded once to float.  The same
// code runs on the host and the device; every product that matters is an
// explicit fma, so both give the same bits.  Correct rounding is checked
// exhaustively over all 2^23 inputs against 113-bit logq.
#pragma once
#include <stdint.h>
#include <string.h>
#include <hip/hip_runtime.h>

#include "acclog_table.h"

static __constant__ double acclog_invc_dev[128] = {ACCLOG_INVC_LIST};
static __constant__ double acclog_logc_dev[128] = {ACCLOG_LOGC_LIST};
static const double acclog_invc_host[128] = {ACCLOG_INVC_LIST};
static const double acclog_logc_host[128] = {ACCLOG_LOGC_LIST};

__host__ __device__ static inline double acclog_bits_to_double(uint64_t b)
{
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}

__host__ __device__ static inline uint64_t acclog_double_to_bits(double d)
{
    uint64_t b;
    memcpy(&b, &d, sizeof b);
    return b;
}

// log(u) in double-double precision, u = m * 2^-24 with m odd in [1, 2^24),
// with the table at (invc_tab, logc_tab) (device constant memory, host
// memory, or a copy in LDS).
__host__ __device__ static inline double acclog_u01_double_tab(uint32_t r,
                                                              const double *invc_tab,
                                                              const double *logc_tab)
{
    const uint32_t m = 1u | (r >> 8);
    const double x = (double)m * 0x1p-24;
    const uint64_t ix = acclog_double_to_bits(x);
    const uint64_t tmp = ix - ACCLOG_OFF;
    const int i = (int)((tmp >> 45) & 127u);
    const int64_t k = (int64_t)tmp >> 52;
    const double z = acclog_bits_to_double(ix - (tmp & (0xfffULL << 52)));
    const double invc = invc_tab[i];
    const double logc = logc_tab[i];
    const double rr = fma(z, invc, -1.0);           // |rr| <= 2^-7
    const double kf = (double)k;
    const double kh = kf * ACCLOG_LN2HI;             // exact: |k| <= 24
    const double t1 = kh + logc;
    const double e1 = (kh - t1) + logc;              // |kh| >= |logc| or kh == 0
    const double t2 = t1 + rr;
    const double e2 = (t1 - t2) + rr;                // |t1| >= |rr| or t1 == 0
    // log1p(rr) - rr = rr^2 * P(rr), Taylor series to rr^9
    double p = 1.0 / 9.0;
    p = fma(p, rr, -1.0 / 8.0);
    p = fma(p, rr, 1.0 / 7.0);
    p = fma(p, rr, -1.0 / 6.0);
    p = fma(p, rr, 1.0 / 5.0);
    p = fma(p, rr, -1.0 / 4.0);
    p = fma(p, rr, 1.0 / 3.0);
    p = fma(p, rr, -0.5);
    const double r2 = rr * rr;
    const double lo = fma(r2, p, fma(kf, ACCLOG_LN2LO, e1 + e2));
    return t2 + lo;
}

__host__ __device__ static inline double acclog_u01_double(uint32_t r)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return acclog_u01_double_tab(r, acclog_invc_dev, acclog_logc_dev);
#else
    return acclog_u01_double_tab(r, acclog_invc_host, acclog_logc_host);
#endif
}

__host__ __device__ static inline float acclog_u01(uint32_t r)
{
    return (float)acclog_u01_double(r);
}

// ---------------------------------------------------------------------------
// Fast screening of the acceptance test.  The reference accepts iff
// diff >= 0 or log(u) < diff; since log(u) < 0 for every u01fixedpt value
// that is exactly "log(u) < diff".  la = v_log_f32(u) * ln2 differs from the
// correctly rounded log by at most ACCLOG_FAST_REL * |la| + ACCLOG_FAST_ABS
// for every one of the 2^23 inputs (half of the band below; checked
// exhaustively on the device by the test suite).  So
//   la + band <  diff  => accept,   la - band >= diff  => reject,
// and only a diff inside the band needs the exact log.
#define ACCLOG_FAST_REL 0x1p-20f
#define ACCLOG_FAST_ABS 0x1p-36f

__device__ static inline float acclog_fast_u01(uint32_t r)
{
    const float u = (float)(1u | (r >> 8)) * 0x1p-24f;
    return __builtin_amdgcn_logf(u) * 0x1.62e430p-1f;
}

__device__ static inline float acclog_fast_band(float la)
{
    return fabsf(la) * ACCLOG_FAST_REL + ACCLOG_FAST_ABS;
}

// device: copy the 128-entry table into LDS (invc at dst[0..127], logc at dst[128..255])
__device__ static inline void acclog_table_to_lds(double *dst, int tid, int nthreads)
{
    for (int i = tid; i < 128; i += nthreads) {
        dst[i] = acclog_invc_dev[i];
        dst[128 + i] = acclog_logc_dev[i];
    }
}
