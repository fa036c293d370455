// moments.h -- exact per-D moment sums of the produce path (count, sum |M|,
// sum M^2, sum M^4 per D and walker block), shared by the update kernels.
// Layout of a block's sums: kMomWords arrays of N+1 uint64 words (kernels.h).
#pragma once
#include "kernels.h"

namespace muca {

// A 128-bit value (vl, vh < 2^32) added to three limb words w0 + w1 2^32 +
// w2 2^64 with non-returning atomics (no carry to detect: each word gains
// less than 2^32 per add).
__device__ __forceinline__ void atomic_add_limbs(uint64_t *w0, uint64_t *w1, uint64_t *w2,
                                                 uint64_t vl, uint64_t vh)
{
    const uint64_t l0 = vl & 0xFFFFFFFFull, l1 = vl >> 32;
    if (l0) atomicAdd((unsigned long long *)w0, (unsigned long long)l0);
    if (l1) atomicAdd((unsigned long long *)w1, (unsigned long long)l1);
    if (vh) atomicAdd((unsigned long long *)w2, (unsigned long long)vh);
}

// A walker's production attempts at one D, summed in registers and added to
// its block's per-D moments when D changes: count, sum |M|, and 128-bit sums
// of M^2 and M^4 (|M| <= N <= 2^16, so M^4 <= 2^64 needs a high word too).
struct MomRun {
    uint64_t n, s1, s2l, s2h, s4l, s4h;
    int32_t d;
};

__device__ __forceinline__ void mom_add(MomRun &r, int32_t m)
{
    const uint64_t am = (uint64_t)(m < 0 ? -m : m);
    const uint64_t m2 = am * am;
    const uint64_t m4l = m2 * m2, m4h = __umul64hi(m2, m2);
    r.n += 1;
    r.s1 += am;
    r.s2l += m2;
    r.s2h += r.s2l < m2 ? 1u : 0u;
    r.s4l += m4l;
    r.s4h += m4h + (r.s4l < m4l ? 1u : 0u);
}

__device__ __forceinline__ void mom_flush(const MomRun &r, uint64_t *mb, uint64_t nb)
{
    if (r.n == 0) return;
    const uint64_t d = (uint64_t)r.d;
    atomicAdd((unsigned long long *)&mb[d], (unsigned long long)r.n);
    atomicAdd((unsigned long long *)&mb[nb + d], (unsigned long long)r.s1);
    // a run holds at most NUPDATES < 2^32 attempts, so the high words are < 2^32
    atomic_add_limbs(&mb[2 * nb + d], &mb[3 * nb + d], &mb[4 * nb + d], r.s2l, r.s2h);
    atomic_add_limbs(&mb[5 * nb + d], &mb[6 * nb + d], &mb[7 * nb + d], r.s4l, r.s4h);
}

}  // namespace muca
