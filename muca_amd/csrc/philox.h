// philox.h -- Philox4x32-R counter-based generator (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), as keyed by
// the reference through Random123 1.09 (Random123/include/Random123/philox.h:285-331;
// muca_delta_multigpu.cu:21 selects R = 7).  On gfx950 each round's two
// 32x32->64 products map to v_mad_u64_u32.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

struct philox4 {
    uint32_t v[4];
};

#define MUCA_PHILOX_M0 0xD2511F53u
#define MUCA_PHILOX_M1 0xCD9E8D57u
#define MUCA_PHILOX_W0 0x9E3779B9u
#define MUCA_PHILOX_W1 0xBB67AE85u

template <int R>
__host__ __device__ static inline philox4 philox4x32(uint32_t c0, uint32_t c1, uint32_t c2,
                                                     uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (r > 0) {
            k0 += MUCA_PHILOX_W0;
            k1 += MUCA_PHILOX_W1;
        }
        const uint64_t p0 = (uint64_t)MUCA_PHILOX_M0 * c0;
        const uint64_t p1 = (uint64_t)MUCA_PHILOX_M1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    philox4 o;
    o.v[0] = c0;
    o.v[1] = c1;
    o.v[2] = c2;
    o.v[3] = c3;
    return o;
}

__host__ __device__ static inline philox4 philox4x32_rounds(int rounds, uint32_t c0, uint32_t c1,
                                                            uint32_t c2, uint32_t c3,
                                                            uint32_t k0, uint32_t k1)
{
    return rounds == 10 ? philox4x32<10>(c0, c1, c2, c3, k0, k1)
                        : philox4x32<7>(c0, c1, c2, c3, k0, k1);
}
