// probe_valu.hip -- microbenchmarks of the update kernels' VALU ceiling,
// built into libmuca_probe.so (not part of the libmuca ABI; bench.py loads it).
//
// The integer work every flip attempt needs is half a Philox4x32-7 block
// (muca_delta_multigpu.cu:794-797: two blocks, one per key, per four
// attempts), i.e. 7 rounds x 2 32x32->64 products (v_mad_u64_u32 on gfx950)
// plus xors and key bumps.  k_philox_rate runs nothing but that: every lane
// draws the two keyed streams of a walker for `iters` counters and folds the
// words into one value, so its rate in blocks/ns, times two, bounds the
// attempts/ns of any update kernel that draws the reference's random numbers.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.h"

namespace {

template <int R>
__global__ __launch_bounds__(256) void k_philox_rate(uint32_t iters, uint32_t seed, uint32_t *out)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0;
    for (uint32_t c = 1; c <= iters; c++) {
        const philox4 a = philox4x32<R>(c, seed, 0, 0xBADC0DEDu, 2u * w, 0xdecafbadu);
        const philox4 b = philox4x32<R>(c, seed, 0, 0xBADC0DEDu, 2u * w + 50000u, 0x11cafeadu);
        acc ^= (a.v[0] ^ a.v[1]) + (a.v[2] ^ a.v[3]) + (b.v[0] ^ b.v[1]) + (b.v[2] ^ b.v[3]);
    }
    out[w] = acc;
}

}  // namespace

extern "C" {

// Philox4x32-R blocks per ns over `reps` launches of grid x 256 lanes, each
// lane computing 2 * iters blocks; ms receives the mean launch duration.
// Returns 0, or a HIP error code.
int muca_probe_philox_rate(int device, int rounds, uint32_t grid, uint32_t iters, int reps,
                           double *blocks_per_ns, float *ms)
{
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    const uint32_t lanes = grid * 256u;
    uint32_t *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float total = 0;
    void (*k)(uint32_t, uint32_t, uint32_t *) = rounds == 10 ? k_philox_rate<10> : k_philox_rate<7>;
    if ((e = hipMalloc((void **)&out, (size_t)lanes * 4)) != hipSuccess) return (int)e;
    if ((e = hipEventCreate(&e0)) != hipSuccess) goto done;
    if ((e = hipEventCreate(&e1)) != hipSuccess) goto done;
    hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, 0, iters, 1234u, out);   // warm-up
    if ((e = hipDeviceSynchronize()) != hipSuccess) goto done;
    for (int r = 0; r < reps; r++) {
        hipEventRecord(e0, 0);
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, 0, iters, 1234u + r, out);
        hipEventRecord(e1, 0);
        if ((e = hipEventSynchronize(e1)) != hipSuccess) goto done;
        float t = 0;
        hipEventElapsedTime(&t, e0, e1);
        total += t;
    }
    if (ms) *ms = total / reps;
    if (blocks_per_ns) *blocks_per_ns = 2.0 * lanes * (double)iters / (total / reps * 1e6);
done:
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(out);
    return (int)e;
}

}  // extern "C"
