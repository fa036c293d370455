// muca_lane.hip -- the lane update kernel: one lane per walker, for even L.
//
// For large walker counts (enough wavefronts to hide memory latency) the
// update is bound by per-attempt VALU work.  This kernel trims it:
//  - the reference's even-L site order (muca_delta_multigpu.cu:173: even
//    linear indices, then odd ones) is walked with scalar arithmetic instead
//    of the order/neighbour tables;
//  - ln W and the D histogram live in LDS;
//  - the five spins of attempt i+1 are loaded while attempt i is decided.
//    For even L >= 4 the site of attempt i is never among attempt i+1's
//    site and neighbours, so the prefetched values equal the sequential ones;
//  - the acceptance "diff >= 0 or log(u) < diff" (equivalently log(u) <
//    diff, since log(u) < 0) is screened with the hardware log: outside a
//    band proven exhaustively around it, the decision needs no exact log;
//    inside, the correctly rounded acclog decides.
// Results are bit-identical to the generic kernel and the reference chain.
#include "kernels.h"

#include "acclog.h"
#include "philox.h"

namespace muca {

// The walkers wbase .. wbase + blockDim.x - 1 of the launch described by a
// (every lane of the workgroup calls it together).
template <int R, bool FP64, int KEYING, bool LDSHIST, bool JOINT>
__device__ __forceinline__ void lane_run(const McsArgs &a, uint32_t lds_w, uint64_t wbase,
                                         unsigned char *lds)
{
    typedef typename std::conditional<FP64, double, float>::type T;
    __syncthreads();   // LDS may still be read by a previous call's flush
    T *wt = (T *)lds;
    uint32_t *hist = (uint32_t *)(lds + lds_w);
    const uint32_t N = a.N, L = a.L, Rw = L / 2;
    for (uint32_t i = threadIdx.x; i <= N; i += blockDim.x) {
        wt[i] = FP64 ? (T)a.lnw64[i] : (T)a.lnw32[i];
        if (LDSHIST) hist[i] = a.hist_preload;
    }
    __syncthreads();
    const uint64_t w = wbase + threadIdx.x;
    if (w < a.W) {
        const uint32_t wid = a.w_offset + (uint32_t)w;
        const uint32_t k1a = KEYING == 0 ? 2u * wid : wid;
        const uint32_t k1b = 0xdecafbadu;
        const uint32_t k3a = KEYING == 0 ? 2u * wid + 50000u : wid + 40000u;
        const uint32_t k3b = KEYING == 0 ? 0x11cafeadu : 0xdecafeadu;
        const int8_t *sr = a.lat + w;
        int8_t *sw = a.lat + w;
        const uint64_t W = a.W;
        int32_t e = a.E[w], d = a.D[w], m = a.M[w];
        uint64_t *jbase = nullptr;
        if (JOINT) jbase = a.joint + (w / a.block_walkers) * a.joint_stride;
        // spins of the attempt at stripe position (h, y, xi): x = 2 xi + h
        auto load5 = [&](uint32_t h, uint32_t y, uint32_t xi, int32_t &old_s, int32_t &nb) {
            const uint32_t x = 2 * xi + h;
            const uint32_t xr = (x + 1 == L) ? 0 : x + 1, xl = (x == 0) ? L - 1 : x - 1;
            const uint32_t yu = (y + 1 == L) ? 0 : y + 1, yd = (y == 0) ? L - 1 : y - 1;
            old_s = sr[(uint64_t)(y * L + x) * W];
            nb = sr[(uint64_t)(y * L + xr) * W] + sr[(uint64_t)(y * L + xl) * W] +
                 sr[(uint64_t)(yu * L + x) * W] + sr[(uint64_t)(yd * L + x) * W];
        };
        uint32_t c0 = 0;
        // a workgroup adds at most blockDim NUPDATES counts to one u32 bin
        const bool may_wrap = a.hist_preload != 0 || (uint64_t)blockDim.x * a.nup >= (1ull << 32);
        for (int phase = 0; phase < 2; phase++) {
            const uint64_t n = phase ? a.nup : a.therm;
            if (n == 0) continue;
            uint32_t h = 0, y = 0, xi = 0;          // position of the current attempt
            int32_t old_s, nb;
            load5(h, y, xi, old_s, nb);
            for (uint64_t i0 = 0; i0 < n; i0 += 4) {
                ++c0;
                const philox4 r1 = philox4x32<R>(c0, a.seed, a.iter, 0xBADC0DEDu, k1a, k1b);
                const philox4 r3 = philox4x32<R>(c0, a.seed, a.iter, 0xBADC0DEDu, k3a, k3b);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (i0 + j < n) {
                        // next position and its prefetch
                        uint32_t nh = h, ny = y, nxi = xi + 1;
                        if (nxi == Rw) {
                            nxi = 0;
                            if (++ny == L) {
                                ny = 0;
                                nh ^= 1u;
                            }
                        }
                        int32_t n_old = 0, n_nb = 0;
                        if (i0 + j + 1 < n) load5(nh, ny, nxi, n_old, n_nb);
                        // the attempt at (h, y, xi)
                        int32_t new_s = old_s + (r3.v[j] < 0x80000000u ? 1 : -1);
                        new_s = new_s == 2 ? -1 : (new_s == -2 ? 1 : new_s);
                        const int32_t ds = new_s - old_s;
                        const int32_t dE = -ds * nb;
                        const int32_t dDel = new_s * new_s - old_s * old_s;
                        T diff = (T)(-(double)dE * a.beta);
                        diff = diff + wt[d + dDel];
                        diff = diff - wt[d];
                        const float la = acclog_fast_u01(r1.v[j]);
                        const float band = acclog_fast_band(la);
                        bool acc = (T)(la + band) < diff;
                        const bool amb = !acc && !((T)(la - band) >= diff);
                        if (__builtin_amdgcn_ballot_w64(amb)) {
                            if (amb) acc = (T)acclog_u01(r1.v[j]) < diff;
                        }
                        if (acc) {
                            sw[(uint64_t)(y * L + 2 * xi + h) * W] = (int8_t)new_s;
                            e += dE;
                            d += dDel;
                            m += ds;
                        }
                        if (phase) {
                            if (LDSHIST) {
                                if (may_wrap) {
                                    // a u32 bin that wraps carries 2^32 into HBM
                                    if (atomicAdd(&hist[d], 1u) == 0xFFFFFFFFu)
                                        atomicAdd((unsigned long long *)&a.hist[d], 1ull << 32);
                                } else {
                                    atomicAdd(&hist[d], 1u);
                                }
                            }
                            else
                                atomicAdd((unsigned long long *)&a.hist[d], 1ull);
                            if (JOINT) {
                                const int64_t mm = min(max((int64_t)m, -(int64_t)d), (int64_t)d);
                                atomicAdd((unsigned long long *)&jbase[(mm + (int64_t)d * d +
                                                                        2 * (int64_t)d) / 2],
                                          1ull);
                            }
                        }
                        h = nh;
                        y = ny;
                        xi = nxi;
                        old_s = n_old;
                        nb = n_nb;
                    }
                }
            }
        }
        a.E[w] = e;
        a.D[w] = d;
        a.M[w] = m;
    }
    if (LDSHIST) {
        __syncthreads();
        // counts since the start (modulo 2^32; a wrap was carried already)
        for (uint32_t i = threadIdx.x; i <= N; i += blockDim.x) {
            const uint32_t c = hist[i];
            if (c != a.hist_preload)
                atomicAdd((unsigned long long *)&a.hist[i],
                          (unsigned long long)((int64_t)c - (int64_t)a.hist_preload));
        }
    }
}

template <int R, bool FP64, int KEYING, bool LDSHIST, bool JOINT>
__global__ __launch_bounds__(1024) void k_mcs_lane(McsArgs a, uint32_t lds_w)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    lane_run<R, FP64, KEYING, LDSHIST, JOINT>(a, lds_w, (uint64_t)blockIdx.x * blockDim.x, lds);
}

bool lane_eligible(const McsArgs &a, int fp64, size_t lds_max)
{
    const size_t tsz = fp64 ? 8 : 4;
    return a.L % 2 == 0 && a.L >= 4 && (a.N + 1) * tsz <= lds_max;
}

template <int R, bool FP64, int KEYING>
static hipError_t launch_lane_r(hipStream_t st, const Launch &cfg, const McsArgs &a)
{
    const int block = cfg.block;
    const size_t tsz = FP64 ? 8 : 4;
    const uint32_t lds_w = (uint32_t)(((a.N + 1) * tsz + 15) & ~(size_t)15);
    const size_t lds_h = (size_t)(a.N + 1) * 4;
    const bool hist_lds = lds_w + lds_h <= cfg.lds_max;
    const size_t bytes = lds_w + (hist_lds ? lds_h : 0);
    const bool joint = a.joint != nullptr;
    void (*k)(McsArgs, uint32_t);
    if (hist_lds)
        k = joint ? k_mcs_lane<R, FP64, KEYING, true, true> : k_mcs_lane<R, FP64, KEYING, true, false>;
    else
        k = joint ? k_mcs_lane<R, FP64, KEYING, false, true>
                  : k_mcs_lane<R, FP64, KEYING, false, false>;
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3((unsigned)((a.W + block - 1) / block)), dim3(block), bytes, st, a,
                       lds_w);
    return hipGetLastError();
}

hipError_t launch_mcs_lane(hipStream_t st, const Launch &cfg, const McsArgs &a)
{
    if (cfg.rounds == 10) {
        if (cfg.fp64)
            return cfg.keying ? launch_lane_r<10, true, 1>(st, cfg, a)
                              : launch_lane_r<10, true, 0>(st, cfg, a);
        return cfg.keying ? launch_lane_r<10, false, 1>(st, cfg, a)
                          : launch_lane_r<10, false, 0>(st, cfg, a);
    }
    if (cfg.fp64)
        return cfg.keying ? launch_lane_r<7, true, 1>(st, cfg, a)
                          : launch_lane_r<7, true, 0>(st, cfg, a);
    return cfg.keying ? launch_lane_r<7, false, 1>(st, cfg, a)
                      : launch_lane_r<7, false, 0>(st, cfg, a);
}

// ---------------------------------------------------------------------------
// batched engine, lane tiles: a persistent grid takes tiles of blockDim
// walkers of one instance from a shared counter, largest lattices first.
// ---------------------------------------------------------------------------
template <bool FP64>
__global__ __launch_bounds__(1024) void k_batch_lane(BatchArgs b)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    __shared__ uint32_t s_tile;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_tile = batch_take(b);
        __syncthreads();
        const uint32_t t = s_tile;
        if (t >= b.ntiles) break;
        const BatchTile tile = b.tiles[t];
        const BatchInst *I = b.inst + tile.inst;
        if (I->converged) continue;
        const McsArgs a = batch_mcs_args(b, *I);
        const bool hist_lds = I->lane_hist;
        if (hist_lds)
            lane_run<7, FP64, 0, true, false>(a, I->lane_lds_w, tile.wbase, lds);
        else
            lane_run<7, FP64, 0, false, false>(a, I->lane_lds_w, tile.wbase, lds);
    }
}

hipError_t launch_batch_lane(hipStream_t st, int fp64, const BatchArgs &b, uint32_t grid,
                             uint32_t block, size_t lds_bytes)
{
    void (*k)(BatchArgs) = fp64 ? k_batch_lane<true> : k_batch_lane<false>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds_bytes, st, b);
    return hipGetLastError();
}

hipError_t prepare_batch_lane(int fp64, size_t lds_bytes)
{
    void (*k)(BatchArgs) = fp64 ? k_batch_lane<true> : k_batch_lane<false>;
    return hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_bytes);
}

}  // namespace muca
