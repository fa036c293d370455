// muca_kernels.hip -- HIP kernels of libmuca for gfx950 (MI355X).
//
// Update-kernel design: every walker is a strict serial Metropolis chain
// (fixed site order, acceptance depends on the running D = sum s^2), so the
// parallelism is across walkers: one lane per walker, walkers interleaved
// site-major in HBM so that each site access of a wavefront is one
// contiguous 64-byte segment.  The D histogram is privatised per workgroup
// in LDS (uint32) and flushed to HBM with 64-bit atomics.  Nothing in the
// update is a dense contraction, so there is no MFMA here.
#include "kernels.h"
#include "moments.h"

#include "acclog.h"
#include "philox.h"

namespace muca {

#define MUCA_CTR3 0xBADC0DEDu

// ---------------------------------------------------------------------------
// initial lattice (muca_delta_multigpu.cu:340-351; produce2_multigpu.cu:126-138)
// one thread per (walker, block of four sites); stores are coalesced over walkers
// ---------------------------------------------------------------------------
// strat_W > 0 selects the stratified start (no counterpart in the reference):
// walker wid of strat_W gets a nonzero spin with probability
// p = (wid + 1/2) / strat_W, sign - for u < p/2 and + for p/2 <= u < p, from
// the same Philox stream, so that the walkers start spread over D = 0..N.
template <int R>
__global__ __launch_bounds__(256) void k_init_lattice(int8_t *lat, uint64_t W, uint32_t N,
                                                      int keying, uint32_t mm, uint32_t gpu,
                                                      uint32_t seed, uint32_t w_offset,
                                                      uint32_t strat_W)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nblk = (N + 3) / 4;
    if (t >= W * nblk) return;
    const uint64_t w = t % W;
    const uint32_t blk = (uint32_t)(t / W);
    const uint32_t wid = w_offset + (uint32_t)w;
    philox4 r;
    if (keying == 0)
        r = philox4x32<R>(gpu * 2u + blk + 1u, seed, 0xBADCAB1Eu, 0xBADC0DEDu,
                          10u * wid + mm * 2u, 0xdecafbadu);
    else
        r = philox4x32<R>(gpu * 1000u, seed, blk + 1u, 0xBADCAFE3u, wid * 10u, 0xdecafbadu);
    const double p = strat_W ? ((double)wid + 0.5) / (double)strat_W : 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t site = 4u * blk + (uint32_t)j;
        if (site < N) {
            const double u = (double)((float)(1u | (r.v[j] >> 8)) * 0x1p-24f);
            const int8_t sv = strat_W ? (int8_t)(u < p ? (u < 0.5 * p ? -1 : 1) : 0)
                                      : (int8_t)((u < 1 / 3.0) + (u < 2 / 3.0) - 1);
            lat[(uint64_t)site * W + w] = sv;
        }
    }
}

hipError_t launch_init_lattice(hipStream_t st, int8_t *lat, uint64_t W, uint32_t N,
                               int keying, int rounds, uint32_t mm, uint32_t gpu,
                               uint32_t seed, uint32_t w_offset, uint32_t strat_W)
{
    const uint64_t total = W * ((N + 3) / 4);
    const dim3 grid((unsigned)((total + 255) / 256));
    if (rounds == 10)
        hipLaunchKernelGGL(k_init_lattice<10>, grid, dim3(256), 0, st, lat, W, N, keying, mm,
                           gpu, seed, w_offset, strat_W);
    else
        hipLaunchKernelGGL(k_init_lattice<7>, grid, dim3(256), 0, st, lat, W, N, keying, mm,
                           gpu, seed, w_offset, strat_W);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// per-walker E_J, D, M (muca_delta_multigpu.cu:69-83,548-571; produce2_multigpu.cu:66-89)
// ---------------------------------------------------------------------------
// sixteen lanes per walker: lane l sums sites l, l+16, ...; the row of
// sixteen lanes is then reduced (integer sums, so the order does not matter)
__global__ __launch_bounds__(256) void k_state(const int8_t *lat, uint64_t W, uint32_t N,
                                               const uint32_t *naver, int32_t *E, int32_t *D,
                                               int32_t *M)
{
    const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const uint32_t tl = threadIdx.x & 15;
    if (w >= W) return;   // whole rows of 16 lanes share w
    int32_t e = 0, d = 0, m = 0;
    for (uint32_t i = tl; i < N; i += 16) {
        const int32_t s = lat[(uint64_t)i * W + w];
        const uint32_t *nv = naver + 4u * i;
        const int32_t nb = lat[(uint64_t)nv[0] * W + w] + lat[(uint64_t)nv[1] * W + w] +
                           lat[(uint64_t)nv[2] * W + w] + lat[(uint64_t)nv[3] * W + w];
        e += -s * nb;
        d += s * s;
        m += s;
    }
    for (int off = 8; off > 0; off >>= 1) {
        e += __shfl_xor(e, off, 16);
        d += __shfl_xor(d, off, 16);
        m += __shfl_xor(m, off, 16);
    }
    if (tl == 0) {
        E[w] = e >> 1;
        D[w] = d;
        if (M) M[w] = m;
    }
}

hipError_t launch_state(hipStream_t st, const int8_t *lat, uint64_t W, uint32_t N,
                        const uint32_t *naver, int32_t *E, int32_t *D, int32_t *M)
{
    hipLaunchKernelGGL(k_state, dim3((unsigned)((W * 16 + 255) / 256)), dim3(256), 0, st, lat, W,
                       N, naver, E, D, M);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The update kernel, reference site order from tables (any L).
// Per walker (muca_delta_multigpu.cu:735-829; produce2_multigpu.cu:401-496):
//   keys  weight : k1 = {2w, 0xdecafbad}, k3 = {2w+50000, 0x11cafead}
//         produce: k1 = {w, 0xdecafbad},  k3 = {w+40000, 0xdecafead}
//   ctr = {++c0, seed, iter, 0xBADC0DED} once per four attempts, c0 shared
//   by the therm and production phases, attempt index restarting at 0.
//   step  = +1 if r3 < 2^31 (u01fixedpt<float>(r3) < 0.5) else -1, new in {-1,0,1}
//   dE    = -(new - old) * (sum of the four neighbours)   [equals the
//           reference's ((old==0) + old*(new-old)) * (-tmp) * sum for every
//           old, new in {-1,0,1}, tmp being its temporary write]
//   diff  = (T)(-dE*beta) + lnW[D+dDel] - lnW[D]   (T = float | double)
//   accept iff diff >= 0 or acclog(r1) < diff
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T lnw_at(const McsArgs &a, int32_t i);
template <>
__device__ __forceinline__ float lnw_at<float>(const McsArgs &a, int32_t i)
{
    return a.lnw32[i];
}
template <>
__device__ __forceinline__ double lnw_at<double>(const McsArgs &a, int32_t i)
{
    return a.lnw64[i];
}

// n entries of D histogram bin d: LDS (uint32, a workgroup adds at most
// block * nup < 2^32 to a bin) or HBM
template <bool LDSHIST>
__device__ __forceinline__ void hist_add(const McsArgs &a, uint32_t *lds_hist, int32_t d,
                                         uint64_t n)
{
    if (n == 0) return;
    if (LDSHIST)
        atomicAdd(&lds_hist[d], (uint32_t)n);
    else
        atomicAdd((unsigned long long *)&a.hist[d], (unsigned long long)n);
}

// OBS: 0 = D histogram only; 1 = joint H(D,M) per walker block
// (produce2_multigpu.cu:489); 2 = exact per-D moments per walker block
template <int R, bool FP64, int KEYING, bool LDSHIST, int OBS>
__global__ __launch_bounds__(256) void k_mcs_generic(McsArgs a)
{
    typedef typename std::conditional<FP64, double, float>::type T;
    extern __shared__ uint32_t lds_hist[];
    const uint32_t nbins = a.N + 1;
    if (LDSHIST) {
        for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) lds_hist[b] = 0;
        __syncthreads();
    }
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < a.W) {
        const uint32_t wid = a.w_offset + (uint32_t)w;
        const uint32_t k1a = KEYING == 0 ? 2u * wid : wid;
        const uint32_t k1b = 0xdecafbadu;
        const uint32_t k3a = KEYING == 0 ? 2u * wid + 50000u : wid + 40000u;
        const uint32_t k3b = KEYING == 0 ? 0x11cafeadu : 0xdecafeadu;
        int8_t *s = a.lat + w;
        // walker counts are 32-bit (muca_params.walkers): site * W is one
        // 32 x 32 -> 64-bit multiply-add
        const uint32_t W = (uint32_t)a.W;
        int32_t e = a.E[w], d = a.D[w], m = a.M[w];
        uint64_t *jbase = nullptr;
        if (OBS == 1) jbase = a.joint + (w / a.block_walkers) * a.joint_stride;
        // per-D moment arrays of this walker's block, mstride words apart
        const uint64_t mstride = (uint64_t)a.N + 1;
        uint64_t *mbase = nullptr;
        if (OBS == 2) mbase = a.moms + (w / a.block_walkers) * (uint64_t)kMomWords * mstride;
        // the current D's moment sums and the previous D's: a walker that
        // steps back to the D it came from flushes nothing
        MomRun run{}, prev{};
        run.d = -1;
        prev.d = -1;
        // Histogram entries are counted in registers while the bin stays the
        // same (D, or (D, M) for the joint histogram, changes only on some
        // accepted moves) and added with one atomic when it changes: the
        // integer sums are the same, the atomics far fewer.
        // The bin in use and the one before it are both kept, so a walker
        // stepping back and forth between two bins adds nothing until it
        // reaches a third.
        uint64_t hrun = 0, hprun = 0, jrun = 0, jprun = 0;
        int32_t hbin = 0, hpbin = 0;
        int64_t jbin = 0, jpbin = 0;
        // Attempt t+1's site, neighbours and spins are loaded before attempt t
        // computes, so their memory latency overlaps attempt t's ln W gathers,
        // store and histogram atomic.  That is exact when attempt t's site is
        // neither attempt t+1's site nor one of its neighbours: for even L >= 4
        // consecutive sites of the order 0, 2, .., N-2, 1, 3, .., N-1 differ by
        // 2 (or are N-2 -> 1 and N-1 -> 0), never by 0, +-1, +-L, 1-L, L-1 or
        // N-L modulo N.  Other L load after the store.
        const bool ahead = (a.L % 2 == 0) && a.L >= 4;
        uint32_t c0 = 0;
        for (int phase = 0; phase < 2; phase++) {
            const uint64_t n = phase ? a.nup : a.therm;
            if (n == 0) continue;
            uint32_t pos = 0;
            // spins of the attempt about to run: site, and sum of neighbours
            uint32_t idx = a.order[0];
            int32_t old_s = s[(uint64_t)idx * W];
            int32_t nsum;
            {
                const uint32_t *nv = a.naver + 4u * idx;
                nsum = s[(uint64_t)nv[0] * W] + s[(uint64_t)nv[1] * W] + s[(uint64_t)nv[2] * W] +
                       s[(uint64_t)nv[3] * W];
            }
            philox4 r1{}, r3{};
            for (uint64_t i = 0; i < n; i++) {
                if ((i & 3) == 0) {
                    ++c0;
                    r1 = philox4x32<R>(c0, a.seed, a.iter, MUCA_CTR3, k1a, k1b);
                    r3 = philox4x32<R>(c0, a.seed, a.iter, MUCA_CTR3, k3a, k3b);
                }
                const uint32_t j = (uint32_t)(i & 3);
                const uint32_t u1 = j == 0 ? r1.v[0] : j == 1 ? r1.v[1] : j == 2 ? r1.v[2] : r1.v[3];
                const uint32_t u3 = j == 0 ? r3.v[0] : j == 1 ? r3.v[1] : j == 2 ? r3.v[2] : r3.v[3];
                const bool more = i + 1 < n;
                const uint32_t npos = (pos + 1 == a.N) ? 0 : pos + 1;
                uint32_t nidx = 0;
                int32_t nold = 0, nnsum = 0;
                if (ahead && more) {
                    nidx = a.order[npos];
                    const uint32_t *nv = a.naver + 4u * nidx;
                    nold = s[(uint64_t)nidx * W];
                    nnsum = s[(uint64_t)nv[0] * W] + s[(uint64_t)nv[1] * W] +
                            s[(uint64_t)nv[2] * W] + s[(uint64_t)nv[3] * W];
                }
                int32_t new_s = old_s + (u3 < 0x80000000u ? 1 : -1);
                new_s = new_s == 2 ? -1 : (new_s == -2 ? 1 : new_s);
                const int32_t dE = -(new_s - old_s) * nsum;
                const int32_t dDel = new_s * new_s - old_s * old_s;
                // D stays in [0, N] for a consistent (lattice, D) pair; the
                // clamp keeps every table access in bounds for any input
                const int32_t dn = min(max(d + dDel, 0), (int32_t)a.N);
                T diff = (T)(-(double)dE * a.beta);
                diff = diff + lnw_at<T>(a, dn);
                diff = diff - lnw_at<T>(a, d);
                // log(u) < diff (diff >= 0 included: log(u) < 0), screened with
                // the hardware log: exact unless diff lies within the proven
                // error band of la, where the correctly rounded log decides
                const float la = acclog_fast_u01(u1);
                const float band = acclog_fast_band(la);
                bool acc = (T)(la + band) < diff;
                const bool amb = !acc && !((T)(la - band) >= diff);
                if (__builtin_amdgcn_ballot_w64(amb)) {
                    if (amb) acc = (T)acclog_u01(u1) < diff;
                }
                if (acc) {
                    s[(uint64_t)idx * W] = (int8_t)new_s;
                    e += dE;
                    d = dn;
                    m += new_s - old_s;
                }
                if (phase) {
                    // with per-D moments the D histogram is the blocks' S0,
                    // summed after the launch (k_hist_from_moments)
                    if (OBS != 2) {
                        if (d != hbin) {
                            if (d == hpbin && hprun) {
                                const uint64_t tr = hrun;
                                hrun = hprun;
                                hprun = tr;
                                hpbin = hbin;
                            } else {
                                hist_add<LDSHIST>(a, lds_hist, hpbin, hprun);
                                hprun = hrun;
                                hpbin = hbin;
                                hrun = 0;
                            }
                            hbin = d;
                        }
                        hrun++;
                    }
                    if (OBS == 1) {
                        const int64_t mm = min(max((int64_t)m, -(int64_t)d), (int64_t)d);
                        const int64_t jb = (mm + (int64_t)d * d + 2 * (int64_t)d) / 2;
                        if (jb != jbin) {
                            if (jb == jpbin && jprun) {
                                const uint64_t tr = jrun;
                                jrun = jprun;
                                jprun = tr;
                                jpbin = jbin;
                            } else {
                                if (jprun) atomicAdd((unsigned long long *)&jbase[jpbin], jprun);
                                jprun = jrun;
                                jpbin = jbin;
                                jrun = 0;
                            }
                            jbin = jb;
                        }
                        jrun++;
                    }
                    if (OBS == 2) {
                        if (d != run.d) {
                            if (d == prev.d) {
                                const MomRun t = run;
                                run = prev;
                                prev = t;
                            } else {
                                mom_flush(prev, mbase, mstride);
                                prev = run;
                                run = MomRun{};
                                run.d = d;
                            }
                        }
                        mom_add(run, min(max(m, -d), d));
                    }
                }
                if (more) {
                    if (!ahead) {
                        nidx = a.order[npos];
                        const uint32_t *nv = a.naver + 4u * nidx;
                        nold = s[(uint64_t)nidx * W];
                        nnsum = s[(uint64_t)nv[0] * W] + s[(uint64_t)nv[1] * W] +
                                s[(uint64_t)nv[2] * W] + s[(uint64_t)nv[3] * W];
                    }
                    idx = nidx;
                    old_s = nold;
                    nsum = nnsum;
                }
                pos = npos;
            }
        }
        hist_add<LDSHIST>(a, lds_hist, hbin, hrun);
        hist_add<LDSHIST>(a, lds_hist, hpbin, hprun);
        if (OBS == 1 && jrun) atomicAdd((unsigned long long *)&jbase[jbin], jrun);
        if (OBS == 1 && jprun) atomicAdd((unsigned long long *)&jbase[jpbin], jprun);
        if (OBS == 2) {
            mom_flush(run, mbase, mstride);
            mom_flush(prev, mbase, mstride);
        }
        a.E[w] = e;
        a.D[w] = d;
        a.M[w] = m;
    }
    if (LDSHIST) {
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) {
            const uint32_t c = lds_hist[b];
            if (c) atomicAdd((unsigned long long *)&a.hist[b], (unsigned long long)c);
        }
    }
}

template <int R, bool FP64, int KEYING, bool LDSHIST, int OBS>
static void launch_generic_v(hipStream_t st, dim3 grid, int block, size_t lds, const McsArgs &a)
{
    auto k = k_mcs_generic<R, FP64, KEYING, LDSHIST, OBS>;
    if (LDSHIST)
        hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
    hipLaunchKernelGGL(k, grid, dim3(block), LDSHIST ? lds : 0, st, a);
}

template <int R, bool FP64, int KEYING, bool LDSHIST>
static void launch_generic_h(hipStream_t st, dim3 grid, int block, size_t lds, const McsArgs &a)
{
    // the produce path computes in FP64, so only FP64 builds the moments variant
    if (a.joint)
        launch_generic_v<R, FP64, KEYING, LDSHIST, 1>(st, grid, block, lds, a);
    else if (FP64 && a.moms)
        launch_generic_v<R, FP64, KEYING, LDSHIST, FP64 ? 2 : 0>(st, grid, block, lds, a);
    else
        launch_generic_v<R, FP64, KEYING, LDSHIST, 0>(st, grid, block, lds, a);
}

// D histogram of a launch that kept per-D moments: hist[D] = sum over blocks
// of S0[D] (every production attempt adds 1 to S0 of its walker's block)
__global__ __launch_bounds__(256) void k_hist_from_moments(const uint64_t *moms, uint32_t N,
                                                           uint32_t blocks, uint64_t *hist)
{
    const uint64_t nb = (uint64_t)N + 1;
    const uint64_t D = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (D >= nb) return;
    uint64_t h = 0;
    for (uint32_t b = 0; b < blocks; b++) h += moms[(uint64_t)b * kMomWords * nb + D];
    hist[D] = h;
}

template <int R, bool FP64, int KEYING>
static hipError_t launch_generic_r(hipStream_t st, const Launch &cfg, const McsArgs &a)
{
    const int block = (cfg.block > 0 && cfg.block <= 256) ? cfg.block : 256;   // launch bounds
    const dim3 grid((unsigned)((a.W + block - 1) / block));
    const size_t lds = (size_t)(a.N + 1) * sizeof(uint32_t);
    // uint32 LDS bins cannot overflow: a workgroup adds at most block * nup
    const bool lds_ok = lds <= cfg.lds_max && (uint64_t)block * a.nup < (1ull << 32);
    if (lds_ok)
        launch_generic_h<R, FP64, KEYING, true>(st, grid, block, lds, a);
    else
        launch_generic_h<R, FP64, KEYING, false>(st, grid, block, lds, a);
    if (FP64 && a.moms && !a.joint && a.nup > 0)
        return launch_hist_from_moments(st, a.moms, a.N, a.joint_blocks, a.hist);
    return hipGetLastError();
}

hipError_t launch_hist_from_moments(hipStream_t st, const uint64_t *moms, uint32_t N,
                                    uint32_t blocks, uint64_t *hist)
{
    const uint64_t nb = (uint64_t)N + 1;
    hipLaunchKernelGGL(k_hist_from_moments, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st,
                       moms, N, blocks, hist);
    return hipGetLastError();
}

template <bool FP64, int KEYING>
static hipError_t launch_generic_k(hipStream_t st, const Launch &cfg, const McsArgs &a)
{
    return cfg.rounds == 10 ? launch_generic_r<10, FP64, KEYING>(st, cfg, a)
                            : launch_generic_r<7, FP64, KEYING>(st, cfg, a);
}

hipError_t launch_mcs(hipStream_t st, const Launch &cfg, const McsArgs &a)
{
    if (a.W == 0 || (a.therm == 0 && a.nup == 0)) return hipSuccess;
    // per-D moments: the generic kernel, or (FP64) the team kernel
    const bool moments = a.moms != nullptr && a.joint == nullptr;
    if (cfg.kernel == 2 && (!moments || cfg.fp64)) {
        TeamLayout lay{};
        if (team_layout(a, cfg.fp64, cfg.lds_max, lay)) return launch_mcs_team(st, cfg, a, lay);
    }
    if (cfg.kernel == 3 && !moments && lane_eligible(a, cfg.fp64, cfg.lds_max))
        return launch_mcs_lane(st, cfg, a);
    if (cfg.fp64)
        return cfg.keying == 0 ? launch_generic_k<true, 0>(st, cfg, a)
                               : launch_generic_k<true, 1>(st, cfg, a);
    return cfg.keying == 0 ? launch_generic_k<false, 0>(st, cfg, a)
                           : launch_generic_k<false, 1>(st, cfg, a);
}

// ---------------------------------------------------------------------------
// acceptance log evaluated on the device (self-check)
// ---------------------------------------------------------------------------
__global__ void k_acclog(uint32_t k0, uint32_t n, float *out)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = acclog_u01((k0 + t) << 9);
}

hipError_t launch_acclog(hipStream_t st, uint32_t k0, uint32_t n, float *out)
{
    hipLaunchKernelGGL(k_acclog, dim3((n + 255) / 256), dim3(256), 0, st, k0, n, out);
    return hipGetLastError();
}

__global__ void k_acclog_fast(uint32_t k0, uint32_t n, float *out)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = acclog_fast_u01((k0 + t) << 9);
}

hipError_t launch_acclog_fast(hipStream_t st, uint32_t k0, uint32_t n, float *out)
{
    hipLaunchKernelGGL(k_acclog_fast, dim3((n + 255) / 256), dim3(256), 0, st, k0, n, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// produce path: exact per-D moments (S0, S1, limb words of S2 and S4) -> FP64
// in the layout of k_joint_moments' output
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_moments_fp64(const uint64_t *moms, uint32_t N,
                                                      uint32_t blocks, double *mom,
                                                      uint32_t *range_flag)
{
    const uint64_t nb = (uint64_t)N + 1;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb * blocks) return;
    const uint64_t b = t / nb, D = t - b * nb;
    const uint64_t *s = moms + b * kMomWords * nb + D;
    double *o = mom + b * 4 * nb + D;
    const double two32 = 4294967296.0, two64 = 18446744073709551616.0;
    if (s[0] >> 32) atomicOr(range_flag, 1u);   // limb words may have wrapped
    o[0] = (double)s[0];
    o[nb] = (double)s[nb];
    o[2 * nb] = (double)s[4 * nb] * two64 + ((double)s[3 * nb] * two32 + (double)s[2 * nb]);
    o[3 * nb] = (double)s[7 * nb] * two64 + ((double)s[6 * nb] * two32 + (double)s[5 * nb]);
}

hipError_t launch_moments_fp64(hipStream_t st, const uint64_t *moms, uint32_t N,
                               uint32_t blocks, double *mom, uint32_t *range_flag)
{
    const uint64_t n = ((uint64_t)N + 1) * blocks;
    hipLaunchKernelGGL(k_moments_fp64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       moms, N, blocks, mom, range_flag);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// produce path: joint H(D,M) -> per-D moments, one wavefront per (block, D);
// lanes stride over m so the reads of a shell are contiguous.
// mom layout: [block][4][N+1] with S0 = sum H, S1 = sum |m| H, S2 = sum m^2 H,
// S4 = sum m^4 H (FP64).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_joint_moments(const uint64_t *joint, uint32_t N,
                                                       uint64_t stride, uint32_t blocks,
                                                       double *mom)
{
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t nb = N + 1;
    if (wave >= nb * blocks) return;
    const uint32_t b = wave / (uint32_t)nb;
    const uint64_t D = wave % nb;
    const uint64_t *h = joint + b * stride + (D * D + D) / 2;   // m = -D at offset 0
    double a0 = 0, a1 = 0, a2 = 0, a4 = 0;
    for (uint64_t j = lane; j <= D; j += 64) {
        const double c = (double)h[j];
        const double am = fabs((double)(2 * (int64_t)j - (int64_t)D));
        const double m2 = am * am;
        a0 += c;
        a1 += am * c;
        a2 += m2 * c;
        a4 += m2 * m2 * c;
    }
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_down(a0, off, 64);
        a1 += __shfl_down(a1, off, 64);
        a2 += __shfl_down(a2, off, 64);
        a4 += __shfl_down(a4, off, 64);
    }
    if (lane == 0) {
        double *o = mom + (uint64_t)b * 4 * nb;
        o[D] = a0;
        o[nb + D] = a1;
        o[2 * nb + D] = a2;
        o[3 * nb + D] = a4;
    }
}

hipError_t launch_joint_moments(hipStream_t st, const uint64_t *joint, uint32_t N,
                                uint64_t stride, uint32_t blocks, double *mom)
{
    const uint64_t waves = (uint64_t)(N + 1) * blocks;
    hipLaunchKernelGGL(k_joint_moments, dim3((unsigned)((waves * 64 + 255) / 256)), dim3(256), 0,
                       st, joint, N, stride, blocks, mom);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// reweighting (produce2_multigpu.cu:295-297,391-400,498-518), one workgroup
// per (delta, block): shift = max_D(-lnW_D - beta*delta*D), then
// Z = sum_D exp(-beta*delta*D - lnW_D - shift) S0_D, A_k likewise with S_k.
// Three planes of ndelta*blocks values: out[k*plane + id*blocks + b] with
// k = 0, 1, 2 for <|M|>, chi, Binder cumulant.
// ---------------------------------------------------------------------------
__device__ static double block_reduce_sum(double v, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int wv = threadIdx.x / 64, ln = threadIdx.x & 63;
    __syncthreads();
    if (ln == 0) sh[wv] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / 64); i++) t += sh[i];
    return t;
}

__device__ static double block_reduce_max(double v, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    const int wv = threadIdx.x / 64, ln = threadIdx.x & 63;
    __syncthreads();
    if (ln == 0) sh[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sh[0];
        for (int i = 1; i < (int)(blockDim.x / 64); i++) t = fmax(t, sh[i]);
        sh[8] = t;
    }
    __syncthreads();
    return sh[8];
}

__global__ __launch_bounds__(256) void k_reweight(const double *mom, const double *lnw,
                                                  uint32_t N, uint32_t blocks, double beta,
                                                  const double *deltas, double *out)
{
    __shared__ double sh[16];
    const uint32_t id = blockIdx.x / blocks, b = blockIdx.x % blocks;
    const double bd = beta * deltas[id];
    const uint64_t nb = N + 1;
    // The shift cancels in every ratio; the reference takes it over all bins
    // (produce2_multigpu.cu:391-400), which underflows every visited bin to 0
    // (and the ratios to 0/0) when the maximum sits at an unvisited D.  The
    // maximum over the block's visited bins gives the same ratios wherever
    // the reference's are finite, and finite ratios wherever a bin is visited;
    // unvisited bins are left out of the sums (their terms are 0 x exp(...)).
    const double *S = mom + (uint64_t)b * 4 * nb;
    double mx = -INFINITY;
    for (uint32_t i = threadIdx.x; i <= N; i += blockDim.x)
        if (S[i] > 0) mx = fmax(mx, -lnw[i] - bd * (double)i);
    mx = block_reduce_max(mx, sh);
    double z = 0, a1 = 0, a2 = 0, a4 = 0;
    for (uint32_t i = threadIdx.x; i <= N; i += blockDim.x) {
        if (!(S[i] > 0)) continue;   // unvisited: contributes nothing (and may overflow)
        const double bz = exp(-bd * (double)i - lnw[i] - mx);
        z += bz * S[i];
        a1 += bz * S[nb + i];
        a2 += bz * S[2 * nb + i];
        a4 += bz * S[3 * nb + i];
    }
    z = block_reduce_sum(z, sh);
    a1 = block_reduce_sum(a1, sh);
    a2 = block_reduce_sum(a2, sh);
    a4 = block_reduce_sum(a4, sh);
    if (threadIdx.x == 0) {
        const double m1 = a1 / z, m2 = a2 / z, m4 = a4 / z;
        const uint64_t plane = (uint64_t)gridDim.x;
        double *o = out + (uint64_t)id * blocks + b;
        o[0] = m1;
        o[plane] = (m2 - m1 * m1) * beta / (double)N;
        o[2 * plane] = 1.0 - m4 / (3.0 * m2 * m2);
    }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

hipError_t launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t n, uint32_t v)
{
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_reweight(hipStream_t st, const double *mom, const double *lnw, uint32_t N,
                           uint32_t blocks, double beta, const double *deltas,
                           uint32_t ndelta, double *out)
{
    hipLaunchKernelGGL(k_reweight, dim3(ndelta * blocks), dim3(256), 0, st, mom, lnw, N, blocks,
                       beta, deltas, out);
    return hipGetLastError();
}

}  // namespace muca
