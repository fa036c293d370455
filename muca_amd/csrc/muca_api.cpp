// muca_api.cpp -- the C ABI of libmuca (include/muca.h): contexts, device
// buffers, launches and the O(N) host-side pieces of the weight iteration.
#include "../../include/muca.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "acclog.h"
#include "ctx.h"
#include "kernels.h"

// the thread's last error message (muca_last_error)
thread_local std::string g_muca_err;

namespace {

// neighbour and order tables exactly as muca_delta_multigpu.cu:157-175
void make_tables(uint32_t L, std::vector<uint32_t> &naver, std::vector<uint32_t> &order)
{
    const uint32_t N = L * L;
    naver.assign(4ull * N, 0);
    order.assign(N, 0);
    for (uint32_t i = 0; i < N; i++) {
        int right = (int)i + 1, left = (int)i - 1, up = (int)(i + L), down = (int)i - (int)L;
        if (right % (int)L == 0) right -= (int)L;
        if (i % L == 0) left += (int)L;
        if (up > (int)N - 1) up -= (int)N;
        if (down < 0) down += (int)N;
        naver[4ull * i + 0] = (uint32_t)right;
        naver[4ull * i + 1] = (uint32_t)left;
        naver[4ull * i + 2] = (uint32_t)up;
        naver[4ull * i + 3] = (uint32_t)down;
        const uint32_t half = N / 2;
        const uint32_t blk = half ? (i / half) % 2 : 0;
        order[i] = ((0 == blk) * (2 * i)) % N + ((1 == blk) * (2 * i + 1)) % N;
    }
}

void free_ctx(muca_ctx *c)
{
    if (!c) return;
    if (c->p.device >= 0) hipSetDevice(c->p.device);
    void *bufs[] = {c->lat, c->E, c->D, c->M, c->lnw32, c->lnw64, c->order, c->naver,
                    c->hist, c->joint, c->moms, c->mom, c->deltas, c->rw_out,
                    c->rw_flag};
    for (void *b : bufs)
        if (b) hipFree(b);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
}

template <typename T>
int dev_alloc(T **p, uint64_t count, const char *what)
{
    hipError_t e = hipMalloc((void **)p, count * sizeof(T) ? count * sizeof(T) : 1);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(MUCA_E_NOMEM, std::string("hipMalloc(") + what + ", " +
                                      std::to_string(count * sizeof(T)) +
                                      " bytes): " + hipGetErrorString(e));
    }
    return MUCA_OK;
}

}  // namespace

extern "C" {

const char *muca_last_error(void)
{
    return g_muca_err.c_str();
}

int muca_default_walkers(int32_t device, uint32_t *walkers)
{
    if (!walkers) return fail(MUCA_E_ARG, "walkers is NULL");
    int mt = 0, cu = 0;
    HIP_TRY(hipDeviceGetAttribute(&mt, hipDeviceAttributeMaxThreadsPerMultiProcessor, device));
    HIP_TRY(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device));
    *walkers = (uint32_t)mt * (uint32_t)cu;
    return MUCA_OK;
}

int muca_create(const muca_params *params, muca_ctx **out)
{
    if (!params || !out) return fail(MUCA_E_ARG, "NULL argument");
    *out = nullptr;
    const muca_params &p = *params;
    if (p.L < 2 || p.L > 46340) return fail(MUCA_E_ARG, "L must be in [2, 46340]");
    if (p.rounds != 7 && p.rounds != 10) return fail(MUCA_E_ARG, "rounds must be 7 or 10");
    if (p.precision != MUCA_REF32 && p.precision != MUCA_FP64)
        return fail(MUCA_E_ARG, "precision must be MUCA_REF32 or MUCA_FP64");
    if (p.keying != MUCA_KEYING_WEIGHT && p.keying != MUCA_KEYING_PRODUCE)
        return fail(MUCA_E_ARG, "keying must be MUCA_KEYING_WEIGHT or MUCA_KEYING_PRODUCE");
    if (!(p.beta > 0) || !std::isfinite(p.beta)) return fail(MUCA_E_ARG, "beta must be > 0");
    if (p.observables != MUCA_OBS_JOINT && p.observables != MUCA_OBS_MOMENTS)
        return fail(MUCA_E_ARG, "observables must be MUCA_OBS_JOINT or MUCA_OBS_MOMENTS");
    const bool moments = p.joint_blocks > 0 && p.observables == MUCA_OBS_MOMENTS;
    if (moments && p.precision != MUCA_FP64)
        return fail(MUCA_E_ARG, "per-D moments are accumulated in MUCA_FP64 mode only");
    if (moments && p.kernel == MUCA_KERNEL_LANE)
        return fail(MUCA_E_ARG, "per-D moments are accumulated by the generic and team kernels only");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (p.device < 0 || p.device >= ndev)
        return fail(MUCA_E_ARG, "device " + std::to_string(p.device) + " not present (" +
                                    std::to_string(ndev) + " visible)");
    HIP_TRY(hipSetDevice(p.device));
    muca_ctx *c = new muca_ctx();
    c->p = p;
    c->N = p.L * p.L;
    uint32_t W = p.walkers;
    if (W == 0) {
        int rc = muca_default_walkers(p.device, &W);
        if (rc) {
            free_ctx(c);
            return rc;
        }
    }
    c->W = W;
    c->p.walkers = W;
    int cu = 0, lds = 0;
    hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, p.device);
    hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, p.device);
    c->cus = (uint32_t)cu;
    c->cfg.rounds = (int)p.rounds;
    c->cfg.fp64 = p.precision == MUCA_FP64;
    c->cfg.keying = p.keying;
    // Kernel choice (measured on MI355X, L = 72): the team kernel keeps every
    // SIMD busy when walkers are few (31 attempts/ns at 4096 walkers against
    // 6.7 for one lane per walker); one lane per walker wins once there are
    // enough wavefronts to hide its memory latency (58/ns at 65536 walkers
    // against 28).  The crossover sits near 128 walkers per compute unit.
    // Beyond that the lane kernel leads (213 against 147 at 524288 walkers),
    // with 512-lane workgroups once there are enough of them to fill every
    // compute unit twice (279 at 524288 walkers; 256 lanes stay ahead at
    // 65536).  muca_autotune replaces this rule by measurement.
    const uint64_t ncu = (uint64_t)(cu > 0 ? cu : 256);
    if (p.kernel == MUCA_KERNEL_AUTO) {
        // eligibility as the launch decides it (layout in this device's LDS)
        muca::McsArgs probe{};
        probe.N = p.L * p.L;
        probe.L = p.L;
        probe.W = W;
        probe.nup = 1;
        const size_t lds_probe = (size_t)(lds > 0 ? lds : 65536);
        muca::TeamLayout lay_probe;
        const bool team_ok = muca::team_layout(probe, p.precision == MUCA_FP64, lds_probe, lay_probe);
        const bool lane_ok = muca::lane_eligible(probe, p.precision == MUCA_FP64, lds_probe);
        if (moments)
            c->cfg.kernel = team_ok && (uint64_t)W <= 128ull * ncu ? MUCA_KERNEL_TEAM
                                                                   : MUCA_KERNEL_GENERIC;
        else if (team_ok && (uint64_t)W <= 128ull * ncu)
            c->cfg.kernel = MUCA_KERNEL_TEAM;
        else if (lane_ok)
            c->cfg.kernel = MUCA_KERNEL_LANE;
        else
            c->cfg.kernel = MUCA_KERNEL_GENERIC;
    } else {
        c->cfg.kernel = p.kernel;
    }
    if (p.block != 0 && (p.block < 64 || p.block > 1024 || p.block % 64 != 0)) {
        delete c;
        return fail(MUCA_E_ARG, "block must be 0 or a multiple of 64 in [64, 1024]");
    }
    c->cfg.block = p.block ? (int)p.block : 256;
    if (p.block == 0 && c->cfg.kernel == MUCA_KERNEL_GENERIC) {
        // one walker per lane and no LDS lattice: the grid should reach every
        // compute unit before workgroups grow (16 384 walkers at 256 lanes
        // would occupy 64 of 256 compute units)
        while (c->cfg.block > 64 && (uint64_t)W < (uint64_t)c->cfg.block * ncu)
            c->cfg.block /= 2;
    }
    if (p.block == 0 && c->cfg.kernel == MUCA_KERNEL_LANE && (uint64_t)W >= 1024ull * ncu) {
        // Enough walkers for two 512-lane workgroups per compute unit; when
        // the lane kernel's LDS (ln W and the histogram) admits only one
        // workgroup per compute unit, 1024 lanes keep twice the wavefronts
        // resident (measured at L = 128, 262 144 walkers: 273 attempts/ns
        // against 158 with 512; at L = 72, 524 288 walkers, 512 leads: 279).
        const size_t tsz = p.precision == MUCA_FP64 ? 8 : 4;
        const size_t lds_lane = (size_t)(p.L * p.L + 1) * (tsz + 4);
        const size_t lds_cu = (size_t)(lds > 0 ? lds : 65536);
        c->cfg.block = lds_lane * 2 > lds_cu ? 1024 : 512;
    }
    c->cfg.lds_max = (size_t)(lds > 0 ? lds : 65536);
    const uint64_t N = c->N;
    int rc = MUCA_OK;
    if (!rc) rc = dev_alloc(&c->lat, N * W, "lattice");
    if (!rc) rc = dev_alloc(&c->E, W, "E");
    if (!rc) rc = dev_alloc(&c->D, W, "D");
    if (!rc) rc = dev_alloc(&c->M, W, "M");
    if (!rc) rc = dev_alloc(&c->lnw32, N + 1, "lnw32");
    if (!rc) rc = dev_alloc(&c->lnw64, N + 1, "lnw64");
    if (!rc) rc = dev_alloc(&c->order, N, "order");
    if (!rc) rc = dev_alloc(&c->naver, 4 * N, "naver");
    if (!rc) rc = dev_alloc(&c->hist, N + 1, "hist");
    if (!rc && p.joint_blocks > 0) {
        if (p.joint_blocks > W) {
            free_ctx(c);
            return fail(MUCA_E_ARG, "joint_blocks exceeds walkers");
        }
        c->block_walkers = (W + p.joint_blocks - 1) / p.joint_blocks;
        if (p.observables == MUCA_OBS_MOMENTS) {
            rc = dev_alloc(&c->moms, (uint64_t)muca::kMomWords * (N + 1) * p.joint_blocks,
                           "per-D moments");
        } else {
            c->joint_stride = (N + 1) * (N + 2) / 2;
            rc = dev_alloc(&c->joint, c->joint_stride * p.joint_blocks, "joint histogram");
        }
    }
    if (rc) {
        free_ctx(c);
        return rc;
    }
    std::vector<uint32_t> naver, order;
    make_tables(p.L, naver, order);
    hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess)
        e = hipMemcpy(c->naver, naver.data(), naver.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(c->order, order.data(), order.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(c->lat, 0, N * W);
    if (e == hipSuccess) e = hipMemset(c->E, 0, W * 4);
    if (e == hipSuccess) e = hipMemset(c->D, 0, W * 4);
    if (e == hipSuccess) e = hipMemset(c->M, 0, W * 4);
    if (e == hipSuccess && c->moms)
        e = hipMemset(c->moms, 0, (uint64_t)muca::kMomWords * (N + 1) * p.joint_blocks * 8);
    if (e != hipSuccess) {
        free_ctx(c);
        return fail(MUCA_E_HIP, std::string("context setup: ") + hipGetErrorString(e));
    }
    *out = c;
    return MUCA_OK;
}

int muca_destroy(muca_ctx *c)
{
    free_ctx(c);
    return MUCA_OK;
}

int muca_query(const muca_ctx *c, muca_info *info)
{
    if (!c || !info) return fail(MUCA_E_ARG, "NULL argument");
    info->L = c->p.L;
    info->N = c->N;
    info->walkers = (uint32_t)c->W;
    info->nbins = c->N + 1;
    info->device = c->p.device;
    info->kernel = c->cfg.kernel;
    info->cus = c->cus;
    info->block = (uint32_t)c->cfg.block;
    info->lattice_bytes = (uint64_t)c->N * c->W;
    info->joint_bins = c->joint_stride * c->p.joint_blocks;
    info->lds_per_block = (uint32_t)c->cfg.lds_max;
    muca::McsArgs a{};
    a.N = c->N;
    a.L = c->p.L;
    a.W = c->W;
    a.nup = 1;
    muca::TeamLayout lay{};
    info->lds_lattice = c->cfg.kernel == MUCA_KERNEL_TEAM &&
                        muca::team_layout(a, c->cfg.fp64, c->cfg.lds_max, lay) && lay.lat_lds;
    return MUCA_OK;
}

int muca_init_lattice(muca_ctx *c, uint32_t mm, uint32_t gpu_index, uint32_t seed)
{
    if (!c) return fail(MUCA_E_ARG, "NULL context");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(muca::launch_init_lattice(c->st, c->lat, c->W, c->N, c->p.keying, (int)c->p.rounds,
                                      mm, gpu_index, seed, 0));
    HIP_TRY(hipStreamSynchronize(c->st));
    return MUCA_OK;
}

int muca_init_lattice_stratified(muca_ctx *c, uint32_t mm, uint32_t gpu_index, uint32_t seed)
{
    if (!c) return fail(MUCA_E_ARG, "NULL context");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(muca::launch_init_lattice(c->st, c->lat, c->W, c->N, c->p.keying, (int)c->p.rounds,
                                      mm, gpu_index, seed, 0, (uint32_t)c->W));
    HIP_TRY(hipStreamSynchronize(c->st));
    return MUCA_OK;
}

int muca_set_lattice(muca_ctx *c, const int8_t *src, uint64_t bytes)
{
    if (!c || !src) return fail(MUCA_E_ARG, "NULL argument");
    if (bytes != (uint64_t)c->N * c->W) return fail(MUCA_E_ARG, "lattice size mismatch");
    for (uint64_t i = 0; i < bytes; i++)
        if (src[i] < -1 || src[i] > 1) return fail(MUCA_E_ARG, "spins must be in {-1,0,1}");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipMemcpy(c->lat, src, bytes, hipMemcpyHostToDevice));
    return MUCA_OK;
}

int muca_get_lattice(muca_ctx *c, int8_t *dst, uint64_t bytes)
{
    if (!c || !dst) return fail(MUCA_E_ARG, "NULL argument");
    if (bytes != (uint64_t)c->N * c->W) return fail(MUCA_E_ARG, "lattice size mismatch");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->st));
    HIP_TRY(hipMemcpy(dst, c->lat, bytes, hipMemcpyDeviceToHost));
    return MUCA_OK;
}

int muca_compute_state(muca_ctx *c)
{
    if (!c) return fail(MUCA_E_ARG, "NULL context");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(muca::launch_state(c->st, c->lat, c->W, c->N, c->naver, c->E, c->D, c->M));
    HIP_TRY(hipStreamSynchronize(c->st));
    return MUCA_OK;
}

int muca_get_state(muca_ctx *c, int32_t *E, int32_t *D, int32_t *M)
{
    if (!c) return fail(MUCA_E_ARG, "NULL context");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->st));
    if (E) HIP_TRY(hipMemcpy(E, c->E, c->W * 4, hipMemcpyDeviceToHost));
    if (D) HIP_TRY(hipMemcpy(D, c->D, c->W * 4, hipMemcpyDeviceToHost));
    if (M) HIP_TRY(hipMemcpy(M, c->M, c->W * 4, hipMemcpyDeviceToHost));
    return MUCA_OK;
}

int muca_set_state(muca_ctx *c, const int32_t *E, const int32_t *D, const int32_t *M)
{
    if (!c) return fail(MUCA_E_ARG, "NULL context");
    if (D)
        for (uint64_t w = 0; w < c->W; w++)
            if (D[w] < 0 || (uint32_t)D[w] > c->N) return fail(MUCA_E_ARG, "D out of [0, N]");
    HIP_TRY(hipSetDevice(c->p.device));
    if (E) HIP_TRY(hipMemcpy(c->E, E, c->W * 4, hipMemcpyHostToDevice));
    if (D) HIP_TRY(hipMemcpy(c->D, D, c->W * 4, hipMemcpyHostToDevice));
    if (M) HIP_TRY(hipMemcpy(c->M, M, c->W * 4, hipMemcpyHostToDevice));
    return MUCA_OK;
}

int muca_set_log_weight(muca_ctx *c, const double *lnw, uint32_t nbins)
{
    if (!c || !lnw) return fail(MUCA_E_ARG, "NULL argument");
    if (nbins != c->N + 1) return fail(MUCA_E_ARG, "nbins must be N+1");
    std::vector<float> f(nbins);
    for (uint32_t i = 0; i < nbins; i++) {
        if (!std::isfinite(lnw[i])) return fail(MUCA_E_ARG, "ln W must be finite");
        f[i] = (float)lnw[i];
    }
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipMemcpy(c->lnw64, lnw, nbins * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->lnw32, f.data(), nbins * sizeof(float), hipMemcpyHostToDevice));
    c->have_weights = true;
    return MUCA_OK;
}

int muca_run(muca_ctx *c, uint32_t iter, uint32_t seed, uint64_t therm, uint64_t nup,
             uint64_t *hist_out, float *kernel_ms)
{
    if (!c) return fail(MUCA_E_ARG, "NULL context");
    if (!c->have_weights) return fail(MUCA_E_STATE, "muca_set_log_weight was not called");
    HIP_TRY(hipSetDevice(c->p.device));
    const uint64_t N = c->N;
    HIP_TRY(hipMemsetAsync(c->hist, 0, (N + 1) * sizeof(uint64_t), c->st));
    const bool joint = c->joint && c->p.precision == MUCA_FP64 && nup > 0;
    if (joint)
        HIP_TRY(hipMemsetAsync(c->joint, 0, c->joint_stride * c->p.joint_blocks * 8, c->st));
    const bool moments = c->moms && nup > 0;
    if (moments && (nup >> 32))
        return fail(MUCA_E_ARG, "per-D moments take nupdates < 2^32 per launch");
    if (moments)
        HIP_TRY(hipMemsetAsync(c->moms, 0,
                               (uint64_t)muca::kMomWords * (N + 1) * c->p.joint_blocks * 8, c->st));
    muca::McsArgs a{};
    a.lat = c->lat;
    a.E = c->E;
    a.D = c->D;
    a.M = c->M;
    a.lnw32 = c->lnw32;
    a.lnw64 = c->lnw64;
    a.order = c->order;
    a.naver = c->naver;
    a.hist = c->hist;
    a.joint = joint ? c->joint : nullptr;
    a.moms = moments ? c->moms : nullptr;
    a.W = c->W;
    a.N = c->N;
    a.L = c->p.L;
    a.w_offset = 0;
    a.iter = iter;
    a.seed = seed;
    a.therm = therm;
    a.nup = nup;
    a.beta = c->p.beta;
    a.joint_stride = c->joint_stride;
    a.block_walkers = c->block_walkers ? c->block_walkers : c->W;
    a.joint_blocks = c->p.joint_blocks;
    a.hist_preload = c->hist_preload;
    HIP_TRY(hipEventRecord(c->ev0, c->st));
    HIP_TRY(muca::launch_mcs(c->st, c->cfg, a));
    HIP_TRY(hipEventRecord(c->ev1, c->st));
    HIP_TRY(hipEventSynchronize(c->ev1));
    if (kernel_ms) HIP_TRY(hipEventElapsedTime(kernel_ms, c->ev0, c->ev1));
    if (hist_out)
        HIP_TRY(hipMemcpy(hist_out, c->hist, (N + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return MUCA_OK;
}

int muca_set_test_hist_preload(muca_ctx *c, uint32_t preload)
{
    if (!c) return fail(MUCA_E_ARG, "NULL context");
    c->hist_preload = preload;
    return MUCA_OK;
}

int muca_autotune(muca_ctx *c, uint64_t attempts, int32_t *kernel, uint32_t *block, float *ms)
{
    if (!c) return fail(MUCA_E_ARG, "NULL context");
    if (!c->have_weights) return fail(MUCA_E_STATE, "muca_set_log_weight was not called");
    if (attempts == 0) return fail(MUCA_E_ARG, "attempts must be > 0");
    HIP_TRY(hipSetDevice(c->p.device));
    const uint64_t N = c->N, W = c->W;
    int8_t *lat0 = nullptr;
    int32_t *st0 = nullptr;
    int rc = dev_alloc(&lat0, N * W, "autotune lattice copy");
    if (rc) return rc;
    rc = dev_alloc(&st0, 3 * W, "autotune state copy");
    if (rc) {
        hipFree(lat0);
        return rc;
    }
    auto save_restore = [&](bool save) -> hipError_t {
        hipError_t e;
        if (save) {
            e = hipMemcpyAsync(lat0, c->lat, N * W, hipMemcpyDeviceToDevice, c->st);
            if (e == hipSuccess) e = hipMemcpyAsync(st0, c->E, W * 4, hipMemcpyDeviceToDevice, c->st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(st0 + W, c->D, W * 4, hipMemcpyDeviceToDevice, c->st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(st0 + 2 * W, c->M, W * 4, hipMemcpyDeviceToDevice, c->st);
        } else {
            e = hipMemcpyAsync(c->lat, lat0, N * W, hipMemcpyDeviceToDevice, c->st);
            if (e == hipSuccess) e = hipMemcpyAsync(c->E, st0, W * 4, hipMemcpyDeviceToDevice, c->st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(c->D, st0 + W, W * 4, hipMemcpyDeviceToDevice, c->st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(c->M, st0 + 2 * W, W * 4, hipMemcpyDeviceToDevice, c->st);
        }
        return e;
    };
    muca::McsArgs a{};
    a.lat = c->lat;
    a.E = c->E;
    a.D = c->D;
    a.M = c->M;
    a.lnw32 = c->lnw32;
    a.lnw64 = c->lnw64;
    a.order = c->order;
    a.naver = c->naver;
    a.hist = c->hist;
    a.W = W;
    a.N = c->N;
    a.L = c->p.L;
    a.iter = 0xFFFFFFFFu;
    a.seed = 0;
    a.therm = attempts - attempts / 2;
    a.nup = attempts / 2;
    a.beta = c->p.beta;
    a.block_walkers = W;
    struct Cand {
        int kernel, block;
    };
    std::vector<Cand> cands;
    muca::TeamLayout lay{};
    const bool moments = c->moms != nullptr;   // accumulated by the generic and team kernels
    if (moments) {
        // timed as production runs: with the per-D moments
        if (a.nup >> 32) {
            hipFree(lat0);
            hipFree(st0);
            return fail(MUCA_E_ARG, "per-D moments take nupdates < 2^32 per launch");
        }
        a.moms = c->moms;
        a.joint_blocks = c->p.joint_blocks;
        a.block_walkers = c->block_walkers ? c->block_walkers : c->W;
    }
    if (muca::team_layout(a, c->cfg.fp64, c->cfg.lds_max, lay))
        cands.push_back({MUCA_KERNEL_TEAM, 256});
    if (!moments && muca::lane_eligible(a, c->cfg.fp64, c->cfg.lds_max))
        for (int b : {256, 512, 1024}) cands.push_back({MUCA_KERNEL_LANE, b});
    for (int b : {64, 128, 256}) cands.push_back({MUCA_KERNEL_GENERIC, b});
    hipError_t e = save_restore(true);
    float best = 0;
    Cand pick = {c->cfg.kernel, c->cfg.block};
    for (const Cand &cd : cands) {
        muca::Launch cfg = c->cfg;
        cfg.kernel = cd.kernel;
        cfg.block = cd.block;
        float t = 0;
        for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {
            e = save_restore(false);
            if (e == hipSuccess) e = hipMemsetAsync(c->hist, 0, (N + 1) * 8, c->st);
            if (e == hipSuccess && moments)
                e = hipMemsetAsync(c->moms, 0,
                                   (uint64_t)muca::kMomWords * (N + 1) * c->p.joint_blocks * 8,
                                   c->st);
            if (e == hipSuccess) e = hipEventRecord(c->ev0, c->st);
            if (e == hipSuccess) e = muca::launch_mcs(c->st, cfg, a);
            if (e == hipSuccess) e = hipEventRecord(c->ev1, c->st);
            if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
            if (e == hipSuccess) e = hipEventElapsedTime(&t, c->ev0, c->ev1);
        }
        if (e != hipSuccess) break;
        if (best == 0 || t < best) {
            best = t;
            pick = cd;
        }
    }
    if (e == hipSuccess) e = save_restore(false);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    hipFree(lat0);
    hipFree(st0);
    if (e != hipSuccess) return fail(MUCA_E_HIP, std::string("autotune: ") + hipGetErrorString(e));
    c->cfg.kernel = pick.kernel;
    c->cfg.block = pick.block;
    if (kernel) *kernel = pick.kernel;
    if (block) *block = (uint32_t)pick.block;
    if (ms) *ms = best;
    return MUCA_OK;
}

int muca_get_hist(muca_ctx *c, uint64_t *hist, uint32_t nbins, float *copy_ms)
{
    if (!c || !hist) return fail(MUCA_E_ARG, "NULL argument");
    if (nbins != c->N + 1) return fail(MUCA_E_ARG, "nbins must be N+1");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipEventRecord(c->ev0, c->st));
    HIP_TRY(hipMemcpyAsync(hist, c->hist, (uint64_t)nbins * 8, hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipEventRecord(c->ev1, c->st));
    HIP_TRY(hipEventSynchronize(c->ev1));
    if (copy_ms) HIP_TRY(hipEventElapsedTime(copy_ms, c->ev0, c->ev1));
    return MUCA_OK;
}

int muca_get_joint(muca_ctx *c, uint32_t block, uint64_t *dst, uint64_t nbins)
{
    if (!c || !dst) return fail(MUCA_E_ARG, "NULL argument");
    if (!c->joint) return fail(MUCA_E_STATE, "context has no joint histogram");
    if (block >= c->p.joint_blocks) return fail(MUCA_E_ARG, "block out of range");
    if (nbins != c->joint_stride) return fail(MUCA_E_ARG, "nbins must be (N+1)(N+2)/2");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->st));
    HIP_TRY(hipMemcpy(dst, c->joint + (uint64_t)block * c->joint_stride, nbins * 8,
                      hipMemcpyDeviceToHost));
    return MUCA_OK;
}

int muca_get_moments(muca_ctx *c, uint32_t block, uint64_t *dst, uint64_t nwords)
{
    if (!c || !dst) return fail(MUCA_E_ARG, "NULL argument");
    if (!c->moms) return fail(MUCA_E_STATE, "context has no per-D moments");
    if (block >= c->p.joint_blocks) return fail(MUCA_E_ARG, "block out of range");
    const uint64_t nb = (uint64_t)c->N + 1, n = (uint64_t)muca::kMomWords * nb;
    if (nwords != 6 * nb) return fail(MUCA_E_ARG, "nwords must be 6 (N+1)");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->st));
    std::vector<uint64_t> w(n);
    HIP_TRY(hipMemcpy(w.data(), c->moms + (uint64_t)block * n, n * 8, hipMemcpyDeviceToHost));
    // limb words w0 + w1 2^32 + w2 2^64 -> (low, high) words of a 128-bit integer
    for (uint64_t d = 0; d < nb; d++) {
        if (w[d] >> 32) return fail(MUCA_E_RANGE, "a D bin holds 2^32 or more entries");
        dst[d] = w[d];
        dst[nb + d] = w[nb + d];
        for (int k = 0; k < 2; k++) {
            const uint64_t *l = &w[(2 + 3 * k) * nb + d];
            const unsigned __int128 v = (unsigned __int128)l[0] +
                                        ((unsigned __int128)l[nb] << 32) +
                                        ((unsigned __int128)l[2 * nb] << 64);
            dst[(2 + 2 * k) * nb + d] = (uint64_t)v;
            dst[(3 + 2 * k) * nb + d] = (uint64_t)(v >> 64);
        }
    }
    return MUCA_OK;
}

int muca_reweight(muca_ctx *c, const double *deltas, uint32_t ndelta, double *absM,
                  double *chi, double *binder)
{
    if (!c || !deltas || ndelta == 0) return fail(MUCA_E_ARG, "bad arguments");
    if (!c->joint && !c->moms)
        return fail(MUCA_E_STATE, "context has no joint histogram or per-D moments");
    if (!c->have_weights) return fail(MUCA_E_STATE, "muca_set_log_weight was not called");
    HIP_TRY(hipSetDevice(c->p.device));
    const uint32_t B = c->p.joint_blocks;
    const uint64_t nb = c->N + 1;
    if (!c->mom) {
        int rc = dev_alloc(&c->mom, (uint64_t)B * 4 * nb, "moments");
        if (rc) return rc;
    }
    if (!c->rw_flag) {
        int rc = dev_alloc(&c->rw_flag, 1, "range flag");
        if (rc) return rc;
    }
    if (ndelta > c->rw_cap) {
        if (c->deltas) hipFree(c->deltas);
        if (c->rw_out) hipFree(c->rw_out);
        c->deltas = nullptr;
        c->rw_out = nullptr;
        c->rw_cap = 0;
        int rc = dev_alloc(&c->deltas, ndelta, "deltas");
        if (!rc) rc = dev_alloc(&c->rw_out, (uint64_t)ndelta * B * 3, "reweight output");
        if (rc) return rc;
        c->rw_cap = ndelta;
    }
    HIP_TRY(hipMemcpyAsync(c->deltas, deltas, ndelta * 8, hipMemcpyHostToDevice, c->st));
    if (c->joint)
        HIP_TRY(muca::launch_joint_moments(c->st, c->joint, c->N, c->joint_stride, B, c->mom));
    else {
        HIP_TRY(hipMemsetAsync(c->rw_flag, 0, 4, c->st));
        HIP_TRY(muca::launch_moments_fp64(c->st, c->moms, c->N, B, c->mom, c->rw_flag));
    }
    HIP_TRY(muca::launch_reweight(c->st, c->mom, c->lnw64, c->N, B, c->p.beta, c->deltas,
                                  ndelta, c->rw_out));
    std::vector<double> o((size_t)ndelta * B * 3);
    uint32_t range = 0;
    HIP_TRY(hipMemcpyAsync(o.data(), c->rw_out, o.size() * 8, hipMemcpyDeviceToHost, c->st));
    if (!c->joint)
        HIP_TRY(hipMemcpyAsync(&range, c->rw_flag, 4, hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    if (range) return fail(MUCA_E_RANGE, "a D bin holds 2^32 or more entries");
    const size_t plane = (size_t)ndelta * B;
    for (size_t i = 0; i < plane; i++) {
        if (absM) absM[i] = o[i];
        if (chi) chi[i] = o[plane + i];
        if (binder) binder[i] = o[2 * plane + i];
    }
    return MUCA_OK;
}

// ---- host-side pieces of the weight iteration ------------------------------

int muca_hist_range(const uint64_t *hist, uint32_t N, int32_t *start, int32_t *end)
{
    if (!hist || !start || !end) return fail(MUCA_E_ARG, "NULL argument");
    int32_t s = (int32_t)N + 1, e = 0;
    for (uint32_t i = 0; i <= N; i++)
        if (hist[i] > 1) {
            s = (int32_t)i;
            break;
        }
    for (uint32_t i = N + 1; i-- > 0;)
        if (hist[i] > 1) {
            e = (int32_t)i;
            break;
        }
    if (s > e) s = e = 0;
    *start = s;
    *end = e;
    return MUCA_OK;
}

double muca_kull_leibler(const uint64_t *hist, uint32_t N)
{
    double bins = 0, events = 0, kld = 0;
    for (uint32_t i = 0; i <= N; i++)
        if (hist[i] > 0) {
            events = events + (double)hist[i];
            bins = bins + 1.0;
        }
    for (uint32_t i = 0; i <= N; i++)
        if (hist[i] > 0) {
            const double P = (double)hist[i] / events;
            const double Q = 1.0 / bins;
            kld = kld + P * std::log(P / Q);
        }
    return kld;
}

int muca_update_trivial(const uint64_t *hist, uint32_t N, double *lnw)
{
    if (!hist || !lnw) return fail(MUCA_E_ARG, "NULL argument");
    for (uint32_t i = 0; i <= N; i++)
        if (hist[i] > 1) lnw[i] = lnw[i] - std::log((double)hist[i]);
    return MUCA_OK;
}

int muca_update_recursive(const uint64_t *hist, uint32_t N, int32_t start, int32_t end,
                          double *p_acc, double *lnw)
{
    if (!hist || !lnw || !p_acc) return fail(MUCA_E_ARG, "NULL argument");
    if (start < 0 || end > (int32_t)N || start > end) return fail(MUCA_E_ARG, "bad range");
    std::vector<double> bf(lnw, lnw + N + 1);
    for (int32_t i = start; i < end; i++) {
        const uint64_t a = hist[i], b = hist[i + 1];
        if ((a != 0) & (b != 0)) {
            const double p = (double)(a * b) / (double)(a + b);
            const double k = p / (p + p_acc[i]);
            const double dS = -bf[i] + bf[i + 1] + k * std::log((double)a / (double)b);
            lnw[i + 1] = lnw[i] + dS;
            p_acc[i] = p_acc[i] + p;
        }
    }
    return MUCA_OK;
}

int muca_normalise(uint32_t N, double *lnw)
{
    if (!lnw) return fail(MUCA_E_ARG, "NULL argument");
    const double t = lnw[0];
    for (uint32_t i = 0; i <= N; i++) lnw[i] = lnw[i] - t;
    return MUCA_OK;
}

int muca_jackknife(const double *x, uint32_t J, double *avg_out, double *err_out)
{
    if (!x || !avg_out || !err_out || J == 0) return fail(MUCA_E_ARG, "bad arguments");
    double avg = 0;
    for (uint32_t i = 0; i < J; i++) avg += x[i];
    std::vector<double> mi(J);
    for (uint32_t i = 0; i < J; i++) mi[i] = (avg - x[i]) / (double)(J - 1);
    avg = avg / (double)J;
    double dev = 0;
    for (uint32_t i = 0; i < J; i++) dev += (avg - mi[i]) * (avg - mi[i]);
    dev = (J - 1) * dev / (double)J;
    *avg_out = avg;
    *err_out = std::sqrt(dev);
    return MUCA_OK;
}

// ---- self-checks -----------------------------------------------------------

float muca_acclog_host(uint32_t r)
{
    return acclog_u01(r);
}

int muca_acclog_host_range(uint32_t k0, uint32_t n, float *out)
{
    if (!out) return fail(MUCA_E_ARG, "NULL argument");
    for (uint32_t t = 0; t < n; t++) out[t] = acclog_u01((k0 + t) << 9);
    return MUCA_OK;
}

int muca_acclog_device(int32_t device, uint32_t k0, uint32_t n, float *out)
{
    if (!out || n == 0) return fail(MUCA_E_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(device));
    float *d = nullptr;
    int rc = dev_alloc(&d, n, "acclog");
    if (rc) return rc;
    hipError_t e = muca::launch_acclog(nullptr, k0, n, d);
    if (e == hipSuccess) e = hipMemcpy(out, d, (size_t)n * 4, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(MUCA_E_HIP, std::string("acclog: ") + hipGetErrorString(e));
    return MUCA_OK;
}

int muca_acclog_fast_device(int32_t device, uint32_t k0, uint32_t n, float *out)
{
    if (!out || n == 0) return fail(MUCA_E_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(device));
    float *d = nullptr;
    int rc = dev_alloc(&d, n, "acclog_fast");
    if (rc) return rc;
    hipError_t e = muca::launch_acclog_fast(nullptr, k0, n, d);
    if (e == hipSuccess) e = hipMemcpy(out, d, (size_t)n * 4, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess)
        return fail(MUCA_E_HIP, std::string("acclog_fast: ") + hipGetErrorString(e));
    return MUCA_OK;
}

}  // extern "C"
