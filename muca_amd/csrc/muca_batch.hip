// muca_batch.hip -- the batched engine: many independent multicanonical
// weight problems (instances of different L, beta, seed, walker count) advanced
// together, one weight iteration per two launches:
//   1. a persistent update kernel (team or lane tiles, muca_team.hip /
//      muca_lane.hip) that walks every unconverged instance's walkers through
//      therm = N and NUPDATES attempts into per-instance uint64 histograms;
//   2. k_batch_update, one workgroup per instance, which does on the device
//      what the reference's host loop does between launches
//      (muca_delta_multigpu.cu:431-476): hist_range, the width rule,
//      kull_leibler, the convergence test, the trivial update or
//      weight_recursion in FP64, lnW -= lnW[0], the float copy, and the
//      NUPDATES schedule for the next iteration.
// Both launches read all per-iteration state from device memory, so one
// iteration is captured once into a hipGraph and replayed; nothing returns to
// the host between iterations.  The device update uses the device log, which
// may differ from the host libm log in the last bit; the update kernels'
// chains are exact given the weights they are handed.
#include "../../include/muca.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.h"

namespace muca {

namespace {

constexpr int kUpd = 1024;    // threads of the update workgroup
constexpr int kChunk = 2048; // recursion steps staged in LDS at a time

template <typename T, typename Op>
__device__ T block_reduce(T v, T ident, Op op, T *sh)
{
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    __syncthreads();
    if (ln == 0) sh[wv] = v;
    __syncthreads();
    if (threadIdx.x < 64) {
        v = threadIdx.x < (blockDim.x >> 6) ? sh[threadIdx.x] : ident;
        for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
        if (threadIdx.x == 0) sh[0] = v;
    }
    __syncthreads();
    const T r = sh[0];
    __syncthreads();
    return r;
}

struct OpMin {
    __device__ int64_t operator()(int64_t a, int64_t b) const { return a < b ? a : b; }
};
struct OpMax {
    __device__ int64_t operator()(int64_t a, int64_t b) const { return a > b ? a : b; }
};
struct OpAddU {
    __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a + b; }
};
struct OpAddD {
    __device__ double operator()(double a, double b) const { return a + b; }
};

}  // namespace

__global__ __launch_bounds__(kUpd) void k_batch_update(BatchArgs b)
{
    __shared__ int64_t sh_i[16];
    __shared__ uint64_t sh_u[16];
    __shared__ double sh_d[16];
    __shared__ int s_conv, s_rec;
    __shared__ double s_ds[kChunk], s_bn[kChunk], s_x;
    if (blockIdx.x == 0 && threadIdx.x < kBatchQueues) b.queue[threadIdx.x] = 0;   // next iteration
    BatchInst *I = b.inst + blockIdx.x;
    const uint32_t N = I->N;
    uint64_t *h = b.hist + I->w_off;
    uint64_t *hl = b.hist_last + I->w_off;
    double *lnw = b.lnw64 + I->w_off;
    float *l32 = b.lnw32 + I->w_off;
    const int was_converged = I->converged;
    if (!was_converged) {
        // hist_range (muca_delta_multigpu.cu:573-596) and the kull_leibler sums (:598-606)
        int64_t st = (int64_t)N + 1, en = -1;
        uint64_t ev = 0, nb = 0;
        for (uint32_t i = threadIdx.x; i <= N; i += blockDim.x) {
            const uint64_t c = h[i];
            if (c > 1) {
                st = st < (int64_t)i ? st : (int64_t)i;
                en = (int64_t)i;
            }
            if (c > 0) {
                ev += c;
                nb += 1;
            }
        }
        st = block_reduce(st, (int64_t)N + 1, OpMin(), sh_i);
        en = block_reduce(en, (int64_t)-1, OpMax(), sh_i);
        ev = block_reduce(ev, (uint64_t)0, OpAddU(), sh_u);
        nb = block_reduce(nb, (uint64_t)0, OpAddU(), sh_u);
        int32_t start = (int32_t)st, end = en < 0 ? 0 : (int32_t)en;
        if (start > end) start = end = 0;
        // kull_leibler (:609-615); the sums of integer counts are exact
        const double num_event = (double)ev, num_bin = (double)nb;
        double kl = 0;
        for (uint32_t i = threadIdx.x; i <= N; i += blockDim.x) {
            const uint64_t c = h[i];
            if (c > 0) {
                const double P = (double)c / num_event;
                const double Q = 1.0 / num_bin;
                kl = kl + P * log(P / Q);
            }
        }
        kl = block_reduce(kl, 0.0, OpAddD(), sh_d);
        if (threadIdx.x == 0) {
            const uint32_t wt = (uint32_t)(end - start);            // :432-433
            if (wt > I->width) I->width = wt;
            I->start = start;
            I->end = end;
            I->dkl = kl;
            I->attempts += (uint64_t)I->W * ((uint64_t)I->N + I->nup);
            if (I->width == N && !I->reach) {                       // :441-446
                I->reach = 1;
                I->reach_iter = I->iter;
            }
            const int conv = (I->width == N) && (kl < 1e-4);        // :447
            if (conv) {   // the reference leaves the loop before updating (:462)
                I->converged = 1;
                I->done_iter = I->iter;
                I->iter += 1;
            }
            s_conv = conv;
            s_rec = I->recursion;
        }
        __syncthreads();
        if (!s_conv) {
            if (!s_rec) {
                // trivial update (:467)
                for (uint32_t i = threadIdx.x; i <= N; i += blockDim.x) {
                    const uint64_t c = h[i];
                    if (c > 1) lnw[i] = lnw[i] - log((double)c);
                }
            } else {
                // weight_recursion (:848-858).  Each step's increment dS_i and
                // p_acc update read only the pre-update weights bf and the
                // histogram, so they are formed in parallel (same expressions,
                // same order of operations); only lnW[i+1] = lnW[i] + dS_i is a
                // chain.  Chunks of kChunk steps are staged in LDS and one lane
                // walks each chunk in order.  A NaN marks "no step" (dS is finite
                // whenever both counts are nonzero and ln W is finite); a step
                // that is not taken leaves lnW[i+1] at its old value bf[i+1].
                double *bf = b.bf + I->w_off, *pacc = b.p_acc + I->w_off;
                for (uint32_t i = threadIdx.x; i <= N; i += blockDim.x) bf[i] = lnw[i];
                if (threadIdx.x == 0) s_x = lnw[start];
                __syncthreads();
                for (int32_t c0 = start; c0 < end; c0 += kChunk) {
                    const int32_t n = min(kChunk, end - c0);
                    for (int32_t j = (int32_t)threadIdx.x; j < n; j += blockDim.x) {
                        const int32_t i = c0 + j;
                        const uint64_t a0 = h[i], a1 = h[i + 1];
                        double dS = __builtin_nan("");
                        if ((a0 != 0) & (a1 != 0)) {
                            const double p = (double)(a0 * a1) / (double)(a0 + a1);
                            const double k = p / (p + pacc[i]);
                            dS = -bf[i] + bf[i + 1] + k * log((double)a0 / (double)a1);
                            pacc[i] = pacc[i] + p;
                        }
                        s_ds[j] = dS;
                        s_bn[j] = bf[i + 1];
                    }
                    __syncthreads();
                    if (threadIdx.x == 0) {
                        // blocks of 16 steps without a bound test per step,
                        // then the remainder one by one
                        double x = s_x;
                        double *out = lnw + c0 + 1;
                        int32_t j = 0;
                        for (; j + 16 <= n; j += 16) {
                            double dv[16], bv[16];
#pragma unroll
                            for (int u = 0; u < 16; u++) {
                                dv[u] = s_ds[j + u];
                                bv[u] = s_bn[j + u];
                            }
#pragma unroll
                            for (int u = 0; u < 16; u++) {
                                x = (dv[u] == dv[u]) ? x + dv[u] : bv[u];
                                out[j + u] = x;
                            }
                        }
                        for (; j < n; j++) {
                            const double dv = s_ds[j];
                            x = (dv == dv) ? x + dv : s_bn[j];
                            out[j] = x;
                        }
                        s_x = x;
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
            const double t0 = lnw[0];                               // :469-475
            __syncthreads();
            for (uint32_t i = threadIdx.x; i <= N; i += blockDim.x) {
                const double v = lnw[i] - t0;
                lnw[i] = v;
                l32[i] = (float)v;
            }
            if (threadIdx.x == 0) {
                if (I->width == N) {                                 // :379-382, next iteration
                    // 64-bit NUPDATES: with growth 1.01 equal to the
                    // reference's unsigned arithmetic below 2^32; capped
                    // (never 0) at the instance's maximum
                    const uint64_t nu = I->nup + 1;
                    const double grown = (double)nu * I->nup_growth;
                    if (grown >= (double)I->nup_max) {
                        I->nup = I->nup_max;
                        I->nup_capped = 1;
                    } else {
                        I->nup = (uint64_t)grown;
                    }
                }
                I->iter += 1;
            }
        }
    }
    __syncthreads();
    if (!was_converged) {
        for (uint32_t i = threadIdx.x; i <= N; i += blockDim.x) {
            hl[i] = h[i];
            h[i] = 0;
        }
    }
}

hipError_t launch_batch_update(hipStream_t st, const BatchArgs &b)
{
    hipLaunchKernelGGL(k_batch_update, dim3(b.ninst), dim3(kUpd), 0, st, b);
    return hipGetLastError();
}

}  // namespace muca

// ---------------------------------------------------------------------------
// host object and C ABI
// ---------------------------------------------------------------------------
namespace {

thread_local std::string g_berr;

int bfail(int code, const std::string &m)
{
    g_berr = m;
    return code;
}

#define BHIP(expr)                                                                      \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return bfail(MUCA_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct muca_batch {
    int device = 0, fp64 = 0, kind = 0;   // kind: MUCA_KERNEL_TEAM | MUCA_KERNEL_LANE
    uint32_t block = 256, grid = 0, cus = 0;
    size_t lds_bytes = 0;
    int lat_modes = 3;   // team tiles: lattices in LDS (bit 0) / in HBM (bit 1)
    int nrs = 0;         // team tiles: rounds per chunk common to all instances, else 0
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<muca::BatchInst> inst;
    std::vector<uint64_t> init_nup;
    std::vector<muca::BatchTile> tiles;
    muca::BatchArgs a{};
    muca::BatchTile *d_tiles = nullptr;
    hipGraph_t graph = nullptr, graph8 = nullptr;   // one iteration, kGraphIters iterations
    hipGraphExec_t exec = nullptr, exec8 = nullptr;
    uint64_t lat_bytes = 0, walkers = 0, bins = 0;
    bool team_ok = false, lane_ok = false;
    size_t lds_device = 0;
    int stratified = 0;   // start of the last init
    bool inited = false;  // muca_batch_init* ran (ln W holds values, not garbage)
};

namespace {

// iterations replayed by one launch of the larger graph: consecutive
// iterations in one graph start back to back, while separate graph launches
// leave a gap of several microseconds between them
constexpr uint32_t kGraphIters = 8;

void drop_graphs(muca_batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->exec8) hipGraphExecDestroy(b->exec8);
    if (b->graph8) hipGraphDestroy(b->graph8);
    b->exec = b->exec8 = nullptr;
    b->graph = b->graph8 = nullptr;
}

hipError_t launch_iteration(muca_batch *b);

// `iters` iterations captured from the batch's stream into one graph
hipError_t capture_iterations(muca_batch *b, uint32_t iters, hipGraph_t *g, hipGraphExec_t *x)
{
    hipError_t e = hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    for (uint32_t k = 0; k < iters && e == hipSuccess; k++) e = launch_iteration(b);
    hipGraph_t gr = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->st, &gr);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(x, gr, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (gr) hipGraphDestroy(gr);
        return e;
    }
    *g = gr;
    return hipSuccess;
}

void batch_free(muca_batch *b)
{
    if (!b) return;
    hipSetDevice(b->device);
    drop_graphs(b);
    void *bufs[] = {b->a.inst, b->d_tiles, b->a.queue, b->a.lat, b->a.E, b->a.D, b->a.M,
                    b->a.lnw32, b->a.lnw64, b->a.hist, b->a.hist_last, b->a.bf, b->a.p_acc};
    for (void *p : bufs)
        if (p) hipFree(p);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->st) hipStreamDestroy(b->st);
    delete b;
}

template <typename T>
hipError_t balloc(T **p, uint64_t n)
{
    return hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
}

hipError_t launch_iteration(muca_batch *b)
{
    hipError_t e = b->kind == MUCA_KERNEL_TEAM
                       ? muca::launch_batch_team(b->st, b->fp64, b->a, b->grid, b->lds_bytes,
                                                 b->lat_modes, b->nrs)
                       : muca::launch_batch_lane(b->st, b->fp64, b->a, b->grid, b->block,
                                                 b->lds_bytes);
    if (e != hipSuccess) return e;
    return muca::launch_batch_update(b->st, b->a);
}

int batch_init(muca_batch *b, uint32_t mm, int stratified);

// Tile list, grid and LDS size for one tile kind; drops any captured graph.
hipError_t configure(muca_batch *b, int kind, uint32_t block)
{
    drop_graphs(b);
    b->kind = kind;
    size_t need = 0;   // LDS of the largest instance's tile
    for (const auto &I : b->inst)
        need = std::max(need, kind == MUCA_KERNEL_TEAM
                                  ? I.lay.bytes
                                  : (size_t)I.lane_lds_w + (I.lane_hist ? (size_t)(I.N + 1) * 4 : 0));
    const uint32_t per_cu_lds = (uint32_t)std::max<size_t>(1, b->lds_device / (need + 256));
    if (kind == MUCA_KERNEL_LANE && block == 0) {
        // one lane per walker hides latency with many wavefronts: when LDS
        // admits few workgroups per compute unit, make them larger (L = 128:
        // ln W and the histogram take 128 KB, one 1024-lane workgroup per CU)
        block = 256;
        while (block < 1024 && per_cu_lds * block < 2048) block *= 2;
    }
    b->block = kind == MUCA_KERNEL_TEAM ? 256 : block;
    const uint32_t per = kind == MUCA_KERNEL_TEAM ? 16 : b->block;
    const uint32_t n = (uint32_t)b->inst.size();
    std::vector<uint32_t> order(n);   // largest lattices first
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return b->inst[x].N > b->inst[y].N; });
    // groups of tiles that share lattice lines (eight team tiles of 16
    // walkers make 128 bytes; a lane tile fills whole lines by itself) are
    // dealt to the XCD queues in turn, each queue keeping the largest-first
    // order
    const uint32_t group = kind == MUCA_KERNEL_TEAM ? 8 : 1;
    std::vector<muca::BatchTile> q[muca::kBatchQueues];
    uint32_t next = 0;
    for (uint32_t i : order)
        for (uint32_t w = 0; w < b->inst[i].W; w += per) {
            q[next].push_back({i, w});
            if ((w / per) % group == group - 1 || w + per >= b->inst[i].W)
                next = (next + 1) % muca::kBatchQueues;
        }
    b->tiles.clear();
    for (uint32_t x = 0; x < muca::kBatchQueues; x++) {
        b->a.qstart[x] = (uint32_t)b->tiles.size();
        b->tiles.insert(b->tiles.end(), q[x].begin(), q[x].end());
    }
    b->a.qstart[muca::kBatchQueues] = (uint32_t)b->tiles.size();
    b->lds_bytes = need;
    b->lat_modes = 0;
    b->nrs = -1;
    for (const auto &I : b->inst) {
        b->lat_modes |= I.lay.lat_lds ? 1 : 2;
        const int nr = (int)std::min<uint32_t>((I.lay.row + 15) / 16, 4);
        b->nrs = b->nrs == -1 || b->nrs == nr ? nr : 0;
    }
    if (b->nrs < 0) b->nrs = 0;
    const uint32_t per_cu_thr = 2048 / b->block;
    b->grid = b->cus * std::min(per_cu_lds, per_cu_thr);
    b->grid = std::min<uint32_t>(b->grid, (uint32_t)b->tiles.size());
    b->a.ntiles = (uint32_t)b->tiles.size();
    hipError_t e = hipMemcpy(b->d_tiles, b->tiles.data(),
                             b->tiles.size() * sizeof(muca::BatchTile), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    return kind == MUCA_KERNEL_TEAM
               ? muca::prepare_batch_team(b->fp64, b->lds_bytes, b->lat_modes, b->nrs)
               : muca::prepare_batch_lane(b->fp64, b->lds_bytes);
}

}  // namespace

extern "C" {

const char *muca_batch_last_error(void)
{
    return g_berr.c_str();
}

int muca_batch_create(int32_t device, const muca_instance *ins, uint32_t n, int32_t precision,
                      int32_t tiles, muca_batch **out)
{
    if (!ins || !out || n == 0) return bfail(MUCA_E_ARG, "bad arguments");
    *out = nullptr;
    if (precision != MUCA_REF32 && precision != MUCA_FP64)
        return bfail(MUCA_E_ARG, "precision must be MUCA_REF32 or MUCA_FP64");
    // instance checks come before any device call
    for (uint32_t i = 0; i < n; i++) {
        const muca_instance &q = ins[i];
        if (q.L < 4 || q.L % 2 != 0 || q.L > 46340 || q.walkers == 0 || !(q.beta > 0))
            return bfail(MUCA_E_ARG, "instance " + std::to_string(i) +
                                         ": need even L >= 4, walkers > 0, beta > 0");
        if (q.init_nup > MUCA_MAX_NUPDATES || q.nup_max > MUCA_MAX_NUPDATES)
            return bfail(MUCA_E_RANGE, "instance " + std::to_string(i) +
                                           ": init_nup or nup_max exceeds MUCA_MAX_NUPDATES (2^48)");
        if (!(q.nup_growth == 0 || (q.nup_growth >= 1 && q.nup_growth <= 1e6)))
            return bfail(MUCA_E_ARG, "instance " + std::to_string(i) +
                                         ": nup_growth must be 0 (1.01) or in [1, 1e6]");
    }
    int ndev = 0;
    BHIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return bfail(MUCA_E_ARG, "device not present");
    BHIP(hipSetDevice(device));
    int cu = 0, lds = 0;
    BHIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device));
    BHIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    const size_t lds_budget = (size_t)lds - 256;   // static LDS of the persistent kernels
    muca_batch *b = new muca_batch();
    b->device = device;
    b->fp64 = precision == MUCA_FP64;
    b->cus = (uint32_t)cu;
    bool team_ok = true, lane_ok = true;
    uint64_t lat = 0, st = 0, wo = 0;
    for (uint32_t i = 0; i < n; i++) {
        const muca_instance &q = ins[i];
        muca::BatchInst I{};
        I.L = q.L;
        I.N = q.L * q.L;
        I.W = q.walkers;
        I.seed = q.seed;
        I.beta = q.beta;
        I.recursion = q.recursion ? 1 : 0;
        I.nup_growth = q.nup_growth == 0 ? 1.01 : q.nup_growth;
        I.nup_max = q.nup_max == 0 ? MUCA_MAX_NUPDATES : q.nup_max;
        I.lat_off = lat;
        I.st_off = st;
        I.w_off = wo;
        lat += (uint64_t)I.N * I.W;
        st += I.W;
        wo += I.N + 1;
        muca::McsArgs a{};
        a.N = I.N;
        a.L = I.L;
        a.W = I.W;
        a.nup = 1;
        if (!(I.L % 8 == 0 && I.L <= 128 && muca::team_layout(a, b->fp64, lds_budget, I.lay) &&
              !I.lay.lnw_global))
            team_ok = false;
        const size_t tsz = b->fp64 ? 8 : 4;
        I.lane_lds_w = (uint32_t)(((I.N + 1) * tsz + 15) & ~(size_t)15);
        I.lane_hist = I.lane_lds_w + (size_t)(I.N + 1) * 4 <= lds_budget;
        if (I.lane_lds_w > lds_budget) lane_ok = false;
        b->inst.push_back(I);
        b->init_nup.push_back(q.init_nup ? q.init_nup : 1);
    }
    b->lat_bytes = lat;
    b->walkers = st;
    b->bins = wo;
    b->team_ok = team_ok;
    b->lane_ok = lane_ok;
    b->lds_device = (size_t)lds;
    if (tiles == MUCA_KERNEL_TEAM && !team_ok) {
        batch_free(b);
        return bfail(MUCA_E_ARG, "team tiles need every L % 8 == 0 and L <= 128");
    }
    if (tiles == MUCA_KERNEL_LANE && !lane_ok) {
        batch_free(b);
        return bfail(MUCA_E_ARG, "lane tiles need ln W of every instance in LDS");
    }
    int kind = tiles;
    if (tiles != MUCA_KERNEL_TEAM && tiles != MUCA_KERNEL_LANE)
        // same rule as the single-instance contexts: team while walkers are few
        kind = (team_ok && (st <= 128ull * (uint64_t)cu || !lane_ok)) ? MUCA_KERNEL_TEAM
                                                                       : MUCA_KERNEL_LANE;
    if (kind == MUCA_KERNEL_LANE && !lane_ok) {
        batch_free(b);
        return bfail(MUCA_E_ARG, "no tile kind fits every instance");
    }
    uint64_t max_tiles = 0;   // team tiles (16 walkers) are the most numerous
    for (const auto &I : b->inst) max_tiles += (I.W + 15) / 16;
    hipError_t e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    muca::BatchArgs &A = b->a;
    A.ninst = n;
    if (e == hipSuccess) e = balloc(&A.inst, n);
    if (e == hipSuccess) e = balloc(&b->d_tiles, max_tiles);
    if (e == hipSuccess) e = balloc(&A.queue, muca::kBatchQueues);
    if (e == hipSuccess) e = balloc(&A.lat, lat);
    if (e == hipSuccess) e = balloc(&A.E, st);
    if (e == hipSuccess) e = balloc(&A.D, st);
    if (e == hipSuccess) e = balloc(&A.M, st);
    if (e == hipSuccess) e = balloc(&A.lnw32, wo);
    if (e == hipSuccess) e = balloc(&A.lnw64, wo);
    if (e == hipSuccess) e = balloc(&A.hist, wo);
    if (e == hipSuccess) e = balloc(&A.hist_last, wo);
    if (e == hipSuccess) e = balloc(&A.bf, wo);
    if (e == hipSuccess) e = balloc(&A.p_acc, wo);
    A.tiles = b->d_tiles;
    if (e == hipSuccess) e = configure(b, kind, 0);
    if (e != hipSuccess) {
        batch_free(b);
        return bfail(MUCA_E_NOMEM, std::string("batch setup: ") + hipGetErrorString(e));
    }
    *out = b;
    return MUCA_OK;
}

int muca_batch_destroy(muca_batch *b)
{
    batch_free(b);
    return MUCA_OK;
}

int muca_batch_query(const muca_batch *b, int32_t *kind, uint32_t *block, uint32_t *tiles,
                     uint32_t *grid, uint64_t *walkers)
{
    if (!b) return bfail(MUCA_E_ARG, "NULL batch");
    if (kind) *kind = b->kind;
    if (block) *block = b->block;
    if (tiles) *tiles = (uint32_t)b->tiles.size();
    if (grid) *grid = b->grid;
    if (walkers) *walkers = b->walkers;
    return MUCA_OK;
}

int muca_batch_init(muca_batch *b, uint32_t mm)
{
    return batch_init(b, mm, 0);
}

int muca_batch_init_stratified(muca_batch *b, uint32_t mm)
{
    return batch_init(b, mm, 1);
}

}  // extern "C"

namespace {
int batch_init(muca_batch *b, uint32_t mm, int stratified)
{
    if (!b) return bfail(MUCA_E_ARG, "NULL batch");
    b->stratified = stratified;
    BHIP(hipSetDevice(b->device));
    muca::BatchArgs &A = b->a;
    uint32_t maxN = 0;
    for (const auto &I : b->inst) maxN = std::max(maxN, I.N);
    struct DevBuf {
        uint32_t *p = nullptr;
        ~DevBuf()
        {
            if (p) hipFree(p);
        }
    } naver_d;
    BHIP(hipMalloc((void **)&naver_d.p, 16ull * maxN));
    std::vector<uint32_t> naver;
    uint32_t cur_L = 0;
    for (size_t i = 0; i < b->inst.size(); i++) {
        muca::BatchInst &I = b->inst[i];
        BHIP(muca::launch_init_lattice(b->st, A.lat + I.lat_off, I.W, I.N, MUCA_KEYING_WEIGHT, 7,
                                       mm, 0, I.seed, 0, stratified ? I.W : 0));
        if (I.L != cur_L) {   // neighbour table of this L (muca_delta_multigpu.cu:157-172)
            BHIP(hipStreamSynchronize(b->st));   // the previous table may still be read
            naver.assign(4ull * I.N, 0);
            for (uint32_t s = 0; s < I.N; s++) {
                const uint32_t x = s % I.L, y = s / I.L;
                naver[4 * s + 0] = y * I.L + (x + 1) % I.L;
                naver[4 * s + 1] = y * I.L + (x + I.L - 1) % I.L;
                naver[4 * s + 2] = ((y + 1) % I.L) * I.L + x;
                naver[4 * s + 3] = ((y + I.L - 1) % I.L) * I.L + x;
            }
            BHIP(hipMemcpy(naver_d.p, naver.data(), naver.size() * 4, hipMemcpyHostToDevice));
            cur_L = I.L;
        }
        BHIP(muca::launch_state(b->st, A.lat + I.lat_off, I.W, I.N, naver_d.p, A.E + I.st_off,
                                A.D + I.st_off, A.M + I.st_off));
        I.iter = 0;
        I.nup = b->init_nup[i];
        I.width = 10;   // :314
        I.reach = I.converged = I.reach_iter = I.done_iter = I.nup_capped = 0;
        I.start = I.end = 0;
        I.dkl = 0;
        I.attempts = 0;
    }
    BHIP(hipStreamSynchronize(b->st));
    BHIP(hipMemcpy(A.inst, b->inst.data(), b->inst.size() * sizeof(muca::BatchInst),
                   hipMemcpyHostToDevice));
    BHIP(hipMemset(A.queue, 0, 4 * muca::kBatchQueues));
    BHIP(hipMemset(A.lnw32, 0, b->bins * 4));
    BHIP(hipMemset(A.lnw64, 0, b->bins * 8));
    BHIP(hipMemset(A.hist, 0, b->bins * 8));
    BHIP(hipMemset(A.hist_last, 0, b->bins * 8));
    BHIP(hipMemset(A.p_acc, 0, b->bins * 8));
    b->inited = true;
    return MUCA_OK;
}
}  // namespace

extern "C" {

int muca_batch_set_log_weight(muca_batch *b, uint32_t i, const double *lnw, uint32_t nbins)
{
    if (!b || !lnw || i >= b->inst.size()) return bfail(MUCA_E_ARG, "bad arguments");
    const muca::BatchInst &I = b->inst[i];
    if (nbins != I.N + 1) return bfail(MUCA_E_ARG, "nbins must be N+1");
    std::vector<float> f(nbins);
    for (uint32_t k = 0; k < nbins; k++) {
        if (!std::isfinite(lnw[k])) return bfail(MUCA_E_ARG, "ln W must be finite");
        f[k] = (float)lnw[k];
    }
    BHIP(hipSetDevice(b->device));
    BHIP(hipStreamSynchronize(b->st));
    BHIP(hipMemcpy(b->a.lnw64 + I.w_off, lnw, nbins * 8, hipMemcpyHostToDevice));
    BHIP(hipMemcpy(b->a.lnw32 + I.w_off, f.data(), nbins * 4, hipMemcpyHostToDevice));
    return MUCA_OK;
}

int muca_batch_get_log_weight(muca_batch *b, uint32_t i, double *lnw, uint32_t nbins)
{
    if (!b || !lnw || i >= b->inst.size()) return bfail(MUCA_E_ARG, "bad arguments");
    const muca::BatchInst &I = b->inst[i];
    if (nbins != I.N + 1) return bfail(MUCA_E_ARG, "nbins must be N+1");
    BHIP(hipSetDevice(b->device));
    BHIP(hipStreamSynchronize(b->st));
    BHIP(hipMemcpy(lnw, b->a.lnw64 + I.w_off, nbins * 8, hipMemcpyDeviceToHost));
    return MUCA_OK;
}

int muca_batch_get_hist(muca_batch *b, uint32_t i, uint64_t *hist, uint32_t nbins)
{
    if (!b || !hist || i >= b->inst.size()) return bfail(MUCA_E_ARG, "bad arguments");
    const muca::BatchInst &I = b->inst[i];
    if (nbins != I.N + 1) return bfail(MUCA_E_ARG, "nbins must be N+1");
    BHIP(hipSetDevice(b->device));
    BHIP(hipStreamSynchronize(b->st));
    BHIP(hipMemcpy(hist, b->a.hist_last + I.w_off, nbins * 8, hipMemcpyDeviceToHost));
    return MUCA_OK;
}

int muca_batch_get_lattice(muca_batch *b, uint32_t i, int8_t *site_major, uint64_t bytes)
{
    if (!b || !site_major || i >= b->inst.size()) return bfail(MUCA_E_ARG, "bad arguments");
    const muca::BatchInst &I = b->inst[i];
    if (bytes != (uint64_t)I.N * I.W) return bfail(MUCA_E_ARG, "bytes must be N * walkers");
    BHIP(hipSetDevice(b->device));
    BHIP(hipStreamSynchronize(b->st));
    BHIP(hipMemcpy(site_major, b->a.lat + I.lat_off, bytes, hipMemcpyDeviceToHost));
    return MUCA_OK;
}

int muca_batch_get_state(muca_batch *b, uint32_t i, int32_t *E, int32_t *D, int32_t *M)
{
    if (!b || i >= b->inst.size()) return bfail(MUCA_E_ARG, "bad arguments");
    const muca::BatchInst &I = b->inst[i];
    BHIP(hipSetDevice(b->device));
    BHIP(hipStreamSynchronize(b->st));
    int32_t *dst[3] = {E, D, M};
    int32_t *src[3] = {b->a.E + I.st_off, b->a.D + I.st_off, b->a.M + I.st_off};
    for (int k = 0; k < 3; k++)
        if (dst[k]) BHIP(hipMemcpy(dst[k], src[k], (size_t)I.W * 4, hipMemcpyDeviceToHost));
    return MUCA_OK;
}

int muca_batch_iterate(muca_batch *b, uint32_t iterations, int32_t use_graph, float *gpu_ms)
{
    if (!b) return bfail(MUCA_E_ARG, "NULL batch");
    BHIP(hipSetDevice(b->device));
    // both graphs are made on the first call that uses them, so later calls
    // of any length replay without capturing
    if (use_graph && !b->exec) {
        hipError_t e = capture_iterations(b, 1, &b->graph, &b->exec);
        if (e == hipSuccess) e = capture_iterations(b, kGraphIters, &b->graph8, &b->exec8);
        if (e != hipSuccess) {
            drop_graphs(b);
            return bfail(MUCA_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        }
    }
    BHIP(hipEventRecord(b->ev0, b->st));
    uint32_t k = 0;
    if (use_graph) {
        for (; k + kGraphIters <= iterations; k += kGraphIters)
            BHIP(hipGraphLaunch(b->exec8, b->st));
        for (; k < iterations; k++) BHIP(hipGraphLaunch(b->exec, b->st));
    } else {
        for (; k < iterations; k++) BHIP(launch_iteration(b));
    }
    BHIP(hipEventRecord(b->ev1, b->st));
    BHIP(hipEventSynchronize(b->ev1));
    float ms = 0;
    BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (gpu_ms) *gpu_ms = ms;
    return MUCA_OK;
}

int muca_batch_iterate_profiled(muca_batch *b, uint32_t iterations, float *update_ms,
                                float *weight_ms)
{
    if (!b) return bfail(MUCA_E_ARG, "NULL batch");
    BHIP(hipSetDevice(b->device));
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    struct Guard {
        hipEvent_t *e;
        ~Guard()
        {
            for (int i = 0; i < 3; i++)
                if (e[i]) hipEventDestroy(e[i]);
        }
    } guard{ev};
    for (int i = 0; i < 3; i++) BHIP(hipEventCreate(&ev[i]));
    double tu = 0, tw = 0;
    for (uint32_t k = 0; k < iterations; k++) {
        BHIP(hipEventRecord(ev[0], b->st));
        BHIP(b->kind == MUCA_KERNEL_TEAM
                 ? muca::launch_batch_team(b->st, b->fp64, b->a, b->grid, b->lds_bytes,
                                           b->lat_modes, b->nrs)
                 : muca::launch_batch_lane(b->st, b->fp64, b->a, b->grid, b->block, b->lds_bytes));
        BHIP(hipEventRecord(ev[1], b->st));
        BHIP(muca::launch_batch_update(b->st, b->a));
        BHIP(hipEventRecord(ev[2], b->st));
        BHIP(hipEventSynchronize(ev[2]));
        float m1 = 0, m2 = 0;
        BHIP(hipEventElapsedTime(&m1, ev[0], ev[1]));
        BHIP(hipEventElapsedTime(&m2, ev[1], ev[2]));
        tu += m1;
        tw += m2;
    }
    if (update_ms) *update_ms = (float)tu;
    if (weight_ms) *weight_ms = (float)tw;
    return MUCA_OK;
}

int muca_batch_autotune(muca_batch *b, uint32_t mm, int32_t *kind, uint32_t *block, float *ms)
{
    if (!b) return bfail(MUCA_E_ARG, "NULL batch");
    BHIP(hipSetDevice(b->device));
    struct Cand {
        int kind;
        uint32_t block;
    };
    std::vector<Cand> cands;
    if (b->team_ok) cands.push_back({MUCA_KERNEL_TEAM, 256});
    if (b->lane_ok)
        for (uint32_t blk : {256u, 512u, 1024u}) cands.push_back({MUCA_KERNEL_LANE, blk});
    // ln W of an initialised batch (a surrogate seed, say) survives the
    // tuning: each candidate re-initialises the batch, which zeroes ln W
    BHIP(hipStreamSynchronize(b->st));
    const bool keep = b->inited;
    std::vector<double> keep64(keep ? b->bins : 0);
    std::vector<float> keep32(keep ? b->bins : 0);
    if (keep) {
        BHIP(hipMemcpy(keep64.data(), b->a.lnw64, b->bins * 8, hipMemcpyDeviceToHost));
        BHIP(hipMemcpy(keep32.data(), b->a.lnw32, b->bins * 4, hipMemcpyDeviceToHost));
    }
    float best = 0;
    Cand pick{b->kind, b->block};
    for (const Cand &c : cands) {
        BHIP(configure(b, c.kind, c.block));
        int rc = batch_init(b, mm, b->stratified);
        if (rc) return rc;
        float tu = 0;
        rc = muca_batch_iterate_profiled(b, 1, &tu, nullptr);
        if (rc) return rc;
        if (best == 0 || tu < best) {
            best = tu;
            pick = c;
        }
    }
    BHIP(configure(b, pick.kind, pick.block));
    const int rc = batch_init(b, mm, b->stratified);
    if (rc) return rc;
    if (keep) {
        BHIP(hipMemcpy(b->a.lnw64, keep64.data(), b->bins * 8, hipMemcpyHostToDevice));
        BHIP(hipMemcpy(b->a.lnw32, keep32.data(), b->bins * 4, hipMemcpyHostToDevice));
    }
    if (kind) *kind = pick.kind;
    if (block) *block = b->block;
    if (ms) *ms = best;
    return MUCA_OK;
}

int muca_batch_status(muca_batch *b, muca_instance_status *out, uint32_t n)
{
    if (!b || !out || n != b->inst.size()) return bfail(MUCA_E_ARG, "bad arguments");
    BHIP(hipSetDevice(b->device));
    BHIP(hipStreamSynchronize(b->st));
    std::vector<muca::BatchInst> cur(n);
    BHIP(hipMemcpy(cur.data(), b->a.inst, n * sizeof(muca::BatchInst), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; i++) {
        const muca::BatchInst &I = cur[i];
        out[i].iter = I.iter;
        out[i].nup = I.nup;
        out[i].nup_capped = I.nup_capped;
        out[i].width = I.width;
        out[i].converged = I.converged;
        out[i].reach_iter = I.reach ? I.reach_iter : 0xFFFFFFFFu;
        out[i].done_iter = I.converged ? I.done_iter : 0xFFFFFFFFu;
        out[i].start = I.start;
        out[i].end = I.end;
        out[i].dkl = I.dkl;
        out[i].attempts = I.attempts;
    }
    return MUCA_OK;
}

}  // extern "C"
