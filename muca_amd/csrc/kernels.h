// kernels.h -- launch interface between the C-ABI layer (muca_api.cpp) and
// the HIP kernels (muca_kernels.hip).  Internal to libmuca.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace muca {

// Arguments of one update launch.  Lattice layout: site-major,
// s[site * W + walker] (the reference's interleaving, muca_delta_multigpu.cu:151,800),
// with 64-bit indices.
struct McsArgs {
    int8_t *lat;
    int32_t *E, *D, *M;
    const float *lnw32;
    const double *lnw64;
    const uint32_t *order, *naver;
    uint64_t *hist;          // N+1 bins (zeroed by the caller)
    uint64_t *joint;         // nullable: joint_blocks * joint_stride bins
    uint64_t *moms;          // nullable: joint_blocks * kMomWords * (N+1) exact per-D moments
    uint64_t W;              // walkers in the lattice
    uint32_t N, L;
    uint32_t w_offset;       // logical id of walker 0 (for keys)
    uint32_t iter, seed;
    uint64_t therm, nup;
    double beta;
    uint64_t joint_stride;   // (N+1)(N+2)/2
    uint64_t block_walkers;  // walkers per joint block
    uint32_t joint_blocks;
    // test hook: the team and lane kernels start their u32 LDS histogram bins
    // at this value (0 in production) and flush the difference, so that a
    // test reaches the bin-wrap carry with few counts
    uint32_t hist_preload;
};

struct Launch {
    int rounds;      // 7 or 10
    int fp64;        // acceptance precision
    int keying;      // 0 weight, 1 produce
    int kernel;      // MUCA_KERNEL_*
    int block;       // threads per block
    size_t lds_max;  // bytes of LDS available per block
};

// strat_W = 0: the reference's uniform start; > 0: stratified over strat_W walkers
hipError_t launch_init_lattice(hipStream_t st, int8_t *lat, uint64_t W, uint32_t N,
                               int keying, int rounds, uint32_t mm, uint32_t gpu,
                               uint32_t seed, uint32_t w_offset, uint32_t strat_W = 0);
hipError_t launch_state(hipStream_t st, const int8_t *lat, uint64_t W, uint32_t N,
                        const uint32_t *naver, int32_t *E, int32_t *D, int32_t *M);
hipError_t launch_mcs(hipStream_t st, const Launch &cfg, const McsArgs &a);

// team kernel (muca_team.hip): LDS layout of one workgroup of 16 walkers
struct TeamLayout {
    uint32_t row;         // L/2 attempts per half-row
    uint32_t team_bytes;  // per-team Philox ring and ln W window
    uint32_t off_fde;     // (T)(-dE * beta) for dE = -8 .. 8
    uint32_t ring;        // Philox blocks per key held in a team's ring
    uint32_t off_win;     // inside a team's area: its ln W window (ln W in HBM)
    uint32_t off_mom;     // inside a team's area: its per-D moment slots (moments)
    uint32_t off_team, off_tab, off_w, off_hist, off_lat;
    uint32_t lat_stride;  // LDS bytes per walker lattice (>= N, bank-spread)
    size_t bytes;         // dynamic LDS per workgroup
    bool lds_hist, lat_lds;
    bool lnw_global;      // ln W too large for LDS: read from HBM (L2) per window
};
// false when the team kernel cannot run this launch (L % 8 != 0 or L > 256)
bool team_layout(const McsArgs &a, int fp64, size_t lds_max, TeamLayout &lay);
hipError_t launch_mcs_team(hipStream_t st, const Launch &cfg, const McsArgs &a,
                           const TeamLayout &lay);

// lane kernel (muca_lane.hip): even L >= 4 and ln W fits in LDS
bool lane_eligible(const McsArgs &a, int fp64, size_t lds_max);
hipError_t launch_mcs_lane(hipStream_t st, const Launch &cfg, const McsArgs &a);

// ---- batched engine (muca_batch.hip) ----
// One multicanonical weight problem of a batch.  The first block is set on
// the host; the second is the per-iteration state the device update kernel
// advances (the reference's host loop variables, muca_delta_multigpu.cu:314-447).
struct BatchInst {
    uint32_t L, N, W, seed;
    uint64_t lat_off;      // bytes into the batch lattice buffer (site-major per instance)
    uint64_t st_off;       // walkers into the E/D/M arrays
    uint64_t w_off;        // bins into ln W / histogram / scratch arrays
    double beta;
    int32_t recursion;     // 0 trivial update, 1 weight_recursion
    double nup_growth;     // schedule factor once width == N (1.01: the reference)
    uint64_t nup_max;      // schedule cap
    uint32_t lane_lds_w;   // lane tiles: LDS bytes of ln W
    uint32_t lane_hist;    // lane tiles: histogram in LDS
    TeamLayout lay;        // team tiles
    // advanced on the device
    uint32_t iter, width, reach, converged, reach_iter, done_iter;
    int32_t start, end;
    uint32_t nup_capped;   // the schedule reached nup_max
    uint64_t nup;          // NUPDATES of the current iteration
    double dkl;
    uint64_t attempts;     // flip attempts done so far
};

struct BatchTile {
    uint32_t inst, wbase;
};

// Tile queues of the persistent kernels: one per accelerator die (XCD) of
// an MI355X, each with its own L2; workgroup b is dispatched to XCD b mod 8.
// Walkers are the fastest index of a lattice, so the eight 16-walker team
// tiles that share a 128-byte lattice line go to one XCD's queue, and that
// line is fetched into one L2 instead of eight.
constexpr uint32_t kBatchQueues = 8;

struct BatchArgs {
    BatchInst *inst;
    const BatchTile *tiles;   // queue x holds tiles[qstart[x] .. qstart[x+1])
    uint32_t ntiles, ninst;
    uint32_t qstart[kBatchQueues + 1];
    uint32_t *queue;       // kBatchQueues tile counters of the persistent kernel
    int8_t *lat;
    int32_t *E, *D, *M;
    float *lnw32;
    double *lnw64;
    uint64_t *hist, *hist_last;
    double *bf, *p_acc;    // recursion scratch
};

// launch arguments of one instance for its current iteration (therm = N,
// the reference's therm_coeff = 1, muca_delta_multigpu.cu:315-316)
__device__ __forceinline__ McsArgs batch_mcs_args(const BatchArgs &b, const BatchInst &I)
{
    McsArgs a{};
    a.lat = b.lat + I.lat_off;
    a.E = b.E + I.st_off;
    a.D = b.D + I.st_off;
    a.M = b.M + I.st_off;
    a.lnw32 = b.lnw32 + I.w_off;
    a.lnw64 = b.lnw64 + I.w_off;
    a.order = nullptr;
    a.naver = nullptr;
    a.hist = b.hist + I.w_off;
    a.joint = nullptr;
    a.moms = nullptr;
    a.W = I.W;
    a.N = I.N;
    a.L = I.L;
    a.w_offset = 0;
    a.iter = I.iter;
    a.seed = I.seed;
    a.therm = I.N;
    a.nup = I.nup;
    a.beta = I.beta;
    a.joint_stride = 0;
    a.block_walkers = I.W;
    a.joint_blocks = 0;
    return a;
}

// The next tile for this workgroup: from its own XCD's queue, then from the
// others' once that one is empty; ntiles when every queue is empty.  Each
// queue counter passes its length at most once per workgroup, so every
// workgroup ends after at most kBatchQueues failed takes.
__device__ __forceinline__ uint32_t batch_take(const BatchArgs &b)
{
    const uint32_t x0 = blockIdx.x % kBatchQueues;
    for (uint32_t k = 0; k < kBatchQueues; k++) {
        const uint32_t x = (x0 + k) % kBatchQueues;
        const uint32_t n = b.qstart[x + 1] - b.qstart[x];
        if (n == 0) continue;
        const uint32_t t = atomicAdd(b.queue + x, 1u);
        if (t < n) return b.qstart[x] + t;
    }
    return b.ntiles;
}

// persistent kernels: every workgroup takes tiles (batch_take) until none is left
// lat_modes: bit 0 set if some instance keeps its tile's lattices in LDS,
// bit 1 if some reads them from HBM; nrs: the rounds per chunk (1..4) that
// every instance has, else 0.  The kernel built for just those is launched.
hipError_t launch_batch_team(hipStream_t st, int fp64, const BatchArgs &b, uint32_t grid,
                             size_t lds_bytes, int lat_modes, int nrs);
hipError_t launch_batch_lane(hipStream_t st, int fp64, const BatchArgs &b, uint32_t grid,
                             uint32_t block, size_t lds_bytes);
// the kernels' dynamic-LDS limit, set once outside any graph capture
hipError_t prepare_batch_team(int fp64, size_t lds_bytes, int lat_modes, int nrs);
hipError_t prepare_batch_lane(int fp64, size_t lds_bytes);
// one workgroup per instance: histogram bookkeeping, convergence, ln W update
hipError_t launch_batch_update(hipStream_t st, const BatchArgs &b);
hipError_t launch_acclog(hipStream_t st, uint32_t k0, uint32_t n, float *out);
hipError_t launch_acclog_fast(hipStream_t st, uint32_t k0, uint32_t n, float *out);

// Exact per-D moments of a walker block, accumulated by the update kernel in
// place of the joint histogram: kMomWords arrays of N+1 uint64 each,
// S0 = count, S1 = sum |M|, then S2 = sum M^2 and S4 = sum M^4 each as three
// words w0, w1, w2 with S = w0 + w1 2^32 + w2 2^64.  Every add puts less than
// 2^32 into each word, so the words are exact while a bin holds fewer than
// 2^32 entries (S0 < 2^32); launch_moments_fp64 flags bins beyond that.
constexpr int kMomWords = 8;
// team kernel: per team a window of kMomSlots consecutive D values, each
// slot kMomLdsWords uint64 in LDS (count, sum |M|, sum M^2, and sum M^4 as
// three 32-bit limb sums), flushed into the words above when D leaves it
constexpr int kMomSlots = 16;
constexpr int kMomLdsWords = 6;
// FP64 moments [block][4][N+1] (k_joint_moments' layout); *range_flag (zeroed
// by the caller) is set to 1 if some bin has S0 >= 2^32
hipError_t launch_moments_fp64(hipStream_t st, const uint64_t *moms, uint32_t N,
                               uint32_t blocks, double *mom, uint32_t *range_flag);
// joint H(D,M) -> per-D moments (S0, S1, S2, S4) for each block, FP64
hipError_t launch_joint_moments(hipStream_t st, const uint64_t *joint, uint32_t N,
                                uint64_t stride, uint32_t blocks, double *mom);
// p[0..n) = v (a kernel, so that callers may capture it in a graph)
hipError_t launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t n, uint32_t v);
// D histogram of a launch that kept per-D moments: hist[D] = sum over the
// blocks of S0[D]
hipError_t launch_hist_from_moments(hipStream_t st, const uint64_t *moms, uint32_t N,
                                    uint32_t blocks, uint64_t *hist);
// reweighting of per-D moments at ndelta crystal fields; out holds three
// planes of ndelta*blocks values (<|M|>, chi, Binder), [delta][block] each
hipError_t launch_reweight(hipStream_t st, const double *mom, const double *lnw, uint32_t N,
                           uint32_t blocks, double beta, const double *deltas,
                           uint32_t ndelta, double *out);

}  // namespace muca
