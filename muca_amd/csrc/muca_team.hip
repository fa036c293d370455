// muca_team.hip -- the team update kernel's launch side: LDS layout,
// dispatch over the instantiations (compiled in muca_team_r{7,10}_{f32,f64}.hip
// so that the build can run them in parallel), and the
// batched engine's persistent team-tile kernel.
#include "team_run.h"

namespace muca {

#define MUCA_TEAM_EXTERN(R, FP64, KEYING)                                                \
    extern template hipError_t launch_team_r<R, FP64, KEYING>(hipStream_t, const McsArgs &, \
                                                              const TeamLayout &);
MUCA_TEAM_EXTERN(7, false, 0)
MUCA_TEAM_EXTERN(7, false, 1)
MUCA_TEAM_EXTERN(7, true, 0)
MUCA_TEAM_EXTERN(7, true, 1)
MUCA_TEAM_EXTERN(10, false, 0)
MUCA_TEAM_EXTERN(10, false, 1)
MUCA_TEAM_EXTERN(10, true, 0)
MUCA_TEAM_EXTERN(10, true, 1)
#undef MUCA_TEAM_EXTERN

bool team_layout(const McsArgs &a, int fp64, size_t lds_max, TeamLayout &lay)
{
    if (a.L % 8 != 0 || a.L > 256) return false;
    const size_t tsz = fp64 ? 8 : 4;
    const size_t row = a.L / 2;
    auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
    lay.row = (uint32_t)row;
    // Philox ring: a chunk of a half-row (at most 64 attempts) needs chunk/4
    // blocks per key and a generation pass may run up to 7 blocks ahead of
    // them; then the team's ln W window
    lay.ring = (uint32_t)((row < 64 ? row : 64) / 4 + 8);
    lay.off_win = (uint32_t)al(2 * (size_t)lay.ring * 16);
    lay.team_bytes = (uint32_t)(lay.off_win + al((kWin + 2) * tsz));
    // per-D moments (produce path, FP64): the team's window of moment slots;
    // they replace the D histogram
    const bool moms = fp64 && a.moms != nullptr && a.joint == nullptr;
    lay.off_mom = lay.team_bytes;
    if (moms) lay.team_bytes += (uint32_t)(kMomLdsWords * kMomSlots * sizeof(uint64_t));
    size_t off = 0;
    lay.off_team = 0;
    off += al((size_t)lay.team_bytes * kTeams);
    lay.off_tab = (uint32_t)off;
    off += 256 * sizeof(double);
    lay.off_fde = (uint32_t)off;
    off += al(17 * tsz);
    lay.off_w = (uint32_t)off;
    // ln W in LDS when it fits (with the guards), else read from HBM, where it
    // stays in L2 (512 KB at L = 256 in FP64), one window per team at a time
    const size_t wbytes = al((a.N + 1 + 2 * kGuard) * tsz);
    lay.lnw_global = off + wbytes > lds_max;
    if (!lay.lnw_global) off += wbytes;
    // histogram and lattice in LDS when they fit (histogram first)
    const bool hist_fits = !moms && off + al((a.N + 1) * 4) <= lds_max;
    lay.lds_hist = hist_fits;
    lay.off_hist = (uint32_t)off;
    if (hist_fits) off += al((a.N + 1) * 4);
    lay.off_lat = (uint32_t)off;
    // walker-major lattices: the lanes of a team read sites two bytes apart
    // (one dword per two lanes); a per-walker stride of 8 (mod 32) dwords puts
    // the four teams of a wavefront on disjoint LDS banks.
    uint32_t stride = (uint32_t)((a.N + 3) & ~3u);
    while ((stride / 4) % 32 != 8) stride += 4;
    lay.lat_stride = stride;
    lay.lat_lds = off + al((size_t)stride * kTeams) <= lds_max;
    if (lay.lat_lds) off += al((size_t)stride * kTeams);
    lay.bytes = off;
    return off <= lds_max;
}

hipError_t launch_mcs_team(hipStream_t st, const Launch &cfg, const McsArgs &a,
                           const TeamLayout &lay)
{
    if (cfg.rounds == 10) {
        if (cfg.fp64)
            return cfg.keying ? launch_team_r<10, true, 1>(st, a, lay)
                              : launch_team_r<10, true, 0>(st, a, lay);
        return cfg.keying ? launch_team_r<10, false, 1>(st, a, lay)
                          : launch_team_r<10, false, 0>(st, a, lay);
    }
    if (cfg.fp64)
        return cfg.keying ? launch_team_r<7, true, 1>(st, a, lay)
                          : launch_team_r<7, true, 0>(st, a, lay);
    return cfg.keying ? launch_team_r<7, false, 1>(st, a, lay)
                      : launch_team_r<7, false, 0>(st, a, lay);
}

// ---------------------------------------------------------------------------
// batched engine, team tiles: a persistent grid (one workgroup per compute
// unit; each keeps ~100 KB of LDS) takes tiles of sixteen walkers from the
// per-XCD tile queues, largest lattices first, and runs one weight iteration
// of the tile's instance.  Converged instances are skipped.  Weight keying,
// Philox4x32-7, walker ids 0..W-1 within each instance, seed = instance seed.
// ---------------------------------------------------------------------------
// LM: the lattice placements this build of the kernel handles (1: LDS,
// 2: HBM, 3: both); NRS: the rounds per chunk of every instance (0: any).
// A batch of one placement (and one chunk size) launches the build for it
// alone, whose registers and schedule the other variants' code does not
// constrain.
template <bool FP64, int LM, int NRS>
__global__ __launch_bounds__(256) void k_batch_team(BatchArgs b)
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
        const TeamLayout lay = I->lay;
        const uint32_t nr = (lay.row + 15) / 16;
        const int v = (lay.lat_lds ? 1 : 0) | (lay.lds_hist ? 2 : 0);
#define MUCA_BT(NR)                                                                            \
    switch (v) {                                                                               \
    case 0:                                                                                    \
        if constexpr ((LM & 2) != 0)                                                           \
            team_run<7, FP64, 0, false, false, 0, false, NR>(a, lay, tile.wbase, lds);         \
        break;                                                                                 \
    case 1:                                                                                    \
        if constexpr ((LM & 1) != 0)                                                           \
            team_run<7, FP64, 0, true, false, 0, false, NR>(a, lay, tile.wbase, lds);          \
        break;                                                                                 \
    case 2:                                                                                    \
        if constexpr ((LM & 2) != 0)                                                           \
            team_run<7, FP64, 0, false, true, 0, false, NR>(a, lay, tile.wbase, lds);          \
        break;                                                                                 \
    default:                                                                                   \
        if constexpr ((LM & 1) != 0)                                                           \
            team_run<7, FP64, 0, true, true, 0, false, NR>(a, lay, tile.wbase, lds);           \
        break;                                                                                 \
    }
        if constexpr (NRS != 0) {
            MUCA_BT(NRS)
        } else if (nr <= 1) {
            MUCA_BT(1)
        } else if (nr <= 2) {
            MUCA_BT(2)
        } else if (nr <= 3) {
            MUCA_BT(3)
        } else {
            MUCA_BT(4)
        }
#undef MUCA_BT
    }
}

template <bool FP64>
static void (*batch_team_kernel(int lat_modes, int nrs))(BatchArgs)
{
    if (lat_modes == 1) {
        switch (nrs) {
        case 1: return k_batch_team<FP64, 1, 1>;
        case 2: return k_batch_team<FP64, 1, 2>;
        case 3: return k_batch_team<FP64, 1, 3>;
        case 4: return k_batch_team<FP64, 1, 4>;
        default: return k_batch_team<FP64, 1, 0>;
        }
    }
    return lat_modes == 2 ? k_batch_team<FP64, 2, 0> : k_batch_team<FP64, 3, 0>;
}

hipError_t launch_batch_team(hipStream_t st, int fp64, const BatchArgs &b, uint32_t grid,
                             size_t lds_bytes, int lat_modes, int nrs)
{
    void (*k)(BatchArgs) = fp64 ? batch_team_kernel<true>(lat_modes, nrs)
                                : batch_team_kernel<false>(lat_modes, nrs);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kTeams * kTeam), lds_bytes, st, b);
    return hipGetLastError();
}

hipError_t prepare_batch_team(int fp64, size_t lds_bytes, int lat_modes, int nrs)
{
    void (*k)(BatchArgs) = fp64 ? batch_team_kernel<true>(lat_modes, nrs)
                                : batch_team_kernel<false>(lat_modes, nrs);
    return hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_bytes);
}

}  // namespace muca
