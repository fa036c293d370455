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
    uint64_t W;              // walkers in the lattice
    uint32_t N, L;
    uint32_t w_offset;       // logical id of walker 0 (for keys)
    uint32_t iter, seed;
    uint64_t therm, nup;
    double beta;
    uint64_t joint_stride;   // (N+1)(N+2)/2
    uint64_t block_walkers;  // walkers per joint block
    uint32_t joint_blocks;
};

struct Launch {
    int rounds;      // 7 or 10
    int fp64;        // acceptance precision
    int keying;      // 0 weight, 1 produce
    int kernel;      // MUCA_KERNEL_*
    int block;       // threads per block
    size_t lds_max;  // bytes of LDS available per block
};

hipError_t launch_init_lattice(hipStream_t st, int8_t *lat, uint64_t W, uint32_t N,
                               int keying, int rounds, uint32_t mm, uint32_t gpu,
                               uint32_t seed, uint32_t w_offset);
hipError_t launch_state(hipStream_t st, const int8_t *lat, uint64_t W, uint32_t N,
                        const uint32_t *naver, int32_t *E, int32_t *D, int32_t *M);
hipError_t launch_mcs(hipStream_t st, const Launch &cfg, const McsArgs &a);

// team kernel (muca_team.hip): LDS layout of one workgroup of 16 walkers
struct TeamLayout {
    uint32_t row;         // L/2 attempts per half-row
    uint32_t team_bytes;  // per-team Philox words, attempt records and D sequence
    uint32_t off_rec, off_dseq;  // offsets inside a team's area
    uint32_t off_team, off_tab, off_w, off_hist, off_lat;
    uint32_t lat_stride;  // LDS bytes per walker lattice (>= N, bank-spread)
    size_t bytes;         // dynamic LDS per workgroup
    bool lds_hist, lat_lds;
};
// false when the team kernel cannot run this launch (L % 8 != 0, L > 256,
// or ln W does not fit in LDS)
bool team_layout(const McsArgs &a, int fp64, size_t lds_max, TeamLayout &lay);
hipError_t launch_mcs_team(hipStream_t st, const Launch &cfg, const McsArgs &a,
                           const TeamLayout &lay);

// lane kernel (muca_lane.hip): even L >= 4 and ln W fits in LDS
bool lane_eligible(const McsArgs &a, int fp64, size_t lds_max);
hipError_t launch_mcs_lane(hipStream_t st, const Launch &cfg, const McsArgs &a);
hipError_t launch_acclog(hipStream_t st, uint32_t k0, uint32_t n, float *out);
hipError_t launch_acclog_fast(hipStream_t st, uint32_t k0, uint32_t n, float *out);

// joint H(D,M) -> per-D moments (S0, S1, S2, S4) for each block, FP64
hipError_t launch_joint_moments(hipStream_t st, const uint64_t *joint, uint32_t N,
                                uint64_t stride, uint32_t blocks, double *mom);
// reweighting of per-D moments at ndelta crystal fields
hipError_t launch_reweight(hipStream_t st, const double *mom, const double *lnw, uint32_t N,
                           uint32_t blocks, double beta, const double *deltas,
                           uint32_t ndelta, double *out);

}  // namespace muca
