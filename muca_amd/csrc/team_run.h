// team_run.h -- the team update kernel: sixteen lanes (one DPP row) per
// walker, four walkers per wavefront, sixteen per workgroup.
//
// Why: a walker is a strict serial chain, so with few walkers (4096 walkers
// fill only 64 wavefronts) one-lane-per-walker leaves most SIMDs idle and
// every attempt waits on a chain of dependent memory loads.  The reference's
// even-L site order is column-parity stripes (muca_delta_multigpu.cu:173):
// a half-row of L/2 same-parity sites contains no two neighbours, and every
// neighbour of a half-row was last written by an earlier half-row.  So per
// chunk of a half-row (up to 64 attempts) the team
//   A. computes in parallel, attempt j on lane j % 16: the Philox blocks,
//      the proposed spin, dE, dDel, (T)(-dE*beta) and an acceptance bitmask
//      over a window of 8 (weight path) or 6 (produce path) D values around
//      the D the chunk starts from,
//      screened with the hardware log (the exact log only where the screen
//      is inconclusive);
//   B. runs only the D chain serially: attempt j's mask is broadcast from
//      its lane with a DPP row broadcast and the accept bit is read at the
//      current D -- no memory access and no branch on the chain.  A round of
//      16 steps that starts outside the window gets its masks recomputed
//      around the current D; one in which D leaves the window resumes, with
//      masks recomputed, from the step that started outside (and, if D
//      leaves that window too, is replayed with an exact per-step path);
//   C. writes the accepted spins, sums dE and dM and fills the histogram.
// Every acceptance is computed with the reference's operations and rounding:
//   diff = ((T)(-dE*beta) + lnW[D+dDel]) - lnW[D];  accept iff diff >= 0 or
//   acclog(r1) < diff.  Because acclog(r1) < 0 always, that is exactly
//   acclog(r1) < diff.  Results are bit-identical to the one-lane kernel.
// Requires L % 8 == 0 (whole Philox blocks per half-row) and L <= 256.
//
// LDS (dynamic): per team a ring of Philox blocks [2 keys][min(L/2, 64)/4 + 8][4]; the
// acceptance-log table; (T)(-dE*beta) for dE = -8..8; ln W with guard entries
// on each side; the D histogram (u32, when it fits); the workgroup's lattices
// [N][16] int8 (when they fit).
#pragma once
#include "kernels.h"

#include "acclog.h"
#include "moments.h"
#include "philox.h"

namespace muca {

namespace team_detail {

constexpr int kTeam = 16;
constexpr int kTeams = 16;   // per workgroup of 256 lanes
constexpr uint32_t kXcds = 8;   // accelerator dies of an MI355X, one L2 each
// D values per acceptance mask (the most any launch uses).  Narrower
// windows cut the mask work per attempt and make rounds in which D leaves
// the window more frequent.  The weight path (D moves fast in the early
// iterations from ln W = 0) uses 8: L = 72, 4096 walkers, 40.7 attempts/ns
// per weight-loop step against 39.8 with 6.  The produce path (D drifts
// slowly under near-flat weights) uses 6: L = 256, 16 384 walkers, 50.3
// attempts/ns against 41.8 with 8.
constexpr int kWin = 8;
template <int KEYING>
constexpr int kWinFor = KEYING == 1 ? 6 : 8;
constexpr int kGuard = 24;   // ln W guard entries on each side (> kWin / 2 + 2)

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int32_t team_sum(int32_t v)
{
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, kTeam);
    return v;
}

// the word of x that holds its sign bit (bit 31)
__device__ __forceinline__ uint32_t sign_word(float x)
{
    return __float_as_uint(x);
}
__device__ __forceinline__ uint32_t sign_word(double x)
{
    return (uint32_t)(__double_as_longlong(x) >> 32);
}

// value of lane L of this lane's 16-lane row (v_mov_b32_dpp row_newbcast:L;
// every lane of the row reads a valid source, so no old value is needed)
template <int L>
__device__ __forceinline__ int32_t row_bcast(int32_t v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x150 + L, 0xf, 0xf, false);
}

}  // namespace team_detail
using namespace team_detail;

// NR: rounds of 16 attempts per chunk; a half-row of L / 2 attempts runs in
// chunks of up to 16 NR (NR <= 4 keeps the registers of all rounds' attempts
// low enough for more wavefronts per SIMD)
// OBS: 0 the D histogram; 1 joint H(D,M) per walker block (produce2_multigpu.cu:489);
// 2 exact per-D moments per walker block (the D histogram is then their S0).
// LNWG: ln W read from HBM (it does not fit in LDS), one window per team.
// One workgroup's sixteen walkers wbase .. wbase+15 of the launch described
// by a (all 256 lanes of the workgroup call it together).
template <int R, bool FP64, int KEYING, bool LATLDS, bool LDSHIST, int OBS, bool LNWG, int NR>
__device__ __forceinline__ void team_run(const McsArgs &a, const TeamLayout &lay, uint64_t wbase,
                                         unsigned char *lds)
{
    typedef typename std::conditional<FP64, double, float>::type T;
    constexpr int KW = kWinFor<KEYING>;   // this launch's mask width (<= kWin)
    constexpr bool JOINT = OBS == 1, MOMS = OBS == 2, TRACKM = OBS >= 1;
    __syncthreads();   // LDS may still be read by a previous call's flush
    T *wt = (T *)(lds + lay.off_w) + kGuard;   // valid for [-kGuard, N + kGuard]
    uint32_t *hist = (uint32_t *)(lds + lay.off_hist);
    int8_t *latl = (int8_t *)(lds + lay.off_lat);
    double *logtab = (double *)(lds + lay.off_tab);
    T *fdet = (T *)(lds + lay.off_fde) + 8;    // fdet[dE] = (T)(-dE * beta), |dE| <= 8
    const int tid = threadIdx.x, team = tid >> 4, tl = tid & 15;
    const uint32_t N = a.N, L = a.L, Rw = lay.row;
    unsigned char *tbase = lds + lay.off_team + (size_t)team * lay.team_bytes;
    uint32_t *rng = (uint32_t *)tbase;                          // [2][ring][4]
    T *winb = (T *)(tbase + lay.off_win);                       // LNWG: lnW[dlo - 1 .. dlo + KW]
    // ln W at any D: from LDS, or from HBM (clamped into [0, N]; only
    // positions a walker cannot be at read a clamped value)
    auto lnw_at = [&](int32_t i) -> T {
        if (LNWG) {
            const int32_t c = min(max(i, 0), (int32_t)a.N);
            return FP64 ? (T)a.lnw64[c] : (T)a.lnw32[c];
        }
        return wt[i];
    };

    const uint64_t W = a.W;
    const uint64_t w = wbase + team;
    const bool active = w < W;

    if (!LNWG) {
#pragma unroll 4
        for (int i = (int)tid - kGuard; i <= (int)N + kGuard; i += blockDim.x) {
            T v = (T)0;
            if (i >= 0 && i <= (int)N) v = FP64 ? (T)a.lnw64[i] : (T)a.lnw32[i];
            wt[i] = v;
        }
    }
    // MOMS: this team's moment slots, word k of slot s at msl[k * kMomSlots + s]
    uint64_t *msl = (uint64_t *)(tbase + lay.off_mom);
    if (MOMS)
        for (int k = 0; k < kMomLdsWords; k++) msl[k * kMomSlots + tl] = 0;
    if (LDSHIST)
        for (uint32_t i = tid; i <= N; i += blockDim.x) hist[i] = a.hist_preload;
    acclog_table_to_lds(logtab, tid, blockDim.x);
    if (tid < 17) fdet[tid - 8] = (T)(-(double)(tid - 8) * a.beta);
    // LATLDS: the tile's lattices move between the site-major layout in HBM
    // and walker-major LDS rows.  For sixteen whole walkers on 16-byte
    // aligned rows, a lane moves four consecutive sites at a time: four
    // 16-byte global accesses (one per site: the tile's walkers), a 4x4 byte
    // transpose per four walkers, sixteen dword LDS accesses (one per
    // walker); a lane's four loads are all in flight before the first is
    // used.  Otherwise one byte per lane and step.
    const bool lat_vec = LATLDS && W % kTeams == 0 && wbase + kTeams <= W &&
                         ((uintptr_t)a.lat & 15) == 0;   // N % 64 == 0 (L % 8 == 0)
    if (LATLDS) {
        if (lat_vec) {
            for (uint32_t q = tid; q < N / 4; q += blockDim.x) {
                uint32_t s4[4][4];   // [site 4q+j][walkers 4c .. 4c+3]
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint4 v = *(const uint4 *)(a.lat + (uint64_t)(4 * q + j) * W + wbase);
                    s4[j][0] = v.x;
                    s4[j][1] = v.y;
                    s4[j][2] = v.z;
                    s4[j][3] = v.w;
                }
#pragma unroll
                for (int t = 0; t < kTeams; t++) {
                    const int c = t >> 2, sh = (t & 3) * 8;
                    *(uint32_t *)(latl + t * lay.lat_stride + 4 * q) =
                        ((s4[0][c] >> sh) & 0xffu) | (((s4[1][c] >> sh) & 0xffu) << 8) |
                        (((s4[2][c] >> sh) & 0xffu) << 16) | (((s4[3][c] >> sh) & 0xffu) << 24);
                }
            }
        } else {
            for (uint32_t e = tid; e < N * kTeams; e += blockDim.x) {
                const uint32_t site = e >> 4, t = e & 15;
                latl[t * lay.lat_stride + site] =
                    (wbase + t < W) ? a.lat[(uint64_t)site * W + wbase + t] : (int8_t)0;
            }
        }
    }
    __syncthreads();

    if (active) {
        auto sp = [&](uint32_t site) -> int8_t & {
            return LATLDS ? latl[team * lay.lat_stride + site] : a.lat[(uint64_t)site * W + w];
        };
        const uint32_t wid = a.w_offset + (uint32_t)w;
        const uint32_t k1a = KEYING == 0 ? 2u * wid : wid;
        const uint32_t k1b = 0xdecafbadu;
        const uint32_t k3a = KEYING == 0 ? 2u * wid + 50000u : wid + 40000u;
        const uint32_t k3b = KEYING == 0 ? 0x11cafeadu : 0xdecafeadu;
        int32_t d = a.D[w], e = a.E[w], m = a.M[w];
        uint64_t *jbase = nullptr;
        if (JOINT) jbase = a.joint + (w / a.block_walkers) * a.joint_stride;
        // MOMS: the walker's production attempts are summed per D in the
        // team's LDS slots for D = mwb .. mwb + kMomSlots - 1 (mwb the same in
        // every lane of the team) and added to the block's moments in HBM when
        // D leaves that window and at the end
        const uint64_t mstride = (uint64_t)N + 1;
        uint64_t *mbase = nullptr;
        if (MOMS) mbase = a.moms + (w / a.block_walkers) * (uint64_t)kMomWords * mstride;
        int32_t mwb = -(1 << 30);   // no window yet
        // lane tl moves slot tl (D = mwb + tl) to HBM and clears it
        auto flush_slot = [&]() {
            uint64_t v[kMomLdsWords];
#pragma unroll
            for (int k = 0; k < kMomLdsWords; k++) v[k] = msl[k * kMomSlots + tl];
            if (v[0]) {
                const uint64_t dd = (uint64_t)(mwb + (int32_t)tl);   // a visited D, in [0, N]
                atomicAdd((unsigned long long *)&mbase[dd], (unsigned long long)v[0]);
                atomicAdd((unsigned long long *)&mbase[mstride + dd], (unsigned long long)v[1]);
                atomic_add_limbs(&mbase[2 * mstride + dd], &mbase[3 * mstride + dd],
                                 &mbase[4 * mstride + dd], v[2], 0);
                // sum M^4 = v3 + v4 2^32 + v5 2^64 (limb sums) as a 128-bit value
                const uint64_t lo = v[3] + (v[4] << 32);
                const uint64_t hi = v[5] + (v[4] >> 32) + (lo < v[3] ? 1u : 0u);
                atomic_add_limbs(&mbase[5 * mstride + dd], &mbase[6 * mstride + dd],
                                 &mbase[7 * mstride + dd], lo, hi);
#pragma unroll
                for (int k = 0; k < kMomLdsWords; k++) msl[k * kMomSlots + tl] = 0;
            }
        };
        // joint-histogram entries of this lane's attempts, counted for the bin
        // in use and the one before it and added with one atomic when a third
        // bin comes up (a walker stepping back and forth adds nothing)
        uint64_t jrun = 0, jprun = 0;
        int64_t jbin = -1, jpbin = -1;
        uint32_t c0base = 0;
        const uint32_t RB = lay.ring;
        uint32_t gen = 0, gen_slot = 0, cur_slot = 0;   // next block to make, its slot, block blk0's slot
        // PF (lattices in HBM, ln W in LDS): the spins a chunk reads are
        // loaded while the chunk before it runs its D chain, so the loads'
        // latency is hidden.  pf_* hold, per round, the next chunk's site,
        // its right, left and upper neighbours and the lower neighbour as
        // read from memory; with pf_fwd the lower neighbour is instead this
        // lane's own site of the chunk before (same round and lane, the row
        // below), whose final spin fin[] phase C keeps (-9: none).
        constexpr bool PF = !LATLDS && !LNWG;
        int32_t pf_s[NR], pf_r[NR], pf_l[NR], pf_u[NR], pf_d[NR], fin[NR];
        bool pf_have = false, pf_fwd = false;
        int32_t de_lane = 0, dm_lane = 0;
        // a workgroup adds at most 16 NUPDATES counts to one u32 bin
        const bool may_wrap = a.hist_preload != 0 || (uint64_t)kTeams * a.nup >= (1ull << 32);
        for (int phase = 0; phase < 2; phase++) {
            const uint64_t n = phase ? a.nup : a.therm;
            // A half-row is worked in chunks of up to 16 NR attempts (no two
            // sites of a half-row are neighbours, so any split of it is exact).
            // Attempt rs of a phase is at sweep position rs mod N: half h
            // (x parity), row y, place jb in the half-row -- advanced chunk by
            // chunk from 0, without divisions.
            uint32_t h = 0, y = 0, jb = 0;
            for (uint64_t rs = 0; rs < n;) {
                const uint32_t cnt = (uint32_t)min((uint64_t)min(Rw - jb, 16u * NR), n - rs);
                // ---- A1: the walker's Philox stream (block b uses counter b + 1;
                // the stream runs on across chunks and phases) is generated
                // eight blocks x two keys per pass into a ring of `RB` blocks per
                // key, so no lane idles when a chunk's block count is not a
                // multiple of eight.  Blocks [blk0, blk0 + nblk) serve this
                // chunk; a pass may run up to 7 blocks ahead, and RB >= nblk + 8
                // keeps those from overwriting blocks still in use.
                const uint32_t nblk = (cnt + 3) / 4;
                const uint32_t blk0 = c0base + (uint32_t)(rs / 4);
                while (gen < blk0 + nblk) {
                    const uint32_t b = gen + (uint32_t)(tl >> 1), key = tl & 1;
                    const philox4 r = philox4x32<R>(b + 1, a.seed, a.iter, 0xBADC0DEDu,
                                                    key ? k3a : k1a, key ? k3b : k1b);
                    uint32_t slot = gen_slot + (uint32_t)(tl >> 1);
                    slot = slot >= RB ? slot - RB : slot;
                    uint32_t *dst = rng + (key * RB + slot) * 4;
                    dst[0] = r.v[0];
                    dst[1] = r.v[1];
                    dst[2] = r.v[2];
                    dst[3] = r.v[3];
                    gen += 8;
                    gen_slot = gen_slot + 8 >= RB ? gen_slot + 8 - RB : gen_slot + 8;
                }
                // ring slot of this chunk's first block (blocks are contiguous
                // from one chunk to the next, across the phase change too)
                const uint32_t slot0 = cur_slot;
                cur_slot = cur_slot + nblk >= RB ? cur_slot + nblk - RB : cur_slot + nblk;
                wave_sync();
                // ---- A2: attempt j = 16 r + tl on this lane, everything in registers
                int32_t dlo = d - KW / 2;           // mask bit k <-> D = dlo + k
                // lnW[dlo + k] at wcen[k], k = -1 .. KW (LNWG: staged in the
                // team's window from HBM)
                auto window = [&]() -> const T * {
                    if (LNWG) {
                        if (tl < KW + 2) winb[tl] = lnw_at(dlo - 1 + tl);
                        wave_sync();
                        return winb + 1;
                    }
                    return wt + dlo;
                };
                const T *wcen = window();
                T wfrom[KW];                        // lnW[dlo .. dlo + KW - 1]
#pragma unroll
                for (int k = 0; k < KW; k++) wfrom[k] = wcen[k];
                const int32_t dlo_a = dlo;            // the window of every mask made here
                const uint32_t yu = (y + 1 == L) ? 0 : y + 1, yd = (y == 0) ? L - 1 : y - 1;
                const uint32_t s4 = slot0 * 4, ringw = RB * 4;   // ring word of attempt j: s4 + j (mod ringw)
                uint32_t mask[NR];
                // dDel (2-bit signed) | (dE + 8) << 2 | (new - old + 2) << 7 | (new + 1) << 10
                int32_t info[NR];
                // Bits k = 0..KW-1: accept at D = dlo + k, i.e. acclog(r1) <
                // ((fde + lnW[D + dDel]) - lnW[D]), screened with the hardware
                // log and its proven band: la + band < diff accepts, la - band
                // >= diff rejects.  Bit k is the sign of e = thi - diff (e is
                // -0 never, +0 iff thi == diff, which rejects); a diff with
                // tlo < diff <= thi, i.e. 0 <= e < thi - tlo, needs the exact
                // log -- flagged (with margin) by min |e| < 2 (thi - tlo).
                auto accept_bits = [&](const T *wf, const T *wc, uint32_t r1, T fde,
                                       int32_t ddel) -> uint32_t {
                    const T *wto = wc + ddel;
                    T diff[KW];
#pragma unroll
                    for (int k = 0; k < KW; k++) {
                        diff[k] = fde + wto[k];
                        diff[k] = diff[k] - wf[k];
                    }
                    const float la = acclog_fast_u01(r1);
                    const float band = acclog_fast_band(la);
                    const T thi = (T)(la + band), tlo = (T)(la - band);
                    uint32_t mk = 0;
                    T emin = (T)1e30f;
#pragma unroll
                    for (int k = KW - 1; k >= 0; k--) {
                        const T e = thi - diff[k];
                        mk = __builtin_amdgcn_alignbit(mk, sign_word(e), 31);
                        emin = fmin(emin, fabs(e));
                    }
                    const bool amb = emin < (thi - tlo) + (thi - tlo);
                    if (__builtin_amdgcn_ballot_w64(amb)) {   // rare: a diff within the band
                        if (amb) {
                            const T tlog =
                                (T)(float)acclog_u01_double_tab(r1, logtab, logtab + 128);
                            mk = 0;
#pragma unroll
                            for (int k = KW - 1; k >= 0; k--)
                                mk = mk + mk + (tlog < diff[k] ? 1u : 0u);
                        }
                    }
                    return mk;
                };
                // Branch-free over the rounds so that the loads of all rounds
                // can be in flight together: a lane past the end of the
                // chunk works on a clamped (in-bounds) attempt and keeps
                // nothing of it.
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    const uint32_t j = 16 * r + tl;
                    const bool valid = j < cnt;
                    const uint32_t jc = valid ? j : cnt - 1;
                    const uint32_t x = 2 * (jb + jc) + h;
                    const uint32_t xr = (x + 1 == L) ? 0 : x + 1;
                    const uint32_t xl = (x == 0) ? L - 1 : x - 1;
                    int32_t old_s, nb;
                    if (PF && pf_have) {
                        old_s = pf_s[r];
                        nb = pf_r[r] + pf_l[r] + pf_u[r] + ((pf_fwd && fin[r] != -9) ? fin[r] : pf_d[r]);
                    } else {
                        old_s = sp(y * L + x);
                        nb = sp(y * L + xr) + sp(y * L + xl) + sp(yu * L + x) + sp(yd * L + x);
                    }
                    uint32_t widx = s4 + jc;
                    widx = widx >= ringw ? widx - ringw : widx;
                    const uint32_t r1 = rng[widx], r3 = rng[ringw + widx];
                    int32_t new_s = old_s + (r3 < 0x80000000u ? 1 : -1);
                    new_s = new_s == 2 ? -1 : (new_s == -2 ? 1 : new_s);
                    const int32_t ds = new_s - old_s;
                    const int32_t dE = -ds * nb;
                    const int32_t ddel = new_s * new_s - old_s * old_s;
                    // (T)(-dE * beta) by arithmetic (the same operations as
                    // the table in LDS, without a dependent LDS load)
                    const uint32_t mk =
                        accept_bits(wfrom, wcen, r1, (T)(-(double)dE * a.beta), ddel);
                    // bits 0..KW-1: accept at D = dlo + k; bits 16..17: dDel
                    mask[r] = mk | (valid ? ((uint32_t)ddel & 3u) << 16 : 0u);
                    if (JOINT)
                        info[r] = valid ? ((ddel & 3) | ((dE + 8) << 2) | ((ds + 2) << 7) |
                                           ((new_s + 1) << 10))
                                        : ((8 << 2) | (2 << 7));
                    else
                        info[r] = valid ? ((dE + 8) << 2) | ((ds + 2) << 7) | ((new_s + 1) << 10)
                                        : ((8 << 2) | (2 << 7));
                }
                wave_sync();
                if (PF) {
                    // The next chunk (the next phase's first after a phase's
                    // last) and whether its spins may be read now: this chunk
                    // writes sites jb .. jb + cnt - 1 of half-row (y, h) only.
                    // Reads of that half-row are safe when their sites lie
                    // outside it; the lower neighbours of the next row's
                    // chunk at the same jb come from fin[] (lanes within this
                    // chunk) or from memory (lanes past it: not written).
                    uint64_t rs_n = rs + cnt, n_n = n;
                    bool nx = rs_n < n;
                    uint32_t h_n = h, y_n = y, jb_n = jb + cnt;
                    if (jb_n == Rw) {
                        jb_n = 0;
                        if (++y_n == L) {
                            y_n = 0;
                            h_n ^= 1u;
                        }
                    }
                    if (!nx && phase == 0 && a.nup > 0) {
                        rs_n = 0;
                        n_n = a.nup;
                        nx = true;
                        h_n = y_n = jb_n = 0;
                    }
                    const uint32_t cnt_n = (uint32_t)min((uint64_t)min(Rw - jb_n, 16u * NR), n_n - rs_n);
                    const uint32_t yu_n = (y_n + 1 == L) ? 0 : y_n + 1;
                    const uint32_t yd_n = (y_n == 0) ? L - 1 : y_n - 1;
                    const bool ovl = jb_n < jb + cnt && jb < jb_n + cnt_n;
                    bool have = nx;
                    bool fwd = false;
                    if (y_n == y && h_n != h) have = false;          // neighbours in the written half-row
                    if (h_n == h && (y_n == y || yu_n == y) && ovl) have = false;
                    if (h_n == h && yd_n == y && ovl) {
                        if (jb_n == jb) fwd = true;
                        else have = false;
                    }
                    pf_have = have;
                    pf_fwd = fwd;
                    if (have) {
#pragma unroll
                        for (int r = 0; r < NR; r++) {
                            const uint32_t jn = min(16u * r + tl, cnt_n - 1);
                            const uint32_t x = 2 * (jb_n + jn) + h_n;
                            const uint32_t xr = (x + 1 == L) ? 0 : x + 1;
                            const uint32_t xl = (x == 0) ? L - 1 : x - 1;
                            pf_s[r] = sp(y_n * L + x);
                            pf_r[r] = sp(y_n * L + xr);
                            pf_l[r] = sp(y_n * L + xl);
                            pf_u[r] = sp(yu_n * L + x);
                            pf_d[r] = sp(yd_n * L + x);
                        }
                    }
                }
                // ---- B: the D chain.  The chain runs on `off` = D - dlo (the
                // window position); per step one DPP broadcast of the attempt's
                // mask word, the accept bit at the current position, its dDel,
                // and the move.  A position outside the window reads a zero bit
                // (bits KW..15 and 24..31 of the word are zero), so the chain
                // then stays put; the D recovered after the round shows such an
                // excursion, and the round is replayed exactly.
                uint32_t accb[NR];
                int32_t dv[NR], mv[NR];
                const uint32_t sh = (uint32_t)(tid & 48);    // this team's 16 lanes of a ballot
                const uint32_t upto = (2u << tl) - 1u;       // steps 0..tl of a round
                int32_t off = d - dlo;
                // D after this lane's step from the round's start, the accept bits
                // and the dDel of every step (two ballots)
                auto recover = [&](int r, int32_t off0, uint32_t ab) -> int32_t {
                    const int32_t my = __builtin_amdgcn_sbfe((int32_t)mask[r], 16, 2);
                    const uint32_t up =
                        (uint32_t)(__builtin_amdgcn_ballot_w64(my > 0) >> sh) & 0xffffu & ab;
                    const uint32_t dn =
                        (uint32_t)(__builtin_amdgcn_ballot_w64(my < 0) >> sh) & 0xffffu & ab;
                    return off0 + __builtin_popcount(up & upto) - __builtin_popcount(dn & upto);
                };
                // One chain step.  CHK = false: no memory access and no branch.
                // CHK = true: a step that starts outside the window is decided
                // exactly from the attempt's record in LDS.
                // MODE 0: no memory access and no branch.  MODE 1: a step that
                // starts outside the window is decided exactly from the
                // attempt's record.  MODE 2: as 0 from step kres on, steps
                // before kres left alone.
                uint32_t kres = 0;
                auto step = [&](auto lc, auto mode, int r, uint32_t &ab) {
                    constexpr int l = decltype(lc)::value;
                    constexpr int MODE = decltype(mode)::value;
                    constexpr bool CHK = MODE == 1;
                    const uint32_t mj = (uint32_t)row_bcast<l>((int32_t)mask[r]);
                    const int32_t ddel = __builtin_amdgcn_sbfe((int32_t)mj, 16, 2);
                    // v_bfe_u32 reads the offset modulo 32
                    uint32_t acc = __builtin_amdgcn_ubfe(mj, (uint32_t)off, 1);
                    if (MODE == 2) acc = (uint32_t)l >= kres ? acc : 0u;
                    if (CHK) {
                        // a step outside the window: the attempt's dE from its
                        // info word, its r1 from the ring, the screened log
                        // (exact only inside the band)
                        const bool oob = (uint32_t)off >= (uint32_t)KW;
                        if (__builtin_amdgcn_ballot_w64(oob)) {
                            if (oob) {
                                const int32_t ij = row_bcast<l>(info[r]);
                                uint32_t w1 = s4 + 16 * r + l;
                                w1 = w1 >= ringw ? w1 - ringw : w1;
                                const uint32_t r1 = rng[w1];
                                const int32_t dd = dlo + off;
                                T diff = fdet[((ij >> 2) & 31) - 8] + lnw_at(dd + ddel);
                                diff = diff - lnw_at(dd);
                                const float la = acclog_fast_u01(r1);
                                const float band = acclog_fast_band(la);
                                acc = (T)(la + band) < diff ? 1u : 0u;
                                const bool amb = !acc && (T)(la - band) < diff;
                                if (__builtin_amdgcn_ballot_w64(amb)) {
                                    if (amb)
                                        acc = (T)(float)acclog_u01_double_tab(r1, logtab,
                                                                              logtab + 128) < diff
                                                  ? 1u
                                                  : 0u;
                                }
                            }
                        }
                    }
                    off += (int32_t)acc * ddel;
                    ab |= acc << l;
                };
                using Fast = std::integral_constant<int, 0>;
                using Checked = std::integral_constant<int, 1>;
                using Resume = std::integral_constant<int, 2>;
                // round r's masks recomputed around dlo by all lanes of the team
                auto remask = [&](uint32_t &mk_r, int32_t inf_r, int r) {
                    const T *wc = window();
                    T wf[KW];
#pragma unroll
                    for (int k = 0; k < KW; k++) wf[k] = wc[k];
                    uint32_t w1 = s4 + min(16 * (uint32_t)r + tl, cnt - 1);
                    w1 = w1 >= ringw ? w1 - ringw : w1;
                    const int32_t dd = __builtin_amdgcn_sbfe((int32_t)mk_r, 16, 2);
                    const uint32_t mk =
                        accept_bits(wf, wc, rng[w1], fdet[((inf_r >> 2) & 31) - 8], dd);
                    mk_r = mk | (mk_r & 0x30000u);
                };
#define MUCA_STEPS(K, AB)                                                                    \
    if (left >= 16) {                                                                        \
        step(std::integral_constant<int, 0>(), K(), r, AB);                                  \
        step(std::integral_constant<int, 1>(), K(), r, AB);                                  \
        step(std::integral_constant<int, 2>(), K(), r, AB);                                  \
        step(std::integral_constant<int, 3>(), K(), r, AB);                                  \
        step(std::integral_constant<int, 4>(), K(), r, AB);                                  \
        step(std::integral_constant<int, 5>(), K(), r, AB);                                  \
        step(std::integral_constant<int, 6>(), K(), r, AB);                                  \
        step(std::integral_constant<int, 7>(), K(), r, AB);                                  \
        step(std::integral_constant<int, 8>(), K(), r, AB);                                  \
        step(std::integral_constant<int, 9>(), K(), r, AB);                                  \
        step(std::integral_constant<int, 10>(), K(), r, AB);                                 \
        step(std::integral_constant<int, 11>(), K(), r, AB);                                 \
        step(std::integral_constant<int, 12>(), K(), r, AB);                                 \
        step(std::integral_constant<int, 13>(), K(), r, AB);                                 \
        step(std::integral_constant<int, 14>(), K(), r, AB);                                 \
        step(std::integral_constant<int, 15>(), K(), r, AB);                                 \
    } else {                                                                                 \
        if (0 < left) step(std::integral_constant<int, 0>(), K(), r, AB);                    \
        if (1 < left) step(std::integral_constant<int, 1>(), K(), r, AB);                    \
        if (2 < left) step(std::integral_constant<int, 2>(), K(), r, AB);                    \
        if (3 < left) step(std::integral_constant<int, 3>(), K(), r, AB);                    \
        if (4 < left) step(std::integral_constant<int, 4>(), K(), r, AB);                    \
        if (5 < left) step(std::integral_constant<int, 5>(), K(), r, AB);                    \
        if (6 < left) step(std::integral_constant<int, 6>(), K(), r, AB);                    \
        if (7 < left) step(std::integral_constant<int, 7>(), K(), r, AB);                    \
        if (8 < left) step(std::integral_constant<int, 8>(), K(), r, AB);                    \
        if (9 < left) step(std::integral_constant<int, 9>(), K(), r, AB);                    \
        if (10 < left) step(std::integral_constant<int, 10>(), K(), r, AB);                  \
        if (11 < left) step(std::integral_constant<int, 11>(), K(), r, AB);                  \
        if (12 < left) step(std::integral_constant<int, 12>(), K(), r, AB);                  \
        if (13 < left) step(std::integral_constant<int, 13>(), K(), r, AB);                  \
        if (14 < left) step(std::integral_constant<int, 14>(), K(), r, AB);                  \
    }
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    accb[r] = 0;
                    dv[r] = d;
                    mv[r] = m;
                    const uint32_t j0 = 16 * r;
                    if (j0 >= cnt) continue;
                    const uint32_t left = cnt - j0;
                    // the window follows D from round to round: a round that
                    // starts outside the masks' window gets its masks
                    // recomputed around the current D, by all lanes of the
                    // team at once (only this round's: the later rounds'
                    // masks stay relative to dlo_a)
                    dlo = dlo_a;
                    off = d - dlo_a;
                    const bool shift = (uint32_t)off >= (uint32_t)KW;
                    if (__builtin_amdgcn_ballot_w64(shift)) {
                        if (shift) {
                            dlo += off - KW / 2;
                            off = KW / 2;
                            remask(mask[r], info[r], r);
                        }
                    }
                    const int32_t off0 = off;
                    uint32_t ab = 0;
                    MUCA_STEPS(Fast, ab)
                    int32_t o = recover(r, off0, ab);
                    // a step started outside the window iff one before it in the
                    // round ended there (a lane past the chunk's end repeats
                    // the last step's D)
                    const bool out = (uint32_t)o >= (uint32_t)KW && tl + 1 < min(left, 16u);
                    if (__builtin_amdgcn_ballot_w64(out)) {
                        const uint32_t mine =
                            (uint32_t)(__builtin_amdgcn_ballot_w64(out) >> sh) & 0xffffu;
                        if (mine) {
                            // D left the window after step kres - 1: steps
                            // 0 .. kres - 1 stand; the masks are recomputed
                            // around D at the start of step kres and the chain
                            // resumes there
                            kres = (uint32_t)__builtin_ctz(mine) + 1u;
                            const int32_t moved = __shfl(o, (int)kres - 1, kTeam) - KW / 2;
                            dlo += moved;
                            remask(mask[r], info[r], r);
                            const int32_t offr0 = off0 - moved;   // the round's start, new window
                            off = KW / 2;
                            uint32_t abr = 0;
                            MUCA_STEPS(Resume, abr)
                            ab = (ab & ((1u << kres) - 1u)) | abr;
                            o = recover(r, offr0, ab);
                            // rare: D left the new window too -- replay exactly
                            const bool out2 = (uint32_t)o >= (uint32_t)KW &&
                                              tl + 1 < min(left, 16u) && (uint32_t)tl >= kres;
                            if ((uint32_t)(__builtin_amdgcn_ballot_w64(out2) >> sh) & 0xffffu) {
                                off = offr0;
                                ab = 0;
                                MUCA_STEPS(Checked, ab)
                                o = recover(r, offr0, ab);
                            }
                        }
                    }
                    accb[r] = ab;
                    dv[r] = dlo + o;
                    d = dlo + off;
                    if (TRACKM) {
                        // M after this lane's step: an inclusive scan over the
                        // team of the accepted spin changes
                        int32_t x = (int32_t)((ab >> tl) & 1u) * (((info[r] >> 7) & 7) - 2);
                        x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);
                        x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);
                        x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);
                        x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);
                        mv[r] = m + x;
                        m += row_bcast<15>(x);
                    }
                }
#undef MUCA_STEPS
                // ---- C: accepted spins, dE / dM sums, histogram
                int32_t de = 0, dm = 0;
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    const uint32_t j = 16 * r + tl;
                    if (PF) {
                        // the site's spin after this chunk (info: new spin + 1
                        // at bits 10-11, new - old + 2 at bits 7-9)
                        const int32_t ns = ((info[r] >> 10) & 3) - 1;
                        const int32_t os = ns - (((info[r] >> 7) & 7) - 2);
                        fin[r] = j < cnt ? (((accb[r] >> tl) & 1u) ? ns : os) : -9;
                    }
                    if (j < cnt) {
                        if ((accb[r] >> tl) & 1u) {
                            // the new spin is known from phase A: no site of
                            // the chunk is written between its read and here
                            const int32_t ds = ((info[r] >> 7) & 7) - 2;
                            sp(y * L + 2 * (jb + j) + h) = (int8_t)(((info[r] >> 10) & 3) - 1);
                            de += ((info[r] >> 2) & 31) - 8;
                            dm += ds;
                        }
                        if (phase && JOINT) {
                            const int64_t dd = dv[r];
                            const int64_t mm = min(max((int64_t)mv[r], -dd), dd);
                            const int64_t jbn = (mm + dd * dd + 2 * dd) / 2;
                            if (jbn != jbin) {
                                if (jbn == jpbin) {
                                    const uint64_t t = jrun;
                                    jrun = jprun;
                                    jprun = t;
                                    jpbin = jbin;
                                } else {
                                    if (jprun)
                                        atomicAdd((unsigned long long *)&jbase[jpbin], jprun);
                                    jprun = jrun;
                                    jpbin = jbin;
                                    jrun = 0;
                                }
                                jbin = jbn;
                            }
                            jrun++;
                        }
                    }
                }
                if (phase && !MOMS) {
                    // The D histogram: a round's steps split into runs at one D,
                    // starting at step 0 and at every accepted step that
                    // changed D; the first lane of a run adds its length (one
                    // atomic per run instead of sixteen on the same bin)
#pragma unroll
                    for (int r = 0; r < NR; r++) {
                        if (16 * r >= cnt) break;
                        const uint32_t left = min(cnt - 16 * r, 16u);
                        const bool moved = ((accb[r] >> tl) & 1u) && (mask[r] & 0x30000u);
                        const uint32_t starts =
                            ((uint32_t)(__builtin_amdgcn_ballot_w64(moved) >> sh) & 0xffffu) | 1u;
                        const uint32_t later = starts & ~upto;
                        const uint32_t len =
                            (later ? (uint32_t)__builtin_ctz(later) : left) - (uint32_t)tl;
                        if (tl < left && ((starts >> tl) & 1u)) {
                            const int32_t dd = dv[r];
                            if (LDSHIST) {
                                if (may_wrap) {
                                    // a u32 bin that wraps carries 2^32 into HBM
                                    if (atomicAdd(&hist[dd], len) > 0xFFFFFFFFu - len)
                                        atomicAdd((unsigned long long *)&a.hist[dd], 1ull << 32);
                                } else {
                                    atomicAdd(&hist[dd], len);
                                }
                            } else {
                                atomicAdd((unsigned long long *)&a.hist[dd], (unsigned long long)len);
                            }
                        }
                    }
                }
                if (MOMS && phase) {
                    // Attempt j records the (D, M) after its step.  Steps that
                    // are not accepted keep the state, so a round splits into
                    // segments that start at step 0 and at every accepted step
                    // and share one (D, M); the first lane of each segment adds
                    // len x (1, |M|, M^2, M^4) to the slot of its D.
#pragma unroll
                    for (int r = 0; r < NR; r++) {
                        if (16 * r >= cnt) break;
                        const uint32_t left = min(cnt - 16 * r, 16u);
                        const uint32_t ab = accb[r];
                        const bool lead = tl < left && (tl == 0 || ((ab >> tl) & 1u));
                        const uint32_t later = ab & ~upto;
                        const uint32_t len =
                            (later ? (uint32_t)__builtin_ctz(later) : left) - (uint32_t)tl;
                        const int32_t dd = dv[r];
                        int32_t slot = dd - mwb;
                        const bool out = lead && (uint32_t)slot >= (uint32_t)kMomSlots;
                        const uint64_t ob = __builtin_amdgcn_ballot_w64(out);
                        if (ob) {
                            if ((ob >> sh) & 0xffffu) {   // the whole team: new window
                                flush_slot();
                                mwb = row_bcast<0>(dd) - kMomSlots / 2;
                                slot = dd - mwb;
                            }
                        }
                        if (lead) {
                            const int32_t mm = min(max(mv[r], -dd), dd);
                            const uint64_t am = (uint64_t)(mm < 0 ? -mm : mm);
                            const uint64_t m2 = am * am;
                            const uint64_t q = m2 * m2, qh = __umul64hi(m2, m2);   // M^4
                            const uint64_t pl = q * len, ph = qh * len + __umul64hi(q, len);
                            const uint64_t v[kMomLdsWords] = {len, len * am, len * m2,
                                                              pl & 0xFFFFFFFFull, pl >> 32, ph};
                            if ((uint32_t)slot < (uint32_t)kMomSlots) {
                                // no test for zero words: a branch per word
                                // costs more than the add
#pragma unroll
                                for (int k = 0; k < kMomLdsWords; k++)
                                    atomicAdd((unsigned long long *)&msl[k * kMomSlots + slot],
                                              (unsigned long long)v[k]);
                            } else {
                                // D moved more than kMomSlots / 2 within the round
                                atomicAdd((unsigned long long *)&mbase[dd], (unsigned long long)v[0]);
                                atomicAdd((unsigned long long *)&mbase[mstride + dd],
                                          (unsigned long long)v[1]);
                                atomic_add_limbs(&mbase[2 * mstride + dd], &mbase[3 * mstride + dd],
                                                 &mbase[4 * mstride + dd], v[2], 0);
                                atomic_add_limbs(&mbase[5 * mstride + dd], &mbase[6 * mstride + dd],
                                                 &mbase[7 * mstride + dd], pl, ph);
                            }
                        }
                    }
                }
                de_lane += de;
                dm_lane += dm;
                if (!LATLDS) __threadfence_block();
                wave_sync();
                rs += cnt;
                jb += cnt;
                if (jb == Rw) {
                    jb = 0;
                    if (++y == L) {
                        y = 0;
                        h ^= 1u;
                    }
                }
            }
            c0base += (uint32_t)((n + 3) / 4);
        }
        if (JOINT) {
            if (jrun) atomicAdd((unsigned long long *)&jbase[jbin], jrun);
            if (jprun) atomicAdd((unsigned long long *)&jbase[jpbin], jprun);
        }
        if (MOMS) flush_slot();
        // dE and dM of accepted moves were summed per lane; one reduction
        // over the team's sixteen lanes per launch.  In joint mode the
        // chain itself tracks M (it is needed per attempt).
        e += team_sum(de_lane);
        const int32_t dm_team = team_sum(dm_lane);
        if (!TRACKM) m += dm_team;
        if (tl == 0) {
            a.D[w] = d;
            a.E[w] = e;
            a.M[w] = m;
        }
    }
    __syncthreads();
    if (LDSHIST) {
        // counts since the start (modulo 2^32; a wrap was carried already)
        for (uint32_t i = tid; i <= N; i += blockDim.x) {
            const uint32_t c = hist[i];
            if (c != a.hist_preload)
                atomicAdd((unsigned long long *)&a.hist[i],
                          (unsigned long long)((int64_t)c - (int64_t)a.hist_preload));
        }
    }
    if (LATLDS) {
        if (lat_vec) {
            for (uint32_t q = tid; q < N / 4; q += blockDim.x) {
                uint32_t r[kTeams];   // walker t's spins at sites 4q .. 4q+3
#pragma unroll
                for (int t = 0; t < kTeams; t++)
                    r[t] = *(const uint32_t *)(latl + t * lay.lat_stride + 4 * q);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int sh = 8 * j;
                    uint32_t d[4];
#pragma unroll
                    for (int c = 0; c < 4; c++)
                        d[c] = ((r[4 * c] >> sh) & 0xffu) | (((r[4 * c + 1] >> sh) & 0xffu) << 8) |
                               (((r[4 * c + 2] >> sh) & 0xffu) << 16) |
                               (((r[4 * c + 3] >> sh) & 0xffu) << 24);
                    *(uint4 *)(a.lat + (uint64_t)(4 * q + j) * W + wbase) =
                        make_uint4(d[0], d[1], d[2], d[3]);
                }
            }
        } else {
            for (uint32_t e2 = tid; e2 < N * kTeams; e2 += blockDim.x) {
                const uint32_t site = e2 >> 4, t = e2 & 15;
                if (wbase + t < W) a.lat[(uint64_t)site * W + wbase + t] = latl[t * lay.lat_stride + site];
            }
        }
    }
}

// LNWG launches are the large lattices (L >= 136) with many walkers: they
// are held to 128 VGPRs, four wavefronts per SIMD
template <int R, bool FP64, int KEYING, bool LATLDS, bool LDSHIST, int OBS, bool LNWG, int NR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LNWG ? 4 : 1))) void
k_mcs_team(McsArgs a, TeamLayout lay)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    // Workgroups go to the eight XCDs round-robin, and eight consecutive
    // groups of 16 walkers share each 128-byte line of the site-major
    // lattice: workgroup b takes group (b mod 8) (G / 8) + b / 8, so that
    // XCD x runs the consecutive groups x G / 8 .. (x + 1) G / 8 - 1 and the
    // eight sharers of a line use one L2
    const uint32_t G = gridDim.x;
    uint32_t g = blockIdx.x;
    if (G % kXcds == 0) g = (g % kXcds) * (G / kXcds) + g / kXcds;
    team_run<R, FP64, KEYING, LATLDS, LDSHIST, OBS, LNWG, NR>(a, lay, (uint64_t)g * kTeams, lds);
}

template <int R, bool FP64, int KEYING, int NR>
hipError_t launch_team_nr(hipStream_t st, const McsArgs &a, const TeamLayout &lay)
{
    const dim3 grid((unsigned)((a.W + kTeams - 1) / kTeams)), block(kTeams * kTeam);
    const bool joint = a.joint != nullptr;
    // per-D moments (the produce path, FP64) replace the D histogram
    const bool moms = FP64 && a.moms != nullptr && !joint;
    constexpr int MO = FP64 ? 2 : 0;
    void (*k)(McsArgs, TeamLayout) = nullptr;
    if (lay.lnw_global) {
        // ln W too large for LDS (L >= 136): lattices and histogram stay in
        // HBM as well
        if constexpr (NR == 4) {
            k = moms ? k_mcs_team<R, FP64, KEYING, false, false, MO, true, NR>
                : joint ? k_mcs_team<R, FP64, KEYING, false, false, 1, true, NR>
                        : k_mcs_team<R, FP64, KEYING, false, false, 0, true, NR>;
        } else {
            return hipErrorInvalidValue;
        }
    } else if (moms) {
        k = lay.lat_lds ? k_mcs_team<R, FP64, KEYING, true, false, MO, false, NR>
                        : k_mcs_team<R, FP64, KEYING, false, false, MO, false, NR>;
    } else {
#define MUCA_TEAM_PICK(LL, LH)                                                  \
    k = joint ? k_mcs_team<R, FP64, KEYING, LL, LH, 1, false, NR>               \
              : k_mcs_team<R, FP64, KEYING, LL, LH, 0, false, NR>;
        if (lay.lat_lds && lay.lds_hist) {
            MUCA_TEAM_PICK(true, true)
        } else if (lay.lat_lds) {
            MUCA_TEAM_PICK(true, false)
        } else if (lay.lds_hist) {
            MUCA_TEAM_PICK(false, true)
        } else {
            MUCA_TEAM_PICK(false, false)
        }
#undef MUCA_TEAM_PICK
    }
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lay.bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, grid, block, lay.bytes, st, a, lay);
    e = hipGetLastError();
    if (e == hipSuccess && moms && a.nup > 0)
        e = launch_hist_from_moments(st, a.moms, a.N, a.joint_blocks, a.hist);
    return e;
}

template <int R, bool FP64, int KEYING>
hipError_t launch_team_r(hipStream_t st, const McsArgs &a, const TeamLayout &lay)
{
    const uint32_t rounds = (lay.row + 15) / 16;
    if (rounds <= 1) return launch_team_nr<R, FP64, KEYING, 1>(st, a, lay);
    if (rounds <= 2) return launch_team_nr<R, FP64, KEYING, 2>(st, a, lay);
    if (rounds <= 3) return launch_team_nr<R, FP64, KEYING, 3>(st, a, lay);
    return launch_team_nr<R, FP64, KEYING, 4>(st, a, lay);   // longer half-rows in chunks of 64
}

}  // namespace muca
