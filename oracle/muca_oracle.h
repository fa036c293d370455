/*
 * muca_oracle.h -- CPU restatement of the reference's multicanonical
 * Blume-Capel algorithm (null2p/MUCA).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libmuca, the CLIs, the
 * Python package) links or calls this code.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py load it, and there
 * only as the checker or as the timed CPU baseline.
 *
 * Every function cites the reference file:line it restates.  How each one is
 * pinned (known-answer vectors, closed forms, brute force) is listed next to
 * it; the test that pins it lives in tests/test_oracle.py.
 *
 * Conventions shared with the GPU engine:
 *   - a walker's lattice is N = L*L int8 spins in {-1,0,+1}, site i = y*L + x;
 *     the oracle stores one walker contiguously (s[i]);
 *   - histograms are uint64, D axis (D = sum s_i^2), N+1 bins;
 *   - "keying" selects the reference program whose Philox keys are used:
 *       ORC_KEYING_WEIGHT  = muca_delta_multigpu.cu (multigpu_weight)
 *       ORC_KEYING_PRODUCE = produce2_multigpu.cu   (produce_gpu)
 */
#ifndef MUCA_ORACLE_H
#define MUCA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_KEYING_WEIGHT = 0, ORC_KEYING_PRODUCE = 1 };

/* ---- R1: Philox4x32-R (philox.h:285-331, constants :232-251) ---------- */
void orc_philox4x32(unsigned rounds, const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t out[4]);

/* ---- R3: u01fixedpt<float> (uniform.hpp:195-205) ---------------------- */
float orc_u01fixedpt_f32(uint32_t r);

/* ---- R4: correctly rounded float log of u01fixedpt<float>(r).  The
 * reference calls the CUDA device logf (muca_delta_multigpu.cu:781,816);
 * that function is closed, so the engine defines the acceptance log as the
 * correctly rounded one.  This oracle computes it from glibc's double log
 * with a midpoint guard and falls back to 113-bit logq near a midpoint. */
float orc_logf_u01(uint32_t r);
/* exact reference: logq rounded to float (slow, ~1.5 us) */
float orc_logf_u01_quad(uint32_t r);
/* orc_logf_u01((k0+t) << 9) for t in [0, n) into out */
void orc_logf_u01_range(uint32_t k0, uint32_t n, float *out);
/* exhaustive check of orc_logf_u01 against orc_logf_u01_quad over all 2^23
 * distinct inputs, split over nthreads; returns the number of mismatches. */
long orc_logf_u01_exhaustive(int nthreads);

/* ---- R6: neighbour and sweep-order tables (muca_delta_multigpu.cu:157-175)
 * naver[4i..4i+3] = right, left, up, down; order[i] = visiting order. */
void orc_tables(uint32_t L, uint32_t *naver, uint32_t *order);

/* ---- R5: initial lattice of one walker.
 * weight keying  (muca_delta_multigpu.cu:340-351):
 *     key {10*w + 2*mm, 0xdecafbad}, ctr {2*gpu, seed, 0xBADCAB1E, 0xBADC0DED}, ++ctr[0]
 * produce keying (produce2_multigpu.cu:126-138):
 *     key {10*w, 0xdecafbad}, ctr {1000*gpu, seed, 0, 0xBADCAFE3}, ++ctr[2]
 * spin = (u < 1/3.0) + (u < 2/3.0) - 1 with u = u01fixedpt<float>. */
void orc_init_walker(int keying, uint32_t L, uint32_t walker, uint32_t mm, uint32_t gpu,
                     uint32_t seed, unsigned rounds, int8_t *s);
/* The engine's stratified start (no counterpart in the reference): the same
 * stream, spin = (u < p) ? (u < p/2 ? -1 : +1) : 0 with p = (walker + 1/2) /
 * strat_W; strat_W = 0 is orc_init_walker. */
void orc_init_walker_strat(int keying, uint32_t L, uint32_t walker, uint32_t mm, uint32_t gpu,
                           uint32_t seed, unsigned rounds, uint32_t strat_W, int8_t *s);

/* ---- R7: per-walker E_J, D, M (muca_delta_multigpu.cu:69-83,548-571;
 * produce2_multigpu.cu:66-89). */
void orc_state(uint32_t L, const uint32_t *naver, const int8_t *s, int32_t *E, int32_t *D,
               int32_t *M);

/* ---- R8 / R18: the per-walker Metropolis chain of the `mcs` kernels.
 *   weight keying, single precision (ref32): muca_delta_multigpu.cu:735-829
 *   produce keying, double precision:        produce2_multigpu.cu:401-496
 * beta: the reference hard-codes beta = 2 as the integer -dE*2; for any beta
 * the acceptance uses (float|double)(-dE*beta), identical at beta = 2.
 * hist (nullable): N+1 uint64 bins incremented at D after each production
 * attempt.  joint (nullable): (N+1)(N+2)/2 uint64 bins at (M+D*D+2*D)/2.
 * mom (nullable): 6 (N+1) uint64, the per-D moments of orc_moment_add.
 * lnw32 is read in ref32 mode, lnw64 in fp64 mode. */
typedef struct {
    uint32_t L;
    int keying;          /* ORC_KEYING_* */
    int fp64;            /* 0: float acceptance (weight path), 1: double (produce path) */
    unsigned rounds;     /* Philox rounds (7 in the reference) */
    double beta;
    uint32_t iter, seed;
    uint64_t therm, nupdates;
} orc_run_params;

void orc_mcs_walker(const orc_run_params *p, const uint32_t *naver, const uint32_t *order,
                    const float *lnw32, const double *lnw64, uint32_t walker, int8_t *s,
                    int32_t *E, int32_t *D, int32_t *M, uint64_t *hist, uint64_t *joint,
                    uint64_t *mom);
/* per-D moments after one production attempt at (D, M): count, sum |M|,
 * 128-bit sums of M^2 and M^4 as (low, high) words; six arrays of nb words */
void orc_moment_add(uint64_t *mom, uint64_t nb, int32_t d, int32_t m);

/* walkers [w0, w0+nw) with lattices at lat + (w-w0)*N; hist is shared and
 * summed; nthreads pthreads (0 = 1).  joint and mom must be NULL when nthreads > 1. */
void orc_mcs_batch(const orc_run_params *p, const float *lnw32, const double *lnw64,
                   uint32_t w0, uint32_t nw, int8_t *lat, int32_t *E, int32_t *D,
                   int32_t *M, uint64_t *hist, uint64_t *joint, uint64_t *mom, int nthreads);

/* ---- R10: hist_range (muca_delta_multigpu.cu:573-596) ---------------- */
void orc_hist_range(const uint64_t *hist, int N, int *start, int *end);
/* ---- R11: kull_leibler (muca_delta_multigpu.cu:598-618) -------------- */
double orc_kull_leibler(const uint64_t *hist, int N);
/* ---- R13: trivial update (muca_delta_multigpu.cu:466-476) ------------ */
void orc_update_trivial(int N, const uint64_t *hist, double *lnw);
/* ---- R14: weight_recursion (muca_delta_multigpu.cu:833-859) ---------- */
void orc_weight_recursion(int N, const uint64_t *hist, double *p_acc, int start, int end,
                          double *lnw);
/* normalisation lnW -= lnW[0] and the float copy (muca_delta_multigpu.cu:469-475) */
void orc_normalise(int N, double *lnw, float *lnw32);
/* ---- R17: jakknife (muca_delta_multigpu.cu:620-635) ------------------ */
void orc_jackknife(const double *x, unsigned J, double *avg, double *err);
/* ---- R19: weight_shift + boltz (produce2_multigpu.cu:391-400,297) ---- */
void orc_weight_shift(double beta, double delta, int N, double *lnw);
void orc_boltz(double beta, double delta, int N, const double *lnw, double *boltz);
/* ---- R20: observables (produce2_multigpu.cu:498-518), long double sums */
void orc_observables(double beta, unsigned N, const uint64_t *joint, const double *boltz,
                     double *absM, double *chi);
/* per-D moments of a joint histogram: S0 = sum_m H, S1 = sum_m |m| H,
 * S2 = sum_m m^2 H, S4 = sum_m m^4 H  (all as long double sums -> double) */
void orc_joint_moments(unsigned N, const uint64_t *joint, double *S0, double *S1,
                       double *S2, double *S4);

#ifdef __cplusplus
}
#endif
#endif
