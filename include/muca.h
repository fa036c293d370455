/*
 * muca.h -- C ABI of libmuca, the MI355X-native multicanonical Blume-Capel
 * engine.  Plain pointers and sizes only; no C++ or GPU types cross it.
 *
 * The reference (null2p/MUCA) has no library API: its boundary is two CUDA
 * programs, multigpu_weight (muca_delta_multigpu.cu) and produce_gpu
 * (produce2_multigpu.cu), driven by positional argv and text files.  Each
 * entry point below names the reference code it replaces; the drop-in CLIs
 * built on it keep the reference's argv and file formats.
 *
 * Ownership: a context owns all device memory it allocates; host arrays
 * belong to the caller and are only read/written during the call.
 * Threading: contexts are independent; one context must not be used from
 * two threads at once.  Errors: every call returns MUCA_OK (0) or a
 * negative MUCA_E_* code; muca_last_error() gives the calling thread's
 * last message.  Nothing in the library exits the process.
 */
#ifndef MUCA_H
#define MUCA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MUCA_OK 0
#define MUCA_E_ARG (-1)      /* invalid argument or size */
#define MUCA_E_HIP (-2)      /* HIP runtime error */
#define MUCA_E_NOMEM (-3)    /* device allocation failed */
#define MUCA_E_STATE (-4)    /* call out of order (e.g. no weights set) */
#define MUCA_E_IO (-5)       /* file could not be opened or parsed */
#define MUCA_E_RANGE (-6)    /* a count exceeded what its accumulator holds exactly */

/* acceptance precision */
#define MUCA_REF32 0   /* float weights and diff: multigpu_weight (muca_delta_multigpu.cu:775,810) */
#define MUCA_FP64 1    /* double weights and diff: produce_gpu (produce2_multigpu.cu:439,474) */

/* Philox keying of the update and of the initial lattice */
#define MUCA_KEYING_WEIGHT 0   /* muca_delta_multigpu.cu:340-343,740-742 */
#define MUCA_KEYING_PRODUCE 1  /* produce2_multigpu.cu:129-130,406-409 */

/* update-kernel variants (MUCA_KERNEL_AUTO: chosen from the walkers per
 * compute unit and the layout; muca_autotune chooses by measurement) */
#define MUCA_KERNEL_AUTO 0
#define MUCA_KERNEL_GENERIC 1  /* one lane per walker, reference site order from tables */
#define MUCA_KERNEL_TEAM 2     /* sixteen lanes per walker (L % 8 == 0, L <= 256; lattices and
                                  ln W in LDS when they fit) */
#define MUCA_KERNEL_LANE 3     /* one lane per walker, even L, table-free, screened acceptance */

/* what the produce path accumulates per walker block */
#define MUCA_OBS_JOINT 0    /* joint H(D,M), produce2_multigpu.cu:489 */
#define MUCA_OBS_MOMENTS 1  /* per D: count, sum |M|, sum M^2, sum M^4 (exact integers) */

typedef struct muca_ctx muca_ctx;

typedef struct muca_params {
    uint32_t L;          /* lattice side; N = L*L spins (argv[1]) */
    uint32_t walkers;    /* 0: maxThreadsPerCU * CUs, as muca_delta_multigpu.cu:100-106 */
    int32_t device;      /* HIP device ordinal (first_gpu + l) */
    uint32_t rounds;     /* Philox rounds: 7 (muca_delta_multigpu.cu:21) or 10 */
    double beta;         /* inverse temperature; 2.0 is the reference's -dE*2 (:775) */
    int32_t precision;   /* MUCA_REF32 | MUCA_FP64 */
    int32_t keying;      /* MUCA_KEYING_WEIGHT | MUCA_KEYING_PRODUCE */
    uint32_t joint_blocks; /* produce path: 0 = no joint H(D,M); B >= 1 = one joint
                              histogram per contiguous block of walkers (jackknife) */
    int32_t kernel;      /* MUCA_KERNEL_* */
    uint32_t block;      /* workgroup size of the one-lane kernels (0: chosen from
                            walkers and compute units; else 64..1024) */
    int32_t observables; /* produce path with joint_blocks > 0: MUCA_OBS_JOINT keeps
                            the reference's joint H(D,M); MUCA_OBS_MOMENTS keeps
                            exact per-D moments instead (6 (N+1) words per block;
                            MUCA_FP64, generic or team kernel) */
} muca_params;

typedef struct muca_info {
    uint32_t L, N, walkers, nbins;
    int32_t device, kernel;
    uint32_t cus, block;          /* compute units, threads per block of the update */
    uint64_t lattice_bytes, joint_bins;
    uint32_t lds_per_block;       /* bytes of LDS one workgroup may use */
    uint32_t lds_lattice;         /* 1 if the team kernel keeps lattices in LDS */
} muca_info;

const char *muca_last_error(void);
int muca_default_walkers(int32_t device, uint32_t *walkers);

/* Replaces the per-GPU setup of main() (muca_delta_multigpu.cu:93-205;
 * produce2_multigpu.cu:103-275): tables, buffers, constants. */
int muca_create(const muca_params *params, muca_ctx **ctx);
int muca_destroy(muca_ctx *ctx);
int muca_query(const muca_ctx *ctx, muca_info *info);

/* Initial random lattice (muca_delta_multigpu.cu:340-351 for mm = jack index;
 * produce2_multigpu.cu:126-138 ignores mm), generated on the device. */
int muca_init_lattice(muca_ctx *ctx, uint32_t mm, uint32_t gpu_index, uint32_t seed);
/* Stratified start (no counterpart in the reference; for runs seeded with
 * good weights, whose walkers would otherwise spend ~N^2 attempts diffusing
 * out of the random start's D ~ 2N/3): the same Philox stream, but walker w
 * of W has a nonzero spin with probability (w + 1/2)/W, sign - for
 * u < p/2, + for p/2 <= u < p, so the walkers start spread over D. */
int muca_init_lattice_stratified(muca_ctx *ctx, uint32_t mm, uint32_t gpu_index, uint32_t seed);
/* Lattice in the reference's layout s[site * walkers + walker]
 * (muca_delta_multigpu.cu:151,800). */
int muca_set_lattice(muca_ctx *ctx, const int8_t *site_major, uint64_t bytes);
int muca_get_lattice(muca_ctx *ctx, int8_t *site_major, uint64_t bytes);

/* computeEnergies/computeDeltaE (muca_delta_multigpu.cu:69-83),
 * computeEnergy/DeltaE/M (produce2_multigpu.cu:66-89). */
int muca_compute_state(muca_ctx *ctx);
int muca_get_state(muca_ctx *ctx, int32_t *E, int32_t *D, int32_t *M);
int muca_set_state(muca_ctx *ctx, const int32_t *E, const int32_t *D, const int32_t *M);

/* H2D of ln W (muca_delta_multigpu.cu:386; produce2_multigpu.cu:262).  The
 * float copy used in MUCA_REF32 mode is (float)lnw[i] (:473). */
int muca_set_log_weight(muca_ctx *ctx, const double *lnw, uint32_t nbins);

/* One launch of the update kernel `mcs` (muca_delta_multigpu.cu:393,735-829;
 * produce2_multigpu.cu:301,311,401-496): therm attempts without and
 * nupdates attempts with histogram accumulation, per walker.  The D
 * histogram (nbins uint64) is zeroed first and copied to hist_out when it is
 * non-NULL; kernel_ms (nullable) receives the update kernel's duration.
 * In MUCA_FP64 mode with joint_blocks > 0 the joint H(D,M) is zeroed and
 * accumulated too (read it with muca_get_joint), or with MUCA_OBS_MOMENTS
 * the per-D moments (muca_get_moments; nupdates < 2^32), whose S0 summed
 * over the blocks is then the D histogram. */
int muca_run(muca_ctx *ctx, uint32_t iter, uint32_t seed, uint64_t therm, uint64_t nupdates,
             uint64_t *hist_out, float *kernel_ms);
/* Autotuner (no counterpart in the reference, whose launch shape is fixed at
 * 256 threads per block, muca_delta_multigpu.cu:27): every eligible
 * (kernel, workgroup size) candidate runs `attempts` attempts per walker on a
 * scratch copy of the walkers; the fastest becomes the context's choice.
 * The walkers' lattice, E, D and M are left as they were; the last-launch
 * histogram is overwritten.  kernel/block/ms (nullable) report the choice. */
int muca_autotune(muca_ctx *ctx, uint64_t attempts, int32_t *kernel, uint32_t *block,
                  float *ms);
/* Test hook (no counterpart in the reference): the team and lane kernels
 * keep u32 histogram bins in LDS and carry 2^32 into the uint64 HBM bin when
 * one wraps (needed once a workgroup adds 2^32 counts to one bin).  With a
 * nonzero preload those LDS bins start at `preload` and the flush adds the
 * difference, so a test reaches the carry with a few counts per bin while
 * the histogram muca_run returns stays the exact count.  0 (the default)
 * is production behaviour. */
int muca_set_test_hist_preload(muca_ctx *ctx, uint32_t preload);
/* D2H of the last launch's D histogram (muca_delta_multigpu.cu:401-412);
 * copy_ms (nullable) receives the copy's duration */
int muca_get_hist(muca_ctx *ctx, uint64_t *hist, uint32_t nbins, float *copy_ms);
/* joint histogram of one walker block: (N+1)(N+2)/2 bins at (M+D*D+2*D)/2 */
int muca_get_joint(muca_ctx *ctx, uint32_t block, uint64_t *joint, uint64_t nbins);
/* Per-D moments of one walker block (MUCA_OBS_MOMENTS), 6 (N+1) uint64 as six
 * arrays over D: S0 = sum_M H(D,M), S1 = sum_M |M| H, S2 = sum_M M^2 H as
 * (low, high) words of a 128-bit integer, S4 = sum_M M^4 H likewise.  They
 * equal the moments of the joint histogram exactly; the reference reads the
 * same sums off H(D,M) in observables() (produce2_multigpu.cu:498-518).
 * Exact while every bin of the last launch holds fewer than 2^32 entries;
 * MUCA_E_RANGE otherwise (here and in muca_reweight).  In this mode muca_run
 * takes nupdates < 2^32. */
int muca_get_moments(muca_ctx *ctx, uint32_t block, uint64_t *moments, uint64_t nwords);

/* No counterpart in the reference: muca_reweight for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results as
 * muca_reweight, but no allocation of any kind: its scratch (FP64 per-D
 * moments, crystal fields, output planes, range flag) is a fixed 32 MiB
 * device array that each GPU the library runs on holds until the process
 * ends.  MUCA_E_RANGE, before any device work, when 8 * (4 * blocks * (N+1)
 * + 3 * ndelta * blocks + ndelta) bytes plus alignment exceed it.  All calls
 * on one GPU share that array (it keeps the last call's contents); calls
 * that overlap on one GPU, from several threads or contexts, race. */
int muca_reweight_noalloc(muca_ctx *ctx, const double *deltas, uint32_t ndelta, double *absM,
                          double *chi, double *binder);

/* ---- host-side weight iteration pieces (O(N), run between launches) ---- */
/* hist_range (muca_delta_multigpu.cu:573-596) */
int muca_hist_range(const uint64_t *hist, uint32_t N, int32_t *start, int32_t *end);
/* kull_leibler (muca_delta_multigpu.cu:598-618) */
double muca_kull_leibler(const uint64_t *hist, uint32_t N);
/* trivial update, lnW[i] -= log(H[i]) for H[i] > 1 (:467) */
int muca_update_trivial(const uint64_t *hist, uint32_t N, double *lnw);
/* weight_recursion (:833-859); p_acc persists across iterations of one run */
int muca_update_recursive(const uint64_t *hist, uint32_t N, int32_t start, int32_t end,
                          double *p_acc, double *lnw);
/* lnW -= lnW[0] over all bins (:469-475) */
int muca_normalise(uint32_t N, double *lnw);
/* jakknife (:620-635; produce2_multigpu.cu:520-535) */
int muca_jackknife(const double *x, uint32_t J, double *avg, double *err);

/* ---- produce path on the device ---- */
/* Reweighting of the joint histograms (or per-D moments) to canonical <|M|>, chi and the Binder
 * cumulant at ndelta crystal fields (produce2_multigpu.cu:293-358,391-400,
 * 498-518), in FP64 on the device with a max-shifted log-sum-exp.  Uses the
 * ln W last set with muca_set_log_weight.  Outputs are [ndelta][blocks]. */
int muca_reweight(muca_ctx *ctx, const double *deltas, uint32_t ndelta, double *absM,
                  double *chi, double *binder);

/* ---- batched weight iteration on the device ----
 * Many independent weight problems (each one run of multigpu_weight for one
 * jack on one GPU: muca_delta_multigpu.cu:314-476) advanced together: one
 * persistent update launch over every unconverged instance's walkers, then one
 * device launch doing the host loop body (hist_range, width, kull_leibler,
 * convergence, trivial or recursive ln W update in FP64, lnW -= lnW[0], the
 * NUPDATES schedule).  Per instance: weight keying, Philox4x32-7, therm = N
 * (therm_coeff = 1, :315-316), walker ids 0..walkers-1, the instance's seed,
 * start width 10 (:314).  The device update uses the device libm log; the
 * reference's host loop uses the host one. */
typedef struct muca_batch muca_batch;

/* Largest NUPDATES of a batch instance.  The reference keeps NUPDATES in a
 * 32-bit unsigned (muca_delta_multigpu.cu:89,379-382); the batched engine
 * keeps it in 64 bits, so the seeded protocol's c * N^2 updates per walker
 * (6.4e9 at L = 128, c = 24) are taken as given.  Below 2^32 the schedule
 * NUPDATES = (NUPDATES + 1) * 1.01 gives the reference's values. */
#define MUCA_MAX_NUPDATES (1ull << 48)

typedef struct muca_instance {
    uint32_t L;          /* even, >= 4 */
    uint32_t walkers;
    uint32_t seed;
    int32_t recursion;   /* 0 trivial update, 1 weight_recursion */
    uint64_t init_nup;   /* NUPDATES of the first iteration (0: 1); larger than
                            MUCA_MAX_NUPDATES: MUCA_E_RANGE from muca_batch_create */
    double beta;         /* 2.0 reproduces the reference's -dE*2 */
    double nup_growth;   /* once width == N each iteration sets NUPDATES =
                            (NUPDATES + 1) * nup_growth; 0 means 1.01, the
                            reference's schedule (:379-382); else >= 1 */
    uint64_t nup_max;    /* the schedule stops here (0: MUCA_MAX_NUPDATES) */
} muca_instance;

typedef struct muca_instance_status {
    uint32_t iter;       /* iterations done */
    uint32_t width;      /* widest histogram range seen (:432-433) */
    uint32_t converged;  /* width == N and Kullback-Leibler < 1e-4 (:447) */
    uint32_t reach_iter; /* iteration at which width reached N (0xFFFFFFFF: not yet) */
    uint32_t done_iter;  /* iteration that converged (0xFFFFFFFF: not yet) */
    int32_t start, end;  /* hist_range of the last iteration */
    uint32_t nup_capped; /* 1 once the schedule reached the instance's nup_max
                            (NUPDATES then stays there) */
    uint64_t nup;        /* NUPDATES of the next iteration */
    double dkl;          /* Kullback-Leibler divergence of the last iteration */
    uint64_t attempts;   /* flip attempts done (therm + nupdates per walker) */
} muca_instance_status;

const char *muca_batch_last_error(void);
/* tiles: MUCA_KERNEL_AUTO, MUCA_KERNEL_TEAM (every L % 8 == 0, L <= 128) or
 * MUCA_KERNEL_LANE; precision: MUCA_REF32 | MUCA_FP64.  Lane tiles hold
 * 256 walkers, or 512 / 1024 when the largest instance's ln W and histogram
 * leave LDS room for only a few workgroups per compute unit (the autotuner
 * may choose otherwise). */
int muca_batch_create(int32_t device, const muca_instance *instances, uint32_t n,
                      int32_t precision, int32_t tiles, muca_batch **batch);
int muca_batch_destroy(muca_batch *batch);
/* tile kind, threads per workgroup, tile count, persistent grid, total walkers */
int muca_batch_query(const muca_batch *batch, int32_t *tiles, uint32_t *block, uint32_t *ntiles,
                     uint32_t *grid, uint64_t *walkers);
/* random lattices (weight keying, jack index mm), E/D/M, ln W = 0, counters reset */
int muca_batch_init(muca_batch *batch, uint32_t mm);
/* the same with the stratified start of muca_init_lattice_stratified */
int muca_batch_init_stratified(muca_batch *batch, uint32_t mm);
int muca_batch_set_log_weight(muca_batch *batch, uint32_t instance, const double *lnw,
                              uint32_t nbins);
int muca_batch_get_log_weight(muca_batch *batch, uint32_t instance, double *lnw,
                              uint32_t nbins);
/* the D histogram of the instance's last iteration */
int muca_batch_get_hist(muca_batch *batch, uint32_t instance, uint64_t *hist, uint32_t nbins);
/* the instance's walkers in the reference's layout s[site * walkers + walker]
 * (bytes = N * walkers) and their E, D, M (walkers each; NULL skips one) --
 * what muca_get_lattice / muca_get_state give for one context */
int muca_batch_get_lattice(muca_batch *batch, uint32_t instance, int8_t *site_major,
                           uint64_t bytes);
int muca_batch_get_state(muca_batch *batch, uint32_t instance, int32_t *E, int32_t *D,
                         int32_t *M);
/* `iterations` weight iterations of every unconverged instance; use_graph
 * replays a captured hipGraph of up to 8 iterations (the remainder of
 * `iterations` runs eagerly; results are the same either way); gpu_ms
 * (nullable) receives the elapsed device time */
int muca_batch_iterate(muca_batch *batch, uint32_t iterations, int32_t use_graph,
                       float *gpu_ms);
/* the same iterations launched one by one with HIP events around each launch
 * on the batch's stream: summed durations of the persistent update kernel
 * and of the device weight update (either pointer may be NULL) */
int muca_batch_iterate_profiled(muca_batch *batch, uint32_t iterations, float *update_ms,
                                float *weight_ms);
/* Autotuner (no counterpart in the reference, whose launch shape is fixed at
 * 256 threads per block, muca_delta_multigpu.cu:27): times one iteration of
 * every eligible tile kind and workgroup size from freshly initialised walkers
 * (muca_batch_init(mm)), keeps the fastest and leaves the batch initialised
 * with mm, except that in a batch initialised before the call every
 * instance's ln W is what it was before the call (so weights set with
 * muca_batch_set_log_weight survive; call order create, init,
 * set_log_weight, autotune, iterate).  kind/block/ms (nullable) report
 * the choice. */
int muca_batch_autotune(muca_batch *batch, uint32_t mm, int32_t *kind, uint32_t *block,
                        float *ms);
/* (the autotuner re-initialises with the start of the last init call) */
int muca_batch_status(muca_batch *batch, muca_instance_status *status, uint32_t n);

/* ---- weight surrogate (no counterpart in the reference) ----
 * An MLP 4 -> 128 -> 128 -> 128 -> 1 (SiLU) predicting ln W(D)/N from four
 * features of (D/N, 1/L, beta, Delta); its output seeds the init-weight file
 * that multigpu_weight's `continuing` mode reads (muca_delta_multigpu.cu:223-235)
 * in place of the extrapolated initial weights.  FP32 master parameters; the
 * 128 x 128 products run on MFMA with BF16 operands and FP32 accumulation; the
 * input and output layers use BF16-rounded weights in FP32.  Parameters are
 * packed as W1[128][4], b1[128], W2[128][128], b2, W3[128][128], b3, w4[128],
 * b4 (W[o][k]: output o, input k). */
typedef struct muca_mlp muca_mlp;
const char *muca_mlp_last_error(void);
uint32_t muca_mlp_param_count(void);
int muca_mlp_create(int32_t device, muca_mlp **mlp);
int muca_mlp_destroy(muca_mlp *mlp);
/* also resets the Adam moments and step count */
int muca_mlp_set_params(muca_mlp *mlp, const float *params, uint32_t n);
int muca_mlp_get_params(muca_mlp *mlp, float *params, uint32_t n);
/* y[i] for inputs x[i][4] */
int muca_mlp_predict(muca_mlp *mlp, const float *x, uint32_t n, float *y);
/* gradient of the mean squared error over the n samples (no update);
 * grad (nullable) receives muca_mlp_param_count() values, loss (nullable) the MSE */
int muca_mlp_gradient(muca_mlp *mlp, const float *x, const float *t, uint32_t n, float *grad,
                      float *loss);
/* resident training set (the caller shuffles it); w (nullable) weights each
 * sample's squared error (loss = sum w (y - t)^2 / batch) */
int muca_mlp_set_data(muca_mlp *mlp, const float *x, const float *t, const float *w, uint32_t n);
/* `steps` Adam steps on consecutive minibatches of `batch` samples (a multiple
 * of 64), learning rate cosine-annealed from lr to lr_final; loss of the last
 * minibatch and the elapsed device time (both nullable) */
int muca_mlp_train(muca_mlp *mlp, uint32_t steps, uint32_t batch, float lr, float lr_final,
                   float *loss, float *gpu_ms);

/* ---- self-checks usable without a GPU ---- */
/* the acceptance log evaluated by the host instantiation of the device code */
float muca_acclog_host(uint32_t r);
/* host evaluation for r = k << 9, k in [k0, k0+n) */
int muca_acclog_host_range(uint32_t k0, uint32_t n, float *out);
/* the same function evaluated on the device for r = k << 9, k in [k0, k0+n) */
int muca_acclog_device(int32_t device, uint32_t k0, uint32_t n, float *out);
/* the fast screening approximation (hardware log2 * ln 2) the update kernels
 * use before deciding whether the exact log is needed, same inputs */
int muca_acclog_fast_device(int32_t device, uint32_t k0, uint32_t n, float *out);

#ifdef __cplusplus
}
#endif
#endif
