/*
 * muca_oracle.c -- CPU restatement of the reference's algorithm.
 *
 * TEST INFRASTRUCTURE ONLY (see muca_oracle.h).  The product never links it.
 *
 * Built with -ffp-contract=off so that every float/double expression rounds
 * in the order the reference's source states it.
 */
#include "muca_oracle.h"

#include <math.h>
#include <pthread.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------
 * R1  Philox4x32-R.  Restates the round and key schedule of the Philox
 * bijection (Salmon et al., SC'11) as instantiated in
 * Random123/include/Random123/philox.h:285-331 with the multipliers and Weyl
 * increments of :232-251.  Pinned by the Random123 known-answer vectors
 * (kat_vectors:24-29, old_kat_vectors) and by oracle/_ref/r123_probe.
 * --------------------------------------------------------------------- */
void orc_philox4x32(unsigned rounds, const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t out[4])
{
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (unsigned r = 0; r < rounds; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ x1 ^ k0;
        uint32_t n2 = hi0 ^ x3 ^ k1;
        x0 = n0;
        x1 = lo1;
        x2 = n2;
        x3 = lo0;
    }
    out[0] = x0;
    out[1] = x1;
    out[2] = x2;
    out[3] = x3;
}

/* R3  u01fixedpt<float>(uint32) = (1 | (r >> 8)) * 2^-24
 * (Random123/examples/uniform.hpp:195-205 with W=32, M=24). */
float orc_u01fixedpt_f32(uint32_t r)
{
    return (float)(1u | (r >> 8)) * 0x1p-24f;
}

/* ------------------------------------------------------------------------
 * R4  correctly rounded float log of u = u01fixedpt<float>(r).
 * glibc's log is accurate to well under 1 ulp of double.  If the double
 * result lies further than 2^-49 (relative) from every float rounding
 * boundary, rounding it to float gives the correctly rounded result;
 * otherwise the 113-bit logq decides.
 * --------------------------------------------------------------------- */
float orc_logf_u01_quad(uint32_t r)
{
    __float128 u = (__float128)(1u | (r >> 8)) / 16777216.0Q;
    return (float)logq(u);
}

float orc_logf_u01(uint32_t r)
{
    double u = (double)orc_u01fixedpt_f32(r);
    double d = log(u);
    float f = (float)d;
    /* both neighbouring rounding boundaries of f (midpoints to the adjacent
     * floats) are exact in double */
    double up = 0.5 * ((double)f + (double)nextafterf(f, INFINITY));
    double dn = 0.5 * ((double)f + (double)nextafterf(f, -INFINITY));
    double tol = fabs(d) * 0x1p-49;
    if (fabs(d - up) <= tol || fabs(d - dn) <= tol)
        return orc_logf_u01_quad(r);
    return f;
}

void orc_logf_u01_range(uint32_t k0, uint32_t n, float *out)
{
    for (uint32_t t = 0; t < n; t++) out[t] = orc_logf_u01((k0 + t) << 9);
}

typedef struct {
    uint32_t k0, k1;
    long bad;
} logchk_arg;

static void *logchk_worker(void *vp)
{
    logchk_arg *a = (logchk_arg *)vp;
    long bad = 0;
    for (uint32_t k = a->k0; k < a->k1; k++) {
        uint32_t r = k << 9; /* r >> 8 = 2k, so 1 | (r >> 8) = 2k+1 */
        float x = orc_logf_u01(r), y = orc_logf_u01_quad(r);
        if (memcmp(&x, &y, sizeof(float)) != 0) bad++;
    }
    a->bad = bad;
    return NULL;
}

long orc_logf_u01_exhaustive(int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_t th[64];
    logchk_arg args[64];
    const uint32_t total = 1u << 23;
    for (int t = 0; t < nthreads; t++) {
        args[t].k0 = (uint32_t)((uint64_t)total * t / nthreads);
        args[t].k1 = (uint32_t)((uint64_t)total * (t + 1) / nthreads);
        args[t].bad = 0;
        pthread_create(&th[t], NULL, logchk_worker, &args[t]);
    }
    long bad = 0;
    for (int t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        bad += args[t].bad;
    }
    return bad;
}

/* ------------------------------------------------------------------------
 * R6  tables (muca_delta_multigpu.cu:157-175).  The order formula is kept
 * literally, including its behaviour for odd L (it revisits sites).
 * --------------------------------------------------------------------- */
void orc_tables(uint32_t L, uint32_t *naver, uint32_t *order)
{
    uint32_t N = L * L;
    for (uint32_t i = 0; i < N; i++) {
        int right = (int)i + 1;
        int left = (int)i - 1;
        int up = (int)(i + L);
        int down = (int)i - (int)L;
        if (right % (int)L == 0) right -= (int)L;
        if (i % L == 0) left += (int)L;
        if (up > (int)N - 1) up -= (int)N;
        if (down < 0) down += (int)N;
        naver[4 * i + 0] = (uint32_t)right;
        naver[4 * i + 1] = (uint32_t)left;
        naver[4 * i + 2] = (uint32_t)up;
        naver[4 * i + 3] = (uint32_t)down;
        uint32_t half = N / 2;
        uint32_t blk = (i / half) % 2;
        order[i] = ((0 == blk) * (2 * i)) % N + ((1 == blk) * (2 * i + 1)) % N;
    }
}

/* R5  initial lattice (muca_delta_multigpu.cu:340-351; produce2_multigpu.cu:126-138) */
void orc_init_walker_strat(int keying, uint32_t L, uint32_t w, uint32_t mm, uint32_t gpu,
                           uint32_t seed, unsigned rounds, uint32_t strat_W, int8_t *s);

void orc_init_walker(int keying, uint32_t L, uint32_t w, uint32_t mm, uint32_t gpu,
                     uint32_t seed, unsigned rounds, int8_t *s)
{
    orc_init_walker_strat(keying, L, w, mm, gpu, seed, rounds, 0, s);
}

void orc_init_walker_strat(int keying, uint32_t L, uint32_t w, uint32_t mm, uint32_t gpu,
                           uint32_t seed, unsigned rounds, uint32_t strat_W, int8_t *s)
{
    uint32_t N = L * L;
    uint32_t key[2], c[4], r[4] = {0, 0, 0, 0};
    if (keying == ORC_KEYING_WEIGHT) {
        key[0] = 10u * w + mm * 2u;
        key[1] = 0xdecafbadu;
        c[0] = gpu * 2u;
        c[1] = seed;
        c[2] = 0xBADCAB1Eu;
        c[3] = 0xBADC0DEDu;
    } else {
        key[0] = w * 10u;
        key[1] = 0xdecafbadu;
        c[0] = gpu * 1000u;
        c[1] = seed;
        c[2] = 0u;
        c[3] = 0xBADCAFE3u;
    }
    for (uint32_t i = 0; i < N; i++) {
        if (i % 4 == 0) {
            if (keying == ORC_KEYING_WEIGHT)
                ++c[0];
            else
                ++c[2];
            orc_philox4x32(rounds, c, key, r);
        }
        float u = orc_u01fixedpt_f32(r[i % 4]);
        if (strat_W) {
            const double p = ((double)w + 0.5) / (double)strat_W;
            s[i] = (int8_t)((double)u < p ? ((double)u < 0.5 * p ? -1 : 1) : 0);
        } else {
            s[i] = (int8_t)(((double)u < 1 / 3.0) + ((double)u < 2 / 3.0) - 1);
        }
    }
}

/* R7  E_J = (sum_i -s_i * sum_nn s_j) >> 1, D = sum s^2, M = sum s
 * (muca_delta_multigpu.cu:548-571,76-83; produce2_multigpu.cu:82-89) */
void orc_state(uint32_t L, const uint32_t *naver, const int8_t *s, int32_t *E, int32_t *D,
               int32_t *M)
{
    uint32_t N = L * L;
    int32_t sum = 0, d = 0, m = 0;
    for (uint32_t i = 0; i < N; i++) {
        const uint32_t *nv = naver + 4 * i;
        sum += -s[i] * (s[nv[0]] + s[nv[1]] + s[nv[2]] + s[nv[3]]);
        d += s[i] * s[i];
        m += s[i];
    }
    *E = sum >> 1;
    *D = d;
    *M = m;
}

/* ---- per-D moments of the production attempts --------------------------
 * The quantities observables() (produce2_multigpu.cu:498-518) reads off the
 * joint histogram, kept per D instead: mom[0..5][D] (stride nb) = count,
 * sum |M|, sum M^2 (low, high 64-bit words), sum M^4 (low, high), exact in
 * unsigned 128-bit arithmetic.
 * --------------------------------------------------------------------- */
void orc_moment_add(uint64_t *mom, uint64_t nb, int32_t d, int32_t m)
{
    const unsigned __int128 am = (unsigned __int128)(m < 0 ? -(int64_t)m : (int64_t)m);
    const unsigned __int128 m2 = am * am, m4 = m2 * m2;
    unsigned __int128 s2 = ((unsigned __int128)mom[3 * nb + d] << 64) | mom[2 * nb + d];
    unsigned __int128 s4 = ((unsigned __int128)mom[5 * nb + d] << 64) | mom[4 * nb + d];
    s2 += m2;
    s4 += m4;
    mom[d] += 1;
    mom[nb + d] += (uint64_t)am;
    mom[2 * nb + d] = (uint64_t)s2;
    mom[3 * nb + d] = (uint64_t)(s2 >> 64);
    mom[4 * nb + d] = (uint64_t)s4;
    mom[5 * nb + d] = (uint64_t)(s4 >> 64);
}

/* ------------------------------------------------------------------------
 * R8 / R18  per-walker chain (muca_delta_multigpu.cu:735-829;
 * produce2_multigpu.cu:401-496).  Both phases share the counter: ctr[0] is
 * incremented before every block of four attempts, and the attempt index i
 * restarts at 0 for the production phase.
 * --------------------------------------------------------------------- */
void orc_mcs_walker(const orc_run_params *p, const uint32_t *naver, const uint32_t *order,
                    const float *lnw32, const double *lnw64, uint32_t w, int8_t *s,
                    int32_t *E, int32_t *D, int32_t *M, uint64_t *hist, uint64_t *joint,
                    uint64_t *mom)
{
    const uint64_t N = (uint64_t)p->L * p->L;
    uint32_t k1[2], k3[2];
    if (p->keying == ORC_KEYING_WEIGHT) {
        k1[0] = 2u * w;          /* muca_delta_multigpu.cu:740 */
        k1[1] = 0xdecafbadu;
        k3[0] = 2u * w + 50000u; /* :741 */
        k3[1] = 0x11cafeadu;
    } else {
        k1[0] = w;               /* produce2_multigpu.cu:406 */
        k1[1] = 0xdecafbadu;
        k3[0] = w + 40000u;      /* :408 */
        k3[1] = 0xdecafeadu;
    }
    uint32_t c[4] = {0u, p->seed, p->iter, 0xBADC0DEDu};
    uint32_t r1[4] = {0, 0, 0, 0}, r3[4] = {0, 0, 0, 0};
    int32_t e = *E, d = *D, m = *M;
    for (int phase = 0; phase < 2; phase++) {
        uint64_t n = phase ? p->nupdates : p->therm;
        for (uint64_t i = 0; i < n; i++) {
            if (i % 4 == 0) {
                ++c[0];
                orc_philox4x32(p->rounds, c, k1, r1);
                orc_philox4x32(p->rounds, c, k3, r3);
            }
            uint32_t idx = order[i % N];
            const uint32_t *nv = naver + 4 * (uint64_t)idx;
            int old_s = s[idx];
            int new_s = old_s + 2 * ((double)orc_u01fixedpt_f32(r3[i % 4]) < 0.5) - 1;
            new_s = new_s - 3 * (new_s == 2) + 3 * (new_s == -2);
            s[idx] = (int8_t)(new_s * (old_s == 0) + old_s);
            int dE = ((old_s == 0) + old_s * (new_s - old_s)) * (-s[idx]) *
                     (s[nv[0]] + s[nv[1]] + s[nv[2]] + s[nv[3]]);
            int dDel = new_s * new_s - old_s * old_s;
            int accept;
            if (!p->fp64) {
                float fde = (float)(-(double)dE * p->beta);
                float diff = fde + lnw32[d + dDel];
                diff = diff - lnw32[d];
                accept = (diff >= 0) || (orc_logf_u01(r1[i % 4]) < diff);
            } else {
                double diff = -(double)dE * p->beta;
                diff = diff + lnw64[d + dDel];
                diff = diff - lnw64[d];
                accept = (diff >= 0) || ((double)orc_logf_u01(r1[i % 4]) < diff);
            }
            if (accept) {
                s[idx] = (int8_t)new_s;
                e += dE;
                d += dDel;
                m += new_s - old_s;
            } else {
                s[idx] = (int8_t)old_s;
            }
            if (phase) {
                if (hist) hist[d]++;
                if (joint) joint[((int64_t)m + (int64_t)d * d + 2 * (int64_t)d) / 2]++;
                if (mom) orc_moment_add(mom, N + 1, d, m);
            }
        }
    }
    *E = e;
    *D = d;
    *M = m;
}

typedef struct {
    const orc_run_params *p;
    const uint32_t *naver, *order;
    const float *lnw32;
    const double *lnw64;
    uint32_t w0, wa, wb;
    int8_t *lat;
    int32_t *E, *D, *M;
    uint64_t *hist, *joint, *mom;
} batch_arg;

static void *batch_worker(void *vp)
{
    batch_arg *a = (batch_arg *)vp;
    uint64_t N = (uint64_t)a->p->L * a->p->L;
    for (uint32_t w = a->wa; w < a->wb; w++) {
        uint32_t j = w - a->w0;
        int32_t dummyM = 0;
        orc_mcs_walker(a->p, a->naver, a->order, a->lnw32, a->lnw64, w, a->lat + (uint64_t)j * N,
                       a->E + j, a->D + j, a->M ? a->M + j : &dummyM, a->hist, a->joint,
                       a->mom);
    }
    return NULL;
}

void orc_mcs_batch(const orc_run_params *p, const float *lnw32, const double *lnw64,
                   uint32_t w0, uint32_t nw, int8_t *lat, int32_t *E, int32_t *D,
                   int32_t *M, uint64_t *hist, uint64_t *joint, uint64_t *mom, int nthreads)
{
    uint64_t N = (uint64_t)p->L * p->L;
    uint32_t *naver = (uint32_t *)malloc(sizeof(uint32_t) * 4 * N);
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * N);
    orc_tables(p->L, naver, order);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((uint32_t)nthreads > nw) nthreads = nw ? (int)nw : 1;
    if (nthreads == 1) {
        batch_arg a = {p, naver, order, lnw32, lnw64, w0, w0, w0 + nw, lat, E, D, M, hist, joint,
                       mom};
        batch_worker(&a);
    } else {
        pthread_t th[256];
        batch_arg args[256];
        uint64_t **hs = (uint64_t **)calloc(nthreads, sizeof(uint64_t *));
        for (int t = 0; t < nthreads; t++) {
            hs[t] = hist ? (uint64_t *)calloc(N + 1, sizeof(uint64_t)) : NULL;
            uint32_t a0 = w0 + (uint32_t)((uint64_t)nw * t / nthreads);
            uint32_t a1 = w0 + (uint32_t)((uint64_t)nw * (t + 1) / nthreads);
            batch_arg a = {p, naver, order, lnw32, lnw64, w0, a0, a1, lat, E, D, M, hs[t], NULL,
                           NULL};
            args[t] = a;
            pthread_create(&th[t], NULL, batch_worker, &args[t]);
        }
        for (int t = 0; t < nthreads; t++) {
            pthread_join(th[t], NULL);
            if (hist) {
                for (uint64_t b = 0; b <= N; b++) hist[b] += hs[t][b];
                free(hs[t]);
            }
        }
        free(hs);
    }
    free(naver);
    free(order);
}

/* R10  hist_range (muca_delta_multigpu.cu:573-596): first/last bin with H > 1 */
void orc_hist_range(const uint64_t *hist, int N, int *start, int *end)
{
    int i;
    int st = N + 1, en = 0;
    for (i = 0; i < N + 1; i++) {
        if (hist[i] > 1) {
            if (i < st) st = i;
            break;
        }
    }
    for (i = N + 1; i-- > 0;) {
        if (hist[i] > 1) {
            if (i > en) en = i;
            break;
        }
    }
    if (st > en) {
        st = 0;
        en = 0;
    }
    *start = st;
    *end = en;
}

/* R11  kull_leibler (muca_delta_multigpu.cu:598-618) */
double orc_kull_leibler(const uint64_t *hist, int N)
{
    double num_bin = 0, num_event = 0, kld = 0, P, Q;
    for (int i = 0; i < N + 1; i++) {
        if (hist[i] > 0) {
            num_event = num_event + (double)hist[i];
            num_bin = num_bin + 1.0;
        }
    }
    for (int i = 0; i < N + 1; i++) {
        if (hist[i] > 0) {
            P = (double)hist[i] / num_event;
            Q = 1.0 / num_bin;
            kld = kld + P * log(P / Q);
        }
    }
    return kld;
}

/* R13  trivial update, line 467 of muca_delta_multigpu.cu */
void orc_update_trivial(int N, const uint64_t *hist, double *lnw)
{
    for (int i = 0; i <= N; i++)
        if (hist[i] > 1) lnw[i] = lnw[i] - log((double)hist[i]);
}

/* R14  weight_recursion (muca_delta_multigpu.cu:848-858), sequential prefix */
void orc_weight_recursion(int N, const uint64_t *hist, double *p_acc, int start, int end,
                          double *lnw)
{
    double *bf = (double *)malloc(sizeof(double) * (size_t)(N + 1));
    for (int i = 0; i <= N; i++) bf[i] = lnw[i];
    for (int i = start; i < end; i++) {
        if ((hist[i] != 0) & (hist[i + 1] != 0)) {
            double p = (double)(hist[i] * hist[i + 1]) / (double)(hist[i] + hist[i + 1]);
            double k = p / (p + p_acc[i]);
            double dS = -bf[i] + bf[i + 1] + k * log((double)hist[i] / (double)hist[i + 1]);
            lnw[i + 1] = lnw[i] + dS;
            p_acc[i] = p_acc[i] + p;
        }
    }
    free(bf);
}

/* muca_delta_multigpu.cu:469-475 */
void orc_normalise(int N, double *lnw, float *lnw32)
{
    double tmp = lnw[0];
    for (int i = 0; i <= N; i++) {
        lnw[i] = lnw[i] - tmp;
        if (lnw32) lnw32[i] = (float)lnw[i];
    }
}

/* R17  jakknife (muca_delta_multigpu.cu:620-635) */
void orc_jackknife(const double *x, unsigned J, double *avg_out, double *err_out)
{
    double avg = 0;
    for (unsigned i = 0; i < J; i++) avg += x[i];
    double *Mi = (double *)malloc(sizeof(double) * (J ? J : 1));
    for (unsigned i = 0; i < J; i++) Mi[i] = (avg - x[i]) / (double)(J - 1);
    avg = avg / (double)J;
    double dev = 0;
    for (unsigned i = 0; i < J; i++) dev += (avg - Mi[i]) * (avg - Mi[i]);
    dev = (J - 1) * dev / (double)J;
    dev = sqrt(dev);
    free(Mi);
    *avg_out = avg;
    *err_out = dev;
}

/* R19  weight_shift (produce2_multigpu.cu:391-400) and boltz (:297) */
void orc_weight_shift(double beta, double delta, int N, double *lnw)
{
    double mx = 0;
    for (int i = 0; i < N + 1; i++) {
        double v = -lnw[i] - beta * delta * (unsigned)i;
        if (i == 0 || v > mx) mx = v;
    }
    for (int i = 0; i < N + 1; i++) lnw[i] = lnw[i] + mx;
}

void orc_boltz(double beta, double delta, int N, const double *lnw, double *boltz)
{
    for (int j = 0; j < N + 1; j++) boltz[j] = exp(-beta * delta * (unsigned)j - lnw[j]);
}

/* R20  observables (produce2_multigpu.cu:498-518) with 64-bit bin indices */
void orc_observables(double beta, unsigned N, const uint64_t *joint, const double *boltz,
                     double *absM, double *chi)
{
    long double sum = 0, M1 = 0, M2 = 0;
    for (uint64_t i = 0; i <= N; i++) {
        int64_t m;
        for (m = -(int64_t)i; m <= 0; m += 2) {
            uint64_t h = joint[(i * i + 2 * i + (uint64_t)m) / 2];
            sum = sum + (double)h * boltz[i];
            M1 = M1 + (double)((uint64_t)(-m) * h) * boltz[i];
            M2 = M2 + (double)((uint64_t)(m * m) * h) * boltz[i];
        }
        for (; m <= (int64_t)i; m += 2) {
            uint64_t h = joint[(i * i + 2 * i + (uint64_t)m) / 2];
            sum = sum + (double)h * boltz[i];
            M1 = M1 + (double)((uint64_t)m * h) * boltz[i];
            M2 = M2 + (double)((uint64_t)(m * m) * h) * boltz[i];
        }
    }
    *absM = (double)(M1 / sum);
    *chi = (double)((M2 / sum - (M1 / sum) * (M1 / sum)) * beta / (long double)N);
}

void orc_joint_moments(unsigned N, const uint64_t *joint, double *S0, double *S1, double *S2,
                       double *S4)
{
    for (uint64_t i = 0; i <= N; i++) {
        long double a0 = 0, a1 = 0, a2 = 0, a4 = 0;
        for (int64_t m = -(int64_t)i; m <= (int64_t)i; m += 2) {
            long double h = (long double)joint[(i * i + 2 * i + (uint64_t)m) / 2];
            long double am = (long double)(m < 0 ? -m : m);
            a0 += h;
            a1 += am * h;
            a2 += am * am * h;
            a4 += am * am * am * am * h;
        }
        S0[i] = (double)a0;
        S1[i] = (double)a1;
        S2[i] = (double)a2;
        S4[i] = (double)a4;
    }
}
