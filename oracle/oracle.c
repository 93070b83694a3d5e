/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY (see oracle.h).  Plain C + OpenMP.
 *
 * CPU restatement of the BUSseq Gibbs sampler of the reference
 * (BUSseq_gpu.cu; cited as cu:LINE).  The update rules are the reference's;
 * the scan order is the sampler's own, a rotation of the reference's
 * (cu:1267-1398) in which the subtype draw opens the sweep:
 *
 *   W (categorical full conditional) -> pi -> X/Z -> alpha -> gamma0 -> gamma1
 *   -> L -> p, tau0 (from iteration n_unchanged on) -> beta -> nu -> delta -> phi
 *
 * alpha and gamma commute (their full conditionals share no parameter), so
 * their order is immaterial; the W step is an exact Gibbs draw where the
 * reference runs Metropolis-Hastings with a uniform proposal (cu:820-854),
 * which leaves the same posterior.  Every long sum is an exact fixed-point
 * sum (bq_fix), so results are independent of thread count and of the order
 * in which cells are visited.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/busseq/model.h"

#define H BQ_HIST
#define ORC_KMAX 64   /* largest K, as the sampler's */

struct orc {
    int B, G, K, N;
    int* n_b; int* boff; int* batch; int* first;
    uint32_t seed, chain;
    int prec;
    int n_iter, n_burnin, n_unchanged;
    int iter;
    int nthr;

    int32_t* S; int32_t* W; int32_t* L;
    double *alpha, *beta, *nu, *delta, *phi, *gamma, *pi;
    double p, tau0;
    double *mu_alpha, *mu_nu, *mu_delta;
    int64_t* hist_s;   /* B*H static histogram of Y>0 entries with Y<H */

    /* precision-T copies: ea[G], eb[K*G], em[K*G], bt[K*G], en[B*G], ed[N], ph[B*G] */
    void *ea, *eb, *em, *bt, *en, *ed, *ph;

    /* proposals of the current sweep */
    double *alpha_p, *beta_p, *nu_p, *delta_p, *phi_p, *gamma_p;
    int *alpha_acc, *beta_acc;
    /* debug sums of the last sweep */
    double *alpha_sum, *beta_sum, *nu_sum, *delta_sum, *phi_sum, *gamma_sum, *w_logit;

    /* per-sweep exact accumulators */
    uint64_t *acc_alpha, *acc_beta, *acc_nu0, *acc_nu1, *acc_phi;   /* G, K*G, B*K*G, B*K*G, B*G */
    int64_t *hist_d, *n1, *sx1;                                     /* B*H, B, B */
    uint64_t *o0, *o1r, *o1a;                                       /* B each */
    int* cnt;                                                       /* B*K */
    int TR;                                                         /* rows per tile */
    int* perm; int* tiles; int ntiles;                              /* sorted rows, tiles (b,k,start,count) */
    uint64_t* tl;  size_t tl_stride;                                /* per-thread scratch */

    /* posterior accumulators */
    int n_rec;
    double *s_gamma, *s_alpha, *s_beta, *s_nu, *s_delta, *s_phi, *s_pi, s_p, s_tau0;
    int64_t* c_L; int32_t* c_W;

    /* traces of the recorded sweeps, the reference's layout [param * n_pres + iter]
     * (cu:892-936); allocated by orc_keep_traces */
    int n_pres;
    double *tr_gamma, *tr_alpha, *tr_beta, *tr_ptau, *tr_nu, *tr_delta, *tr_phi, *tr_pi;
    int32_t *tr_L, *tr_W;
};

/* Planted mistakes for the tests that pin the oracle (0 = none): each flips
 * one rule of the reference to a plausible wrong one, so that a test can show
 * that its independent check fails on it.  Never set outside tests. */
static int g_mut = 0;
void orc_set_mutation(int m) { g_mut = m; }

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

static int tid(void)
{
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

static void* zal(size_t n) { void* p = calloc(n ? n : 1, 1); return p; }

/* ------------------------------------------------------------------------ */
/* creation / teardown                                                        */
/* ------------------------------------------------------------------------ */
orc_t* orc_create(int B, const int* n_b, int G, int K, const int32_t* Y,
                  uint32_t seed, uint32_t chain, int precision,
                  int n_iter, int n_burnin, int n_unchanged)
{
    if (B < 1 || G < 1 || K < 1 || K > ORC_KMAX) return NULL;
    orc_t* o = (orc_t*)zal(sizeof(orc_t));
    o->B = B; o->G = G; o->K = K;
    o->seed = seed; o->chain = chain; o->prec = precision == 32 ? 32 : 64;
    o->n_iter = n_iter; o->n_burnin = n_burnin; o->n_unchanged = n_unchanged;
    o->n_b = (int*)zal(sizeof(int) * B);
    o->boff = (int*)zal(sizeof(int) * (B + 1));
    int N = 0;
    for (int b = 0; b < B; ++b) { o->n_b[b] = n_b[b]; o->boff[b] = N; N += n_b[b]; }
    o->boff[B] = N; o->N = N;
    o->batch = (int*)zal(sizeof(int) * N);
    o->first = (int*)zal(sizeof(int) * N);
    for (int b = 0; b < B; ++b) {
        for (int i = o->boff[b]; i < o->boff[b + 1]; ++i) o->batch[i] = b;
        if (n_b[b] > 0) o->first[o->boff[b]] = 1;
    }
    size_t NG = (size_t)N * G;
    o->S = (int32_t*)zal(4 * NG);
    #pragma omp parallel for schedule(static)
    for (size_t e = 0; e < NG; ++e) o->S[e] = Y[e] > 0 ? Y[e] : 0;
    o->W = (int32_t*)zal(4 * N);
    o->L = (int32_t*)zal(4 * (size_t)K * G);
    o->alpha = (double*)zal(8 * G);
    o->beta = (double*)zal(8 * (size_t)K * G);
    o->nu = (double*)zal(8 * (size_t)B * G);
    o->delta = (double*)zal(8 * N);
    o->phi = (double*)zal(8 * (size_t)B * G);
    o->gamma = (double*)zal(8 * 2 * B);
    o->pi = (double*)zal(8 * (size_t)B * K);
    o->mu_alpha = (double*)zal(8 * G);
    o->mu_nu = (double*)zal(8 * B);
    o->mu_delta = (double*)zal(8 * N);
    o->hist_s = (int64_t*)zal(8 * (size_t)B * H);
    /* per batch: its cells' rows in parallel, integer counts added (exact in any order) */
    for (int b = 0; b < B; ++b) {
        int64_t* hb = o->hist_s + (size_t)b * H;
        #pragma omp parallel
        {
            int64_t* loc = (int64_t*)zal(8 * (size_t)H);
            #pragma omp for schedule(static)
            for (int i = o->boff[b]; i < o->boff[b + 1]; ++i)
                for (int g = 0; g < G; ++g) {
                    int32_t y = Y[(size_t)i * G + g];
                    if (y > 0 && y < H) loc[y] += 1;
                }
            #pragma omp critical
            for (int x = 0; x < H; ++x) hb[x] += loc[x];
            free(loc);
        }
    }
    size_t ts = o->prec == 32 ? 4 : 8;
    o->ea = zal(ts * G); o->eb = zal(ts * K * G); o->em = zal(ts * K * G);
    o->bt = zal(ts * K * G); o->en = zal(ts * B * G); o->ed = zal(ts * N);
    o->ph = zal(ts * B * G);
    o->alpha_p = (double*)zal(8 * G); o->beta_p = (double*)zal(8 * (size_t)K * G);
    o->nu_p = (double*)zal(8 * (size_t)B * G); o->delta_p = (double*)zal(8 * N);
    o->phi_p = (double*)zal(8 * (size_t)B * G); o->gamma_p = (double*)zal(8 * 2 * B);
    o->alpha_acc = (int*)zal(4 * G); o->beta_acc = (int*)zal(4 * (size_t)K * G);
    o->alpha_sum = (double*)zal(8 * G); o->beta_sum = (double*)zal(8 * (size_t)K * G);
    o->nu_sum = (double*)zal(8 * (size_t)B * G); o->delta_sum = (double*)zal(8 * N);
    o->phi_sum = (double*)zal(8 * (size_t)B * G); o->gamma_sum = (double*)zal(8 * 2 * B);
    o->w_logit = (double*)zal(8 * (size_t)N * K);
    o->acc_alpha = (uint64_t*)zal(8 * G);
    o->acc_beta = (uint64_t*)zal(8 * (size_t)K * G);
    o->acc_nu0 = (uint64_t*)zal(8 * (size_t)B * K * G);
    o->acc_nu1 = (uint64_t*)zal(8 * (size_t)B * K * G);
    o->acc_phi = (uint64_t*)zal(8 * (size_t)B * G);
    o->hist_d = (int64_t*)zal(8 * (size_t)B * H);
    o->n1 = (int64_t*)zal(8 * B); o->sx1 = (int64_t*)zal(8 * B);
    o->o0 = (uint64_t*)zal(8 * B); o->o1r = (uint64_t*)zal(8 * B); o->o1a = (uint64_t*)zal(8 * B);
    o->cnt = (int*)zal(4 * (size_t)B * K);
    o->TR = 64;
    o->perm = (int*)zal(4 * (size_t)N);
    o->tiles = (int*)zal(4 * 4 * ((size_t)N + (size_t)B * K));
#ifdef _OPENMP
    o->nthr = omp_get_max_threads();
#else
    o->nthr = 1;
#endif
    /* per thread: alpha G | beta K*G | nu0 B*K*G | nu1 B*K*G | phi B*G | hist B*H | 6*B | 3*G tile sums */
    o->tl_stride = (size_t)G + (size_t)K * G + 2 * (size_t)B * K * G + (size_t)B * G
                 + (size_t)B * H + 6 * (size_t)B + 3 * (size_t)G;
    o->tl = (uint64_t*)zal(8 * o->tl_stride * o->nthr);
    o->s_gamma = (double*)zal(8 * 2 * B); o->s_alpha = (double*)zal(8 * G);
    o->s_beta = (double*)zal(8 * (size_t)K * G); o->s_nu = (double*)zal(8 * (size_t)B * G);
    o->s_delta = (double*)zal(8 * N); o->s_phi = (double*)zal(8 * (size_t)B * G);
    o->s_pi = (double*)zal(8 * (size_t)B * K);
    o->c_L = (int64_t*)zal(8 * (size_t)K * G); o->c_W = (int32_t*)zal(4 * (size_t)N * K);
    return o;
}

void orc_destroy(orc_t* o)
{
    if (!o) return;
    void* ps[] = { o->n_b, o->boff, o->batch, o->first, o->S, o->W, o->L, o->alpha, o->beta, o->nu,
                   o->delta, o->phi, o->gamma, o->pi, o->mu_alpha, o->mu_nu, o->mu_delta, o->hist_s,
                   o->ea, o->eb, o->em, o->bt, o->en, o->ed, o->ph, o->alpha_p, o->beta_p, o->nu_p,
                   o->delta_p, o->phi_p, o->gamma_p, o->alpha_acc, o->beta_acc, o->alpha_sum,
                   o->beta_sum, o->nu_sum, o->delta_sum, o->phi_sum, o->gamma_sum, o->w_logit,
                   o->acc_alpha, o->acc_beta, o->acc_nu0, o->acc_nu1, o->acc_phi, o->hist_d, o->n1,
                   o->sx1, o->o0, o->o1r, o->o1a, o->cnt, o->perm, o->tiles, o->tl, o->s_gamma, o->s_alpha, o->s_beta,
                   o->s_nu, o->s_delta, o->s_phi, o->s_pi, o->c_L, o->c_W,
                   o->tr_gamma, o->tr_alpha, o->tr_beta, o->tr_ptau, o->tr_nu, o->tr_delta,
                   o->tr_phi, o->tr_pi, o->tr_L, o->tr_W };
    for (size_t j = 0; j < sizeof(ps) / sizeof(ps[0]); ++j) free(ps[j]);
    free(o);
}

int orc_iter(const orc_t* o) { return o->iter; }
void orc_set_tile_rows(orc_t* o, int tr) { if (tr > 0) o->TR = tr; }

/* ------------------------------------------------------------------------ */
/* precision-T copies of the parameters                                       */
/* ------------------------------------------------------------------------ */
#define SETT(arr, idx, v) do { if (o->prec == 32) ((float*)o->arr)[idx] = (float)(v); \
                               else ((double*)o->arr)[idx] = (v); } while (0)

static void refresh_alpha(orc_t* o, int g) { SETT(ea, g, bq_exp(o->alpha[g])); }
static void refresh_beta(orc_t* o, int k, int g)
{
    size_t j = (size_t)k * o->G + g;
    SETT(eb, j, bq_exp(o->beta[j]));
    if (o->prec == 32) ((float*)o->em)[j] = bq_tw_em_f32(o->beta[j]);   /* clamped for the subtype term */
    else ((double*)o->em)[j] = bq_expm1(o->beta[j]);
    SETT(bt, j, o->beta[j]);
}
static void refresh_nu(orc_t* o, int b, int g) { SETT(en, (size_t)b * o->G + g, bq_exp(o->nu[(size_t)b * o->G + g])); }
static void refresh_delta(orc_t* o, int i) { SETT(ed, i, bq_exp(o->delta[i])); }
static void refresh_phi(orc_t* o, int b, int g) { SETT(ph, (size_t)b * o->G + g, o->phi[(size_t)b * o->G + g]); }

static void refresh_all(orc_t* o)
{
    for (int g = 0; g < o->G; ++g) refresh_alpha(o, g);
    for (int k = 0; k < o->K; ++k) for (int g = 0; g < o->G; ++g) refresh_beta(o, k, g);
    for (int b = 0; b < o->B; ++b) for (int g = 0; g < o->G; ++g) { refresh_nu(o, b, g); refresh_phi(o, b, g); }
    for (int i = 0; i < o->N; ++i) refresh_delta(o, i);
}

/* ------------------------------------------------------------------------ */
/* initialisation (cu:1163-1244)                                              */
/* ------------------------------------------------------------------------ */
int orc_init(orc_t* o)
{
    const int B = o->B, G = o->G, K = o->K, N = o->N;
    int64_t* lib = (int64_t*)zal(8 * N);
    /* library sizes s_i = sum_g Y_ig (cu:1204-1205) */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        int64_t s = 0;
        for (int g = 0; g < G; ++g) { int32_t v = o->S[(size_t)i * G + g]; s += v > 0 ? v : 0; }
        lib[i] = s;
    }
    for (int i = 0; i < N; ++i)
        if (lib[i] <= 0) { free(lib); return -1; }
    /* pi0_bk = (k+1) / (K(K+1)/2)  (cu:392-397) */
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < K; ++k) o->pi[b * K + k] = (k + 1) / (K * (K + 1) / 2.0);
    /* W0 ~ Cat(pi0_b), delta0 = log s_i - log s_first(b)  (cu:419-443) */
    for (int i = 0; i < N; ++i) {
        int b = o->batch[i];
        bq_rng r; bq_rng_init(&r, o->seed, o->chain, (uint32_t)i, 0u, BQ_SWEEP_INIT, BQ_PH_INIT_W);
        double u = bq_uniform(&r);
        int k = 0; double tmp = o->pi[b * K];
        while (u > tmp && k < K - 1) { k++; tmp += o->pi[b * K + k]; }
        o->W[i] = k;
        int ref = g_mut == 9 ? o->boff[b + 1] - 1 : o->boff[b];   /* planted: last cell of the batch */
        o->delta[i] = bq_log((double)lib[i]) - bq_log((double)lib[ref]);
    }
    /* raw means r_bkg = mean_{i in b, W_i=k} log(1 + Y_ig / e^delta_i)  (cu:369-382, 1210-1221) */
    uint64_t* rs = (uint64_t*)zal(8 * (size_t)B * K * G);
    int* nbk = (int*)zal(4 * (size_t)B * K);
    for (int i = 0; i < N; ++i) nbk[o->batch[i] * K + o->W[i]] += 1;
    /* exact fixed-point sums: per-thread copies added with integer adds (order-free) */
    #pragma omp parallel
    {
        uint64_t* loc = (uint64_t*)zal(8 * (size_t)B * K * G);
        #pragma omp for schedule(static)
        for (int i = 0; i < N; ++i) {
            int b = o->batch[i], k = o->W[i];
            double ie = bq_exp(-o->delta[i]);
            uint64_t* row = loc + ((size_t)b * K + k) * G;
            for (int g = 0; g < G; ++g) {
                int32_t y = o->S[(size_t)i * G + g];
                row[g] += bq_fix(bq_log1p((double)(y > 0 ? y : 0) * ie));
            }
        }
        #pragma omp critical
        for (size_t j = 0; j < (size_t)B * K * G; ++j) rs[j] += loc[j];
        free(loc);
    }
    double* rm = (double*)zal(8 * (size_t)B * K * G);
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < K; ++k)
            for (int g = 0; g < G; ++g) {
                size_t j = ((size_t)b * K + k) * G + g;
                int den = g_mut == 8 ? o->n_b[b] : nbk[b * K + k];   /* planted: batch size */
                rm[j] = nbk[b * K + k] == 0 ? 0.0 : bq_unfix(rs[j]) / den;
            }
    /* mu_nu_b = mean over batch b of log(1+Y/e^delta) minus batch 0's (cu:470-482, 1232-1241) */
    double* bm = (double*)zal(8 * B);
    for (int b = 0; b < B; ++b) {
        unsigned __int128 tot = 0;
        for (int k = 0; k < K; ++k)
            for (int g = 0; g < G; ++g) tot += (unsigned __int128)(__int128)(int64_t)rs[((size_t)b * K + k) * G + g];
        bm[b] = (double)(__int128)tot * BQ_FIX_INV / ((double)o->n_b[b] * G);
    }
    o->mu_nu[0] = 0.0;
    for (int b = 1; b < B; ++b) o->mu_nu[b] = g_mut == 3 ? bm[0] - bm[b] : bm[b] - bm[0];   /* 3: planted sign */
    /* gamma0 = (0, -0.3)  (cu:384-390) */
    for (int b = 0; b < B; ++b) { o->gamma[b] = BQ_GAMMA0_INIT; o->gamma[B + b] = BQ_GAMMA1_INIT; }
    /* phi0 = 5, nu0_bg = (sum_k r_bkg - sum_k r_0kg)/K  (cu:399-417) */
    for (int b = 0; b < B; ++b)
        for (int g = 0; g < G; ++g) {
            o->phi[(size_t)b * G + g] = BQ_PHI_INIT;
            if (b == 0) { o->nu[g] = 0.0; continue; }
            double t1 = rm[((size_t)b * K) * G + g], t2 = rm[g];
            for (int k = 1; k < K; ++k) { t1 += rm[((size_t)b * K + k) * G + g]; t2 += rm[(size_t)k * G + g]; }
            o->nu[(size_t)b * G + g] = (t1 - t2) / K;
        }
    /* p0 ~ U(0, 0.5), tau0 = 0.005  (cu:1172-1175; the host rand() is replaced by Philox) */
    {
        bq_rng r; bq_rng_init(&r, o->seed, o->chain, 0u, 0u, BQ_SWEEP_INIT, BQ_PH_INIT_P);
        o->p = bq_uniform(&r) * BQ_P_INIT_BOUND;
        o->tau0 = BQ_TAU0_INIT;
    }
    /* alpha0 = r_00g, beta0_kg = r_0kg - r_00g, L0 ~ Bern  (cu:445-462, 1186, 1227) */
    double lrb = bq_log(o->p / (1.0 - o->p)) + (bq_log(BQ_TAU0_INIT) - bq_log(BQ_TAU1)) / 2.0;
    for (int g = 0; g < G; ++g) { o->alpha[g] = rm[g]; o->mu_alpha[g] = rm[g]; o->beta[g] = 0.0; o->L[g] = 0; }
    for (int k = 1; k < K; ++k)
        for (int g = 0; g < G; ++g) {
            size_t j = (size_t)k * G + g;
            o->beta[j] = rm[j] - rm[g];
            double lr = lrb + (o->beta[j] * o->beta[j]) * (1.0 / BQ_TAU0_INIT - 1.0 / BQ_TAU1) / 2.0;
            bq_rng r; bq_rng_init(&r, o->seed, o->chain, (uint32_t)g, (uint32_t)k, BQ_SWEEP_INIT, BQ_PH_INIT_L);
            o->L[j] = bq_bern_logodds(bq_uniform(&r), lr);
        }
    for (int i = 0; i < N; ++i) o->mu_delta[i] = o->delta[i];
    refresh_all(o);
    o->iter = 0;
    o->n_rec = 0;
    free(lib); free(rs); free(nbk); free(rm); free(bm);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* per-element passes, generated for T = float and T = double                 */
/* ------------------------------------------------------------------------ */
#define LANES 64
#define CHUNK 1024
#define TL_ALPHA(tl) (tl)
#define TL_BETA(o, tl) ((tl) + (o)->G)
#define TL_NU0(o, tl) ((tl) + (o)->G + (size_t)(o)->K * (o)->G)
#define TL_NU1(o, tl) (TL_NU0(o, tl) + (size_t)(o)->B * (o)->K * (o)->G)
#define TL_PHI(o, tl) (TL_NU1(o, tl) + (size_t)(o)->B * (o)->K * (o)->G)
#define TL_HIST(o, tl) (TL_PHI(o, tl) + (size_t)(o)->B * (o)->G)
#define TL_SMALL(o, tl) (TL_HIST(o, tl) + (size_t)(o)->B * H)
#define TL_PART(o, tl) ((double*)(TL_SMALL(o, tl) + 6 * (size_t)(o)->B))

/*
 * Canonical reduction structure (shared with the gfx950 kernels):
 *  - a sum over the genes of one cell is split into chunks of 1024 genes
 *    and, within a chunk, into 64 gene classes (g mod 64); each (chunk,
 *    class) group is summed in double precision in increasing g and the
 *    group sums are added exactly in fixed point;
 *  - a sum over cells for one gene runs over tiles: cells sorted stably by
 *    (batch, subtype, index) and cut into tiles of at most TR cells within a
 *    (batch, subtype) group; each tile is summed in double precision in that
 *    order and the tile sums are added exactly in fixed point.
 */
static void build_tiles(orc_t* o)
{
    const int B = o->B, K = o->K, TR = o->TR;
    int* off = (int*)zal(4 * ((size_t)B * K + 1));
    int* cur = (int*)zal(4 * ((size_t)B * K + 1));
    for (int j = 0; j < B * K; ++j) off[j + 1] = off[j] + o->cnt[j];
    memcpy(cur, off, 4 * ((size_t)B * K + 1));
    for (int i = 0; i < o->N; ++i) o->perm[cur[o->batch[i] * K + o->W[i]]++] = i;
    int nt = 0;
    for (int j = 0; j < B * K; ++j)
        for (int st = off[j]; st < off[j + 1]; st += TR) {
            int n = off[j + 1] - st; if (n > TR) n = TR;
            o->tiles[4 * nt] = j / K; o->tiles[4 * nt + 1] = j % K;
            o->tiles[4 * nt + 2] = st; o->tiles[4 * nt + 3] = n;
            nt++;
        }
    o->ntiles = nt;
    free(off); free(cur);
}

/*
 * X/Z update of one entry with Y = 0 (cu:484-512).  s is the encoded state:
 * 0 = (X=0, Z=0), ~x = (X=x, Z=1).  If X = 0, Z ~ Bern(logistic(gamma0)),
 * i.e. Z = 1 when u > 1/(1+e^gamma0) = qz; if Z = 1 an NB(mu, phi)
 * independence proposal x' is accepted when
 *   log u <= log(1+e^{-g0-g1 x}) - log(1+e^{-g0-g1 x'})   (log form of cu:504);
 * if Z = 0 then X = 0.  Returns the new encoded state.
 */
#define DEFINE_PASSES(T, SFX)                                                              \
/* subtype draw: categorical full conditional of W_i (replaces MH cu:820-854) */           \
static void pass_w_##SFX(orc_t* o, int t)                                                  \
{                                                                                          \
    const int G = o->G, K = o->K;                                                          \
    const T* ea = (const T*)o->ea; const T* en = (const T*)o->en; const T* ed = (const T*)o->ed; \
    const T* ph = (const T*)o->ph; const T* bt = (const T*)o->bt; const T* em = (const T*)o->em; \
    _Pragma("omp parallel for schedule(dynamic, 4)")                                       \
    for (int i = 0; i < o->N; ++i) {                                                       \
        int b = o->batch[i];                                                               \
        uint64_t acc[ORC_KMAX];                                                            \
        for (int k = 0; k < K; ++k) acc[k] = 0;                                            \
        for (int c0 = 0; c0 < G; c0 += CHUNK) {                                            \
            T part[ORC_KMAX][LANES];                                                       \
            for (int k = 0; k < K; ++k) for (int l = 0; l < LANES; ++l) part[k][l] = (T)0; \
            int c1 = c0 + CHUNK < G ? c0 + CHUNK : G;                                      \
            for (int g = c0; g < c1; ++g) {                                                \
                int32_t s = o->S[(size_t)i * G + g];                                       \
                int xi = s < 0 ? ~s : s;                                                   \
                T x = (T)xi;                                                               \
                size_t bg = (size_t)b * G + g;                                             \
                T phi = ph[bg];                                                            \
                T mu0 = bq_mu_##SFX(ea[g], (T)1, en[bg], ed[i]);                           \
                T r0 = bq_ratio_##SFX(mu0, phi);                                           \
                for (int k = 1; k < K; ++k) {                                              \
                    size_t kg = (size_t)k * G + g;                                         \
                    part[k][g & (LANES - 1)] += bq_tw_##SFX(x, phi, r0, bt[kg], em[kg]);   \
                }                                                                          \
            }                                                                              \
            for (int k = 1; k < K; ++k) for (int l = 0; l < LANES; ++l) acc[k] += bq_fix((double)part[k][l]); \
        }                                                                                  \
        double lp[ORC_KMAX], mx = -BQ_INF;                                                 \
        for (int k = 0; k < K; ++k) {                                                      \
            double lk = bq_unfix(acc[k]);                                                  \
            o->w_logit[(size_t)i * K + k] = lk;                                            \
            lp[k] = lk + bq_log(o->pi[b * K + k]);                                         \
            if (lp[k] > mx) mx = lp[k];                                                    \
        }                                                                                  \
        double tot = 0.0;                                                                  \
        for (int k = 0; k < K; ++k) { lp[k] = bq_exp(lp[k] - mx); tot += lp[k]; }          \
        bq_rng r; bq_rng_init(&r, o->seed, o->chain, (uint32_t)i, 0u, (uint32_t)t, BQ_PH_W); \
        double tgt = bq_uniform(&r) * tot, cum = 0.0;                                      \
        int kk = K - 1;                                                                    \
        for (int k = 0; k < K; ++k) { cum += lp[k]; if (tgt <= cum) { kk = k; break; } }   \
        o->W[i] = kk;                                                                      \
    }                                                                                      \
}                                                                                          \
/* X/Z imputation (cu:484-512) + alpha sums (cu:608-623) + gamma histogram (cu:524-576) */ \
static void pass_xz_alpha_##SFX(orc_t* o, int t)                                           \
{                                                                                          \
    const int B = o->B, G = o->G;                                                          \
    const T* ea = (const T*)o->ea; const T* eb = (const T*)o->eb; const T* en = (const T*)o->en; \
    const T* ed = (const T*)o->ed; const T* ph = (const T*)o->ph;                          \
    T* da = (T*)malloc(sizeof(T) * G); T* ca = (T*)malloc(sizeof(T) * G);                  \
    for (int g = 0; g < G; ++g) {                                                          \
        double dl = o->alpha_p[g] - o->alpha[g];                                           \
        da[g] = (T)dl; ca[g] = (T)bq_expm1(dl);                                            \
    }                                                                                      \
    double* qz = (double*)malloc(8 * B);                                                   \
    for (int b = 0; b < B; ++b) qz[b] = 1.0 / (1.0 + bq_exp(o->gamma[b]));                 \
    memset(o->tl, 0, 8 * o->tl_stride * o->nthr);                                          \
    _Pragma("omp parallel for schedule(dynamic, 1)")                                       \
    for (int tt = 0; tt < o->ntiles; ++tt) {                                               \
        uint64_t* tl = o->tl + o->tl_stride * (size_t)tid();                               \
        uint64_t* aa = TL_ALPHA(tl);                                                       \
        int64_t* hd = (int64_t*)TL_HIST(o, tl);                                            \
        uint64_t* sm = TL_SMALL(o, tl);                                                    \
        double* pa = TL_PART(o, tl);                                                       \
        int b = o->tiles[4 * tt], k = o->tiles[4 * tt + 1];                                \
        int st = o->tiles[4 * tt + 2], nr = o->tiles[4 * tt + 3];                          \
        double g0 = o->gamma[b], g1 = o->gamma[B + b];                                     \
        double g0p = o->gamma_p[b], g1p = o->gamma_p[B + b];                               \
        for (int g = 0; g < G; ++g) pa[g] = 0.0;                                           \
        for (int rr = 0; rr < nr; ++rr) {                                                  \
            int i = o->perm[st + rr];                                                      \
            for (int g = 0; g < G; ++g) {                                                  \
                size_t e = (size_t)i * G + g, bg = (size_t)b * G + g;                      \
                int32_t s = o->S[e];                                                       \
                T mu = bq_mu_##SFX(ea[g], eb[(size_t)k * G + g], en[bg], ed[i]);           \
                int xi;                                                                    \
                if (s <= 0) {                                                              \
                    bq_rng r; bq_rng_init(&r, o->seed, o->chain, (uint32_t)i, (uint32_t)g, (uint32_t)t, BQ_PH_XZ); \
                    int32_t sn = bq_xz_update(s, (double)mu, o->phi[bg], g0, g1, qz[b], &r);  \
                    o->S[e] = sn;                                                          \
                    xi = sn < 0 ? ~sn : 0;                                                 \
                    if (sn < 0) { sm[0 * B + b] += 1; sm[1 * B + b] += (uint64_t)xi; }    \
                    if (xi < H) hd[(size_t)b * H + xi] += 1;                               \
                } else xi = s;                                                             \
                if (xi >= H) {                                                             \
                    double xd = (double)xi;                                                \
                    double a0 = bq_sp_lin(g0, g1, xd), a1 = bq_sp_lin(g0p, g1, xd);        \
                    double a2 = bq_sp_lin(g0, g1p, xd), a3 = bq_sp_lin(g0p, g1p, xd);      \
                    sm[2 * B + b] += bq_fix(a0 - a1);                                      \
                    sm[3 * B + b] += bq_fix(a0 - a2);                                      \
                    sm[4 * B + b] += bq_fix(a1 - a3);                                      \
                }                                                                          \
                T phi = ph[bg];                                                            \
                T rt = bq_ratio_##SFX(mu, phi);                                            \
                pa[g] += (double)bq_tmh_##SFX((T)xi, phi, rt, da[g], ca[g]);               \
            }                                                                              \
        }                                                                                  \
        for (int g = 0; g < G; ++g) aa[g] += bq_fix(pa[g]);                                \
    }                                                                                      \
    memset(o->acc_alpha, 0, 8 * G); memset(o->hist_d, 0, 8 * (size_t)B * H);               \
    for (int b = 0; b < B; ++b) { o->n1[b] = 0; o->sx1[b] = 0; o->o0[b] = 0; o->o1r[b] = 0; o->o1a[b] = 0; } \
    for (int th = 0; th < o->nthr; ++th) {                                                 \
        uint64_t* tl = o->tl + o->tl_stride * (size_t)th;                                  \
        for (int g = 0; g < G; ++g) o->acc_alpha[g] += tl[g];                              \
        int64_t* hd = (int64_t*)TL_HIST(o, tl);                                            \
        for (size_t j = 0; j < (size_t)B * H; ++j) o->hist_d[j] += hd[j];                  \
        uint64_t* sm = TL_SMALL(o, tl);                                                    \
        for (int b = 0; b < B; ++b) {                                                      \
            o->n1[b] += (int64_t)sm[b]; o->sx1[b] += (int64_t)sm[B + b];                   \
            o->o0[b] += sm[2 * B + b]; o->o1r[b] += sm[3 * B + b]; o->o1a[b] += sm[4 * B + b]; \
        }                                                                                  \
    }                                                                                      \
    free(da); free(ca); free(qz);                                                          \
}                                                                                          \
/* beta (cu:676-694) and nu (cu:715-728) sums; nu under both beta outcomes */              \
static void pass_beta_nu_##SFX(orc_t* o)                                                   \
{                                                                                          \
    const int B = o->B, G = o->G, K = o->K;                                                \
    const T* ea = (const T*)o->ea; const T* eb = (const T*)o->eb; const T* en = (const T*)o->en; \
    const T* ed = (const T*)o->ed; const T* ph = (const T*)o->ph;                          \
    T* db = (T*)malloc(sizeof(T) * K * G); T* cb = (T*)malloc(sizeof(T) * K * G);          \
    T* dn = (T*)malloc(sizeof(T) * B * G); T* cn = (T*)malloc(sizeof(T) * B * G);          \
    for (size_t j = 0; j < (size_t)K * G; ++j) {                                           \
        double dl = o->beta_p[j] - o->beta[j]; db[j] = (T)dl; cb[j] = (T)bq_expm1(dl);     \
    }                                                                                      \
    for (size_t j = 0; j < (size_t)B * G; ++j) {                                           \
        double dl = o->nu_p[j] - o->nu[j]; dn[j] = (T)dl; cn[j] = (T)bq_expm1(dl);         \
    }                                                                                      \
    memset(o->tl, 0, 8 * o->tl_stride * o->nthr);                                          \
    _Pragma("omp parallel for schedule(dynamic, 1)")                                       \
    for (int tt = 0; tt < o->ntiles; ++tt) {                                               \
        uint64_t* tl = o->tl + o->tl_stride * (size_t)tid();                               \
        uint64_t* ab = TL_BETA(o, tl); uint64_t* an0 = TL_NU0(o, tl); uint64_t* an1 = TL_NU1(o, tl); \
        double* pb = TL_PART(o, tl); double* pn0 = pb + G; double* pn1 = pn0 + G;         \
        int b = o->tiles[4 * tt], k = o->tiles[4 * tt + 1];                                \
        int st = o->tiles[4 * tt + 2], nr = o->tiles[4 * tt + 3];                          \
        for (int g = 0; g < G; ++g) { pb[g] = 0.0; pn0[g] = 0.0; pn1[g] = 0.0; }           \
        for (int rr = 0; rr < nr; ++rr) {                                                  \
            int i = o->perm[st + rr];                                                      \
            for (int g = 0; g < G; ++g) {                                                  \
                int32_t s = o->S[(size_t)i * G + g];                                       \
                int xi = s < 0 ? ~s : s;                                                   \
                T x = (T)xi;                                                               \
                size_t bg = (size_t)b * G + g, kg = (size_t)k * G + g;                     \
                T phi = ph[bg];                                                            \
                T mu = bq_mu_##SFX(ea[g], eb[kg], en[bg], ed[i]);                          \
                T rt = bq_ratio_##SFX(mu, phi);                                            \
                T lb = (T)0;                                                               \
                if (k >= 1) {                                                              \
                    lb = LOG1P_##SFX(rt * cb[kg]);                                         \
                    pb[g] += (double)bq_tmh_l_##SFX(x, phi, db[kg], lb);                   \
                }                                                                          \
                if (b >= 1) {                                                              \
                    pn0[g] += (double)bq_tmh_##SFX(x, phi, rt, dn[bg], cn[bg]);            \
                    if (k >= 1) {                                                          \
                        T cbn = (cb[kg] + cn[bg]) + cb[kg] * cn[bg];                       \
                        pn1[g] += (double)bq_tmh2_##SFX(x, phi, rt, dn[bg], cbn, lb);      \
                    }                                                                      \
                }                                                                          \
            }                                                                              \
        }                                                                                  \
        for (int g = 0; g < G; ++g) {                                                      \
            size_t bkg = ((size_t)b * K + k) * G + g;                                      \
            if (k >= 1) ab[(size_t)k * G + g] += bq_fix(pb[g]);                            \
            if (b >= 1) { an0[bkg] += bq_fix(pn0[g]); if (k >= 1) an1[bkg] += bq_fix(pn1[g]); } \
        }                                                                                  \
    }                                                                                      \
    memset(o->acc_beta, 0, 8 * (size_t)K * G);                                             \
    memset(o->acc_nu0, 0, 8 * (size_t)B * K * G); memset(o->acc_nu1, 0, 8 * (size_t)B * K * G); \
    for (int th = 0; th < o->nthr; ++th) {                                                 \
        uint64_t* tl = o->tl + o->tl_stride * (size_t)th;                                  \
        uint64_t* ab = TL_BETA(o, tl); uint64_t* an0 = TL_NU0(o, tl); uint64_t* an1 = TL_NU1(o, tl); \
        for (size_t j = 0; j < (size_t)K * G; ++j) o->acc_beta[j] += ab[j];                \
        for (size_t j = 0; j < (size_t)B * K * G; ++j) { o->acc_nu0[j] += an0[j]; o->acc_nu1[j] += an1[j]; } \
    }                                                                                      \
    free(db); free(cb); free(dn); free(cn);                                                \
}                                                                                          \
/* delta: proposal, row sum (cu:746-761) and decision (cu:763-778) per cell */             \
static void pass_delta_##SFX(orc_t* o, int t)                                              \
{                                                                                          \
    const int G = o->G;                                                                    \
    const T* ea = (const T*)o->ea; const T* eb = (const T*)o->eb; const T* en = (const T*)o->en; \
    T* ed = (T*)o->ed; const T* ph = (const T*)o->ph;                                      \
    _Pragma("omp parallel for schedule(dynamic, 4)")                                       \
    for (int i = 0; i < o->N; ++i) {                                                       \
        if (o->first[i]) { o->delta_p[i] = o->delta[i]; o->delta_sum[i] = 0.0; continue; } \
        int b = o->batch[i], k = o->W[i];                                                  \
        bq_rng r; bq_rng_init(&r, o->seed, o->chain, (uint32_t)i, 0u, (uint32_t)t, BQ_PH_DELTA_P); \
        double dp = o->delta[i] + BQ_RW_SD * bq_normal(&r);                                \
        double dl = dp - o->delta[i];                                                      \
        T dd = (T)dl, cd = (T)bq_expm1(dl);                                                \
        uint64_t acc = 0;                                                                  \
        for (int c0 = 0; c0 < G; c0 += CHUNK) {                                            \
            double part[LANES];                                                            \
            for (int l = 0; l < LANES; ++l) part[l] = 0.0;                                 \
            int c1 = c0 + CHUNK < G ? c0 + CHUNK : G;                                      \
            for (int g = c0; g < c1; ++g) {                                                \
                int32_t s = o->S[(size_t)i * G + g];                                       \
                int xi = s < 0 ? ~s : s;                                                   \
                size_t bg = (size_t)b * G + g;                                             \
                T phi = ph[bg];                                                            \
                T mu = bq_mu_##SFX(ea[g], eb[(size_t)k * G + g], en[bg], ed[i]);           \
                T rt = bq_ratio_##SFX(mu, phi);                                            \
                part[g & (LANES - 1)] += (double)bq_tmh_##SFX((T)xi, phi, rt, dd, cd);     \
            }                                                                              \
            for (int l = 0; l < LANES; ++l) acc += bq_fix(part[l]);                        \
        }                                                                                  \
        double sum = bq_unfix(acc);                                                        \
        double lr = sum + bq_lr_normal(o->delta[i], dp, o->mu_delta[i], BQ_SIGMA2_DELTA);  \
        bq_rng ra; bq_rng_init(&ra, o->seed, o->chain, (uint32_t)i, 0u, (uint32_t)t, BQ_PH_DELTA_A); \
        o->delta_p[i] = dp; o->delta_sum[i] = sum;                                         \
        if (bq_log(bq_uniform(&ra)) <= lr) { o->delta[i] = dp; ed[i] = (T)bq_exp(dp); }    \
    }                                                                                      \
}                                                                                          \
/* phi sums (cu:789-804) with the new delta: the NB parts summed per tile, each   \
 * count's lgamma part added in fixed point on its own */                                  \
static void pass_phi_##SFX(orc_t* o)                                                       \
{                                                                                          \
    const int B = o->B, G = o->G;                                                          \
    const T* ea = (const T*)o->ea; const T* eb = (const T*)o->eb; const T* en = (const T*)o->en; \
    const T* ed = (const T*)o->ed; const T* ph = (const T*)o->ph;                          \
    size_t BG = (size_t)B * G;                                                             \
    T* cst = (T*)malloc(sizeof(T) * 6 * BG);                                               \
    for (size_t j = 0; j < BG; ++j) {                                                      \
        double f = o->phi[j], fp = o->phi_p[j];                                            \
        if (!(fp > 0.0)) fp = f;                                                           \
        double d = fp - f;                                                                 \
        cst[j] = (T)fp; cst[BG + j] = (T)d; cst[2 * BG + j] = (T)(d / f);                  \
        cst[3 * BG + j] = (T)(1.0 / fp); cst[4 * BG + j] = (T)bq_log1p(d / f);             \
        cst[5 * BG + j] = (T)bq_lgamma_diff64(f, d);                                       \
    }                                                                                      \
    memset(o->tl, 0, 8 * o->tl_stride * o->nthr);                                          \
    _Pragma("omp parallel for schedule(dynamic, 1)")                                       \
    for (int tt = 0; tt < o->ntiles; ++tt) {                                               \
        uint64_t* tl = o->tl + o->tl_stride * (size_t)tid();                               \
        uint64_t* ap = TL_PHI(o, tl);                                                      \
        double* pp = TL_PART(o, tl);                                                       \
        int b = o->tiles[4 * tt], k = o->tiles[4 * tt + 1];                                \
        int st = o->tiles[4 * tt + 2], nr = o->tiles[4 * tt + 3];                          \
        for (int g = 0; g < G; ++g) pp[g] = 0.0;                                           \
        for (int rr = 0; rr < nr; ++rr) {                                                  \
            int i = o->perm[st + rr];                                                      \
            for (int g = 0; g < G; ++g) {                                                  \
                int32_t s = o->S[(size_t)i * G + g];                                       \
                int xi = s < 0 ? ~s : s;                                                   \
                size_t bg = (size_t)b * G + g;                                             \
                T mu = bq_mu_##SFX(ea[g], eb[(size_t)k * G + g], en[bg], ed[i]);           \
                pp[g] += (double)bq_tphi_nb_##SFX((T)xi, ph[bg], cst[bg], mu, cst[BG + bg], \
                          cst[2 * BG + bg], cst[3 * BG + bg], cst[4 * BG + bg]);            \
                /* each count's lgamma part, exactly, on its own */                        \
                if (xi > 0)                                                                \
                    ap[bg] += bq_fix((double)bq_tphi_lg_##SFX((T)xi, ph[bg], cst[BG + bg],   \
                                                              cst[5 * BG + bg]));          \
            }                                                                              \
        }                                                                                  \
        for (int g = 0; g < G; ++g) ap[(size_t)b * G + g] += bq_fix(pp[g]);                \
    }                                                                                      \
    memset(o->acc_phi, 0, 8 * BG);                                                         \
    for (int th = 0; th < o->nthr; ++th) {                                                 \
        uint64_t* ap = TL_PHI(o, o->tl + o->tl_stride * (size_t)th);                       \
        for (size_t j = 0; j < BG; ++j) o->acc_phi[j] += ap[j];                            \
    }                                                                                      \
    free(cst);                                                                             \
}

#define LOG1P_f32 bq_log1p_mh_f32
#define LOG1P_f64 bq_log1p
DEFINE_PASSES(float, f32)
DEFINE_PASSES(double, f64)

/* ------------------------------------------------------------------------ */
/* small per-parameter steps                                                  */
/* ------------------------------------------------------------------------ */
static double ph_uniform(const orc_t* o, uint32_t a, uint32_t b, int t, int phase)
{
    bq_rng r; bq_rng_init(&r, o->seed, o->chain, a, b, (uint32_t)t, (uint32_t)phase);
    return bq_uniform(&r);
}

/* pi_b ~ Dirichlet(n_b. + xi) (cu:856-890, with the normalisation race Q1 fixed) */
static void update_pi(orc_t* o, int t)
{
    const int B = o->B, K = o->K;
    memset(o->cnt, 0, 4 * (size_t)B * K);
    for (int i = 0; i < o->N; ++i) o->cnt[o->batch[i] * K + o->W[i]] += 1;
    for (int b = 0; b < B; ++b) {
        bq_rng r; bq_rng_init(&r, o->seed, o->chain, (uint32_t)b, 0u, (uint32_t)t, BQ_PH_PI);
        double gk[ORC_KMAX], tot = 0.0;
        const double xi = g_mut == 1 ? BQ_XI_PI - 1.0 : BQ_XI_PI;   /* 1: planted prior */
        for (int k = 0; k < K; ++k) { gk[k] = bq_rgamma(&r, o->cnt[b * K + k] + xi, 1.0); tot += gk[k]; }
        for (int k = 0; k < K; ++k) o->pi[b * K + k] = gk[k] / tot;
    }
}

/* alpha and gamma proposals (cu:514-522, 602-606) */
static void propose_alpha_gamma(orc_t* o, int t)
{
    const int B = o->B;
    for (int g = 0; g < o->G; ++g) {
        bq_rng r; bq_rng_init(&r, o->seed, o->chain, (uint32_t)g, 0u, (uint32_t)t, BQ_PH_ALPHA_P);
        o->alpha_p[g] = o->alpha[g] + BQ_RW_SD * bq_normal(&r);
    }
    for (int b = 0; b < B; ++b) {
        bq_rng r0; bq_rng_init(&r0, o->seed, o->chain, (uint32_t)b, 0u, (uint32_t)t, BQ_PH_GAMMA0_P);
        o->gamma_p[b] = o->gamma[b] + BQ_RW_SD * bq_normal(&r0);
        bq_rng r1; bq_rng_init(&r1, o->seed, o->chain, (uint32_t)b, 0u, (uint32_t)t, BQ_PH_GAMMA1_P);
        o->gamma_p[B + b] = -bq_rgamma(&r1, -10.0 * o->gamma[B + b], 0.1);
    }
}

/* alpha decision (cu:625-632) */
static void decide_alpha(orc_t* o, int t)
{
    for (int g = 0; g < o->G; ++g) {
        double sum = bq_unfix(o->acc_alpha[g]);
        o->alpha_sum[g] = sum;
        double lr = sum + bq_lr_normal(o->alpha[g], o->alpha_p[g], o->mu_alpha[g], BQ_SIGMA2_ALPHA);
        int acc = bq_log(ph_uniform(o, (uint32_t)g, 0u, t, BQ_PH_ALPHA_A)) <= lr;
        o->alpha_acc[g] = acc;
        if (acc) { o->alpha[g] = o->alpha_p[g]; refresh_alpha(o, g); }
    }
}

/* gamma0 then gamma1 decisions from the X histogram (cu:524-600) */
static void decide_gamma(orc_t* o, int t)
{
    const int B = o->B;
    for (int b = 0; b < B; ++b) {
        const int64_t* hs = o->hist_s + (size_t)b * H;
        const int64_t* hd = o->hist_d + (size_t)b * H;
        double g0 = o->gamma[b], g1 = o->gamma[B + b];
        double g0p = o->gamma_p[b], g1p = o->gamma_p[B + b];
        uint64_t a0 = bq_fix((g0p - g0) * (double)o->n1[b]);
        for (int x = 0; x < H; ++x) {
            int64_t n = hs[x] + hd[x];
            if (n) a0 += bq_fix((double)n * (bq_sp_lin(g0, g1, (double)x) - bq_sp_lin(g0p, g1, (double)x)));
        }
        a0 += o->o0[b];
        double s0 = bq_unfix(a0);
        double lr0 = s0 + bq_lr_normal(g0, g0p, 0.0, BQ_SIGMA2_GAMMA0);
        int acc0 = bq_log(ph_uniform(o, (uint32_t)b, 0u, t, BQ_PH_GAMMA0_A)) <= lr0;
        if (acc0) o->gamma[b] = g0p;
        double g0n = o->gamma[b];
        uint64_t a1 = bq_fix((g1p - g1) * (double)o->sx1[b]);
        for (int x = 0; x < H; ++x) {
            int64_t n = hs[x] + hd[x];
            if (n) a1 += bq_fix((double)n * (bq_sp_lin(g0n, g1, (double)x) - bq_sp_lin(g0n, g1p, (double)x)));
        }
        a1 += acc0 ? o->o1a[b] : o->o1r[b];
        double s1 = bq_unfix(a1);
        double u1 = ph_uniform(o, (uint32_t)b, 0u, t, BQ_PH_GAMMA1_A);
        o->gamma_sum[b] = s0; o->gamma_sum[B + b] = s1;
        if (g1p < 0.0) {
            double lr1 = s1 + bq_lr_gamma1(-g1, -g1p);
            if (bq_log(u1) <= lr1) o->gamma[B + b] = g1p;
        }
    }
}

/* L Gibbs (cu:634-648) */
static void update_L(orc_t* o, int t)
{
    const int G = o->G, K = o->K;
    for (int k = 1; k < K; ++k)
        for (int g = 0; g < G; ++g) {
            size_t j = (size_t)k * G + g;
            double lo = bq_logodds_L(o->p, o->tau0, o->beta[j]);
            o->L[j] = bq_bern_logodds(ph_uniform(o, (uint32_t)g, (uint32_t)k, t, BQ_PH_L), lo);
        }
}

/* p, tau0 Gibbs (cu:650-668) from the current L and beta */
static void update_ptau(orc_t* o, int t)
{
    const int G = o->G, K = o->K;
    int64_t cnt = 0; uint64_t sb = 0;
    for (int k = 1; k < K; ++k)
        for (int g = 0; g < G; ++g) {
            size_t j = (size_t)k * G + g;
            if (o->L[j]) cnt += 1; else sb += bq_fix(o->beta[j] * o->beta[j]);
        }
    double n = (double)G * (K - 1);
    bq_rng r; bq_rng_init(&r, o->seed, o->chain, 0u, 0u, (uint32_t)t, BQ_PH_PTAU);
    const double b_p = g_mut == 2 ? 1.0 : BQ_B_P;   /* 2: planted prior */
    o->p = bq_rbeta(&r, (double)cnt + BQ_A_P, n - (double)cnt + b_p);
    double a = BQ_A_TAU0 + (n - (double)cnt) / (g_mut == 4 ? 1.0 : 2.0);   /* 4: planted shape */
    double bb = BQ_B_TAU0 + bq_unfix(sb) / 2.0;
    o->tau0 = 1.0 / bq_rgamma(&r, a, 1.0 / bb);
}

/* beta and nu random-walk proposals (cu:670-674, 709-713) */
static void propose_beta_nu(orc_t* o, int t)
{
    const int B = o->B, G = o->G, K = o->K;
    for (int g = 0; g < G; ++g) { o->beta_p[g] = o->beta[g]; o->nu_p[g] = o->nu[g]; }
    for (int k = 1; k < K; ++k)
        for (int g = 0; g < G; ++g) {
            bq_rng r; bq_rng_init(&r, o->seed, o->chain, (uint32_t)g, (uint32_t)k, (uint32_t)t, BQ_PH_BETA_P);
            size_t j = (size_t)k * G + g;
            o->beta_p[j] = o->beta[j] + BQ_RW_SD * bq_normal(&r);
        }
    for (int b = 1; b < B; ++b)
        for (int g = 0; g < G; ++g) {
            bq_rng r; bq_rng_init(&r, o->seed, o->chain, (uint32_t)g, (uint32_t)b, (uint32_t)t, BQ_PH_NU_P);
            size_t j = (size_t)b * G + g;
            o->nu_p[j] = o->nu[j] + BQ_RW_SD * bq_normal(&r);
        }
}

/* beta decisions (cu:696-707) then nu decisions (cu:730-737) */
static void decide_beta_nu(orc_t* o, int t)
{
    const int B = o->B, G = o->G, K = o->K;
    for (int k = 1; k < K; ++k)
        for (int g = 0; g < G; ++g) {
            size_t j = (size_t)k * G + g;
            double sum = bq_unfix(o->acc_beta[j]);
            o->beta_sum[j] = sum;
            double lr = sum + bq_lr_beta(o->beta[j], o->beta_p[j], o->L[j], o->tau0);
            int acc = bq_log(ph_uniform(o, (uint32_t)g, (uint32_t)k, t, BQ_PH_BETA_A)) <= lr;
            o->beta_acc[j] = acc;
            if (acc) { o->beta[j] = o->beta_p[j]; refresh_beta(o, k, g); }
        }
    for (int b = 1; b < B; ++b)
        for (int g = 0; g < G; ++g) {
            uint64_t a = 0;
            for (int k = 0; k < K; ++k) {
                size_t bkg = ((size_t)b * K + k) * G + g;
                a += (k >= 1 && o->beta_acc[(size_t)k * G + g]) ? o->acc_nu1[bkg] : o->acc_nu0[bkg];
            }
            size_t j = (size_t)b * G + g;
            double sum = bq_unfix(a);
            o->nu_sum[j] = sum;
            double lr = sum + bq_lr_normal(o->nu[j], o->nu_p[j], o->mu_nu[b], BQ_SIGMA2_NU);
            if (bq_log(ph_uniform(o, (uint32_t)g, (uint32_t)b, t, BQ_PH_NU_A)) <= lr) {
                o->nu[j] = o->nu_p[j]; refresh_nu(o, b, g);
            }
        }
}

/* phi proposals (cu:780-787) */
static void propose_phi(orc_t* o, int t)
{
    for (int b = 0; b < o->B; ++b)
        for (int g = 0; g < o->G; ++g) {
            size_t j = (size_t)b * o->G + g;
            bq_rng r; bq_rng_init(&r, o->seed, o->chain, (uint32_t)g, (uint32_t)b, (uint32_t)t, BQ_PH_PHI_P);
            o->phi_p[j] = bq_rgamma(&r, o->phi[j], 1.0);
        }
}

/* phi decisions (cu:806-818) */
static void decide_phi(orc_t* o, int t)
{
    for (int b = 0; b < o->B; ++b)
        for (int g = 0; g < o->G; ++g) {
            size_t j = (size_t)b * o->G + g;
            double sum = bq_unfix(o->acc_phi[j]);
            o->phi_sum[j] = sum;
            double f = o->phi[j], fp = o->phi_p[j];
            double u = ph_uniform(o, (uint32_t)g, (uint32_t)b, t, BQ_PH_PHI_A);
            if (!(fp > 0.0)) continue;
            double lr = sum + bq_lr_phi(f, fp);
            if (bq_log(u) <= lr) { o->phi[j] = fp; refresh_phi(o, b, g); }
        }
}

/* accumulate the recorded sweep (cu:892-936, 1547-1554) */
static void record(orc_t* o)
{
    const int B = o->B, G = o->G, K = o->K, N = o->N;
    o->n_rec += 1;
    for (int j = 0; j < 2 * B; ++j) o->s_gamma[j] += o->gamma[j];
    for (int g = 0; g < G; ++g) o->s_alpha[g] += o->alpha[g];
    for (size_t j = 0; j < (size_t)K * G; ++j) { o->s_beta[j] += o->beta[j]; o->c_L[j] += o->L[j]; }
    for (size_t j = 0; j < (size_t)B * G; ++j) { o->s_nu[j] += o->nu[j]; o->s_phi[j] += o->phi[j]; }
    for (int i = 0; i < N; ++i) { o->s_delta[i] += o->delta[i]; o->c_W[(size_t)i * K + o->W[i]] += 1; }
    for (size_t j = 0; j < (size_t)B * K; ++j) o->s_pi[j] += o->pi[j];
    o->s_p += o->p; o->s_tau0 += o->tau0;
    const int r = o->n_rec - 1, P = o->n_pres;
    if (!o->tr_alpha || r >= P) return;
    for (int j = 0; j < 2 * B; ++j) o->tr_gamma[(size_t)j * P + r] = o->gamma[j];
    for (int g = 0; g < G; ++g) o->tr_alpha[(size_t)g * P + r] = o->alpha[g];
    for (size_t j = G; j < (size_t)K * G; ++j) {
        o->tr_L[(j - G) * P + r] = o->L[j];
        o->tr_beta[(j - G) * P + r] = o->beta[j];
    }
    o->tr_ptau[r] = o->p; o->tr_ptau[P + r] = o->tau0;
    for (size_t j = 0; j < (size_t)B * G; ++j) {
        if (j >= (size_t)G) o->tr_nu[(j - G) * P + r] = o->nu[j];
        o->tr_phi[j * P + r] = o->phi[j];
    }
    for (int i = 0; i < N; ++i) {
        o->tr_delta[(size_t)i * P + r] = o->delta[i];
        o->tr_W[(size_t)i * P + r] = o->W[i];
    }
    for (size_t j = 0; j < (size_t)B * K; ++j) o->tr_pi[j * P + r] = o->pi[j];
}

static void sweep(orc_t* o)
{
    int t = o->iter;
    propose_alpha_gamma(o, t);
    if (o->prec == 32) pass_w_f32(o, t); else pass_w_f64(o, t);
    update_pi(o, t);
    build_tiles(o);
    if (o->prec == 32) pass_xz_alpha_f32(o, t); else pass_xz_alpha_f64(o, t);
    decide_alpha(o, t);
    decide_gamma(o, t);
    update_L(o, t);
    if (t >= o->n_unchanged) update_ptau(o, t);
    propose_beta_nu(o, t);
    if (o->prec == 32) pass_beta_nu_f32(o); else pass_beta_nu_f64(o);
    decide_beta_nu(o, t);
    if (o->prec == 32) pass_delta_f32(o, t); else pass_delta_f64(o, t);
    propose_phi(o, t);
    if (o->prec == 32) pass_phi_f32(o); else pass_phi_f64(o);
    decide_phi(o, t);
    if (t >= o->n_burnin) record(o);
    o->iter = t + 1;
}

void orc_sweeps(orc_t* o, int n)
{
    for (int s = 0; s < n; ++s) sweep(o);
}

/* ------------------------------------------------------------------------ */
/* state exchange                                                             */
/* ------------------------------------------------------------------------ */
void orc_get_state(const orc_t* o, orc_state* s)
{
    const int B = o->B, G = o->G, K = o->K, N = o->N;
    if (s->S) memcpy(s->S, o->S, 4 * (size_t)N * G);
    if (s->W) memcpy(s->W, o->W, 4 * (size_t)N);
    if (s->alpha) memcpy(s->alpha, o->alpha, 8 * (size_t)G);
    if (s->beta) memcpy(s->beta, o->beta, 8 * (size_t)K * G);
    if (s->L) memcpy(s->L, o->L, 4 * (size_t)K * G);
    if (s->nu) memcpy(s->nu, o->nu, 8 * (size_t)B * G);
    if (s->delta) memcpy(s->delta, o->delta, 8 * (size_t)N);
    if (s->phi) memcpy(s->phi, o->phi, 8 * (size_t)B * G);
    if (s->gamma) memcpy(s->gamma, o->gamma, 8 * 2 * (size_t)B);
    if (s->pi) memcpy(s->pi, o->pi, 8 * (size_t)B * K);
    if (s->ptau) { s->ptau[0] = o->p; s->ptau[1] = o->tau0; }
}

void orc_set_state(orc_t* o, const orc_state* s)
{
    const int B = o->B, G = o->G, K = o->K, N = o->N;
    if (s->S) memcpy(o->S, s->S, 4 * (size_t)N * G);
    if (s->W) memcpy(o->W, s->W, 4 * (size_t)N);
    if (s->alpha) memcpy(o->alpha, s->alpha, 8 * (size_t)G);
    if (s->beta) memcpy(o->beta, s->beta, 8 * (size_t)K * G);
    if (s->L) memcpy(o->L, s->L, 4 * (size_t)K * G);
    if (s->nu) memcpy(o->nu, s->nu, 8 * (size_t)B * G);
    if (s->delta) memcpy(o->delta, s->delta, 8 * (size_t)N);
    if (s->phi) memcpy(o->phi, s->phi, 8 * (size_t)B * G);
    if (s->gamma) memcpy(o->gamma, s->gamma, 8 * 2 * (size_t)B);
    if (s->pi) memcpy(o->pi, s->pi, 8 * (size_t)B * K);
    if (s->ptau) { o->p = s->ptau[0]; o->tau0 = s->ptau[1]; }
    refresh_all(o);
}

void orc_get_priors(const orc_t* o, double* mu_alpha, double* mu_nu, double* mu_delta)
{
    if (mu_alpha) memcpy(mu_alpha, o->mu_alpha, 8 * (size_t)o->G);
    if (mu_nu) memcpy(mu_nu, o->mu_nu, 8 * (size_t)o->B);
    if (mu_delta) memcpy(mu_delta, o->mu_delta, 8 * (size_t)o->N);
}

void orc_get_debug(const orc_t* o, orc_debug* d)
{
    const int B = o->B, G = o->G, K = o->K, N = o->N;
#define CP(f, n) if (d->f) memcpy(d->f, o->f, 8 * (size_t)(n))
    CP(alpha_sum, G); CP(beta_sum, (size_t)K * G); CP(nu_sum, (size_t)B * G);
    CP(delta_sum, N); CP(phi_sum, (size_t)B * G); CP(gamma_sum, 2 * B); CP(w_logit, (size_t)N * K);
#undef CP
    if (d->alpha_prop) memcpy(d->alpha_prop, o->alpha_p, 8 * (size_t)G);
    if (d->beta_prop) memcpy(d->beta_prop, o->beta_p, 8 * (size_t)K * G);
    if (d->nu_prop) memcpy(d->nu_prop, o->nu_p, 8 * (size_t)B * G);
    if (d->delta_prop) memcpy(d->delta_prop, o->delta_p, 8 * (size_t)N);
    if (d->phi_prop) memcpy(d->phi_prop, o->phi_p, 8 * (size_t)B * G);
    if (d->gamma_prop) memcpy(d->gamma_prop, o->gamma_p, 8 * 2 * (size_t)B);
}

/* ------------------------------------------------------------------------ */
/* posterior (cu:1557-1608) and BIC (cu:967-1024, 1815-1833)                  */
/* ------------------------------------------------------------------------ */
void orc_posterior(const orc_t* o, orc_post* p)
{
    const int B = o->B, G = o->G, K = o->K, N = o->N;
    double n = (double)o->n_rec;
    for (int j = 0; j < 2 * B; ++j) p->gamma[j] = o->s_gamma[j] / n;
    for (int g = 0; g < G; ++g) p->alpha[g] = o->s_alpha[g] / n;
    for (size_t j = 0; j < (size_t)K * G; ++j) { p->beta[j] = o->s_beta[j] / n; p->L[j] = (double)o->c_L[j] / n; }
    for (size_t j = 0; j < (size_t)B * G; ++j) { p->nu[j] = o->s_nu[j] / n; p->phi[j] = o->s_phi[j] / n; }
    for (int i = 0; i < N; ++i) {
        p->delta[i] = o->s_delta[i] / (g_mut == 10 ? n + 1.0 : n);   /* 10: planted divisor */
        int best = 0, bv = o->c_W[(size_t)i * K];
        for (int k = 1; k < K; ++k) {
            const int v = o->c_W[(size_t)i * K + k];
            if (v > bv || (g_mut == 5 && v == bv)) { bv = v; best = k; }   /* 5: planted tie-break */
        }
        p->W[i] = best;
    }
    for (size_t j = 0; j < (size_t)B * K; ++j) p->pi[j] = o->s_pi[j] / n;
    p->ptau[0] = o->s_p / n; p->ptau[1] = o->s_tau0 / n;
}

double orc_bic(const orc_t* o, const orc_post* p)
{
    const int B = o->B, G = o->G, K = o->K, N = o->N;
    double* lli = (double*)zal(8 * (size_t)N);
    #pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < N; ++i) {
        int b = o->batch[i];
        double ll[ORC_KMAX], mx = -BQ_INF;
        for (int k = 0; k < K; ++k) {
            uint64_t acc = 0;
            for (int c0 = 0; c0 < G; c0 += CHUNK) {
                double part[LANES];
                for (int l = 0; l < LANES; ++l) part[l] = 0.0;
                int c1 = c0 + CHUNK < G ? c0 + CHUNK : G;
                for (int g = c0; g < c1; ++g) {
                    int32_t s = o->S[(size_t)i * G + g];
                    int y = s > 0 ? s : 0;
                    part[g & (LANES - 1)] += bq_bic_elem(p->alpha[g], p->beta[(size_t)k * G + g],
                                           p->nu[(size_t)b * G + g], p->delta[i], p->gamma[b], p->gamma[B + b],
                                           p->phi[(size_t)b * G + g], y);
                }
                for (int l = 0; l < LANES; ++l) acc += bq_fix(part[l]);
            }
            ll[k] = bq_unfix(acc);
            if (ll[k] > mx) mx = ll[k];
        }
        double s = 0.0;
        for (int k = 0; k < K; ++k) s += (g_mut == 7 ? 1.0 : p->pi[b * K + k]) * bq_exp(ll[k] - mx);   /* 7: planted */
        lli[i] = mx + bq_log(s);
    }
    double tot = 0.0;
    for (int i = 0; i < N; ++i) tot += lli[i];
    free(lli);
    double npar = (double)(B + G) * K + 2.0 * B + (double)G * (B * 2 - 1) + (g_mut == 6 ? 0 : N - B);   /* 6: planted */
    return -2.0 * tot + bq_log((double)G * (double)N) * npar;
}

/* ------------------------------------------------------------------------ */
/* numerics hooks for the tests                                               */
/* ------------------------------------------------------------------------ */
void orc_philox(const uint32_t* c, const uint32_t* k, uint32_t* out)
{
    bq_u32x4 v = bq_philox4x32_10(c[0], c[1], c[2], c[3], k[0], k[1]);
    out[0] = v.v0; out[1] = v.v1; out[2] = v.v2; out[3] = v.v3;
}

/* kind: 0 uniform, 1 normal, 2 gamma(p1, scale p2), 3 poisson(p1), 4 nbinom(mu p1, phi p2),
 *       5 beta(p1, p2), 6 raw u32 words, 7 the X/Z proposal draw nbinom_xz(mu p1, phi p2) */
void orc_stream(uint32_t seed, uint32_t chain, uint32_t a, uint32_t b, uint32_t sweep,
                uint32_t phase, int kind, double p1, double p2, int n, double* out)
{
    bq_rng r; bq_rng_init(&r, seed, chain, a, b, sweep, phase);
    for (int j = 0; j < n; ++j) {
        switch (kind) {
        case 0: out[j] = bq_uniform(&r); break;
        case 1: out[j] = bq_normal(&r); break;
        case 2: out[j] = bq_rgamma(&r, p1, p2); break;
        case 3: out[j] = (double)bq_rpois(&r, p1); break;
        case 4: out[j] = (double)bq_rnbinom(&r, p1, p2); break;
        case 5: out[j] = bq_rbeta(&r, p1, p2); break;
        case 7: out[j] = (double)bq_rnbinom_xz(&r, p1, p2); break;
        default: out[j] = (double)bq_next32(&r); break;
        }
    }
}

/* negative-binomial inversion at given uniforms (admissible (mu, phi) only) */
void orc_nb_inv(const double* u, const double* mu, const double* phi, double* out, int n)
{
    for (int j = 0; j < n; ++j) out[j] = (double)bq_rnbinom_inv(u[j], mu[j], phi[j]);
}

/* fn: 0 log, 1 log1p, 2 exp, 3 expm1, 4 lgamma, 5 lgamma_diff64(x,y), 6 logf, 7 log1pf,
 *     8 lgamma_diff32(x,y), 9 softplus, 10 ndtri, 11 fix/unfix round trip, 12 log1p_mid,
 *     13 log1pf_mid */
void orc_math(int fn, const double* x, const double* y, double* out, int n)
{
    for (int j = 0; j < n; ++j) {
        double v = x[j];
        switch (fn) {
        case 0: out[j] = bq_log(v); break;
        case 1: out[j] = bq_log1p(v); break;
        case 2: out[j] = bq_exp(v); break;
        case 3: out[j] = bq_expm1(v); break;
        case 4: out[j] = bq_lgamma(v); break;
        case 5: out[j] = bq_lgamma_diff64(v, y[j]); break;
        case 6: out[j] = (double)bq_logf((float)v); break;
        case 7: out[j] = (double)bq_log1pf((float)v); break;
        case 8: out[j] = (double)bq_lgamma_diff32((float)v, (float)y[j]); break;
        case 9: out[j] = bq_softplus(v); break;
        case 10: out[j] = bq_ndtri(v); break;
        case 12: out[j] = bq_log1p_mid(v); break;
        case 13: out[j] = (double)bq_log1pf_mid((float)v); break;
        default: out[j] = bq_unfix(bq_fix(v)); break;
        }
    }
}

/* runs the X/Z update of one Y = 0 entry for n steps with fixed parameters;
 * counts[x*2 + z] for x < nx (the stationary law test) */
void orc_xz_chain(double mu, double phi, double g0, double g1, uint32_t seed, int n, int nx, int64_t* counts)
{
    int32_t s = 0;
    double qz = 1.0 / (1.0 + bq_exp(g0));
    for (int t = 0; t < n; ++t) {
        bq_rng r; bq_rng_init(&r, seed, 7u, 0u, 0u, (uint32_t)t, BQ_PH_XZ);
        s = bq_xz_update(s, mu, phi, g0, g1, qz, &r);
        int x = s < 0 ? ~s : 0, z = s < 0;
        if (x < nx) counts[x * 2 + z] += 1;
    }
}

/* element terms for the tests: fn 0 tmh(x,phi,r,dlt,c), 1 tw(x,phi,r0,beta,em),
 * 2 tphi(x,phi,phip,mu) with the per-(b,g) constants derived here, 3 ratio(mu,phi),
 * 4 lr_gamma1(g,gn), 5 lr_phi(phi,phin), 6 logodds_L(p,tau0,beta), 7 lr_normal(old,new,m,s2),
 * 8 bic_elem(a,bk,nb,dl,g0,g1,phi,y), 9 lr_beta(old,new,L,tau0) */
void orc_terms(int fn, int prec, const double* a, int na, double* out, int n)
{
    for (int j = 0; j < n; ++j) {
        const double* v = a + (size_t)j * na;
        double res = 0.0;
        if (fn == 4) { out[j] = bq_lr_gamma1(v[0], v[1]); continue; }
        if (fn == 5) { out[j] = bq_lr_phi(v[0], v[1]); continue; }
        if (fn == 6) { out[j] = bq_logodds_L(v[0], v[1], v[2]); continue; }
        if (fn == 7) { out[j] = bq_lr_normal(v[0], v[1], v[2], v[3]); continue; }
        if (fn == 8) { out[j] = bq_bic_elem(v[0], v[1], v[2], v[3], v[4], v[5], v[6], (int)v[7]); continue; }
        if (fn == 9) { out[j] = bq_lr_beta(v[0], v[1], (int)v[2], v[3]); continue; }
        if (fn == 2) {
            double f = v[1], fp = v[2], d = fp - f;
            if (prec == 32) {
                res = (double)bq_tphi_nb_f32((float)v[0], (float)f, (float)fp, (float)v[3], (float)d,
                                             (float)(d / f), (float)(1.0 / fp), (float)bq_log1p(d / f));
                if (v[0] > 0.0)
                    res += (double)bq_tphi_lg_f32((float)v[0], (float)f, (float)d, (float)bq_lgamma_diff64(f, d));
            } else {
                res = bq_tphi_nb_f64(v[0], f, fp, v[3], d, d / f, 1.0 / fp, bq_log1p(d / f));
                if (v[0] > 0.0) res += bq_tphi_lg_f64(v[0], f, d, bq_lgamma_diff64(f, d));
            }
        } else if (prec == 32) {
            if (fn == 0) res = bq_tmh_f32((float)v[0], (float)v[1], (float)v[2], (float)v[3], (float)v[4]);
            else if (fn == 1) res = bq_tw_f32((float)v[0], (float)v[1], (float)v[2], (float)v[3], bq_tw_clamp_f32((float)v[4]));
            else res = bq_ratio_f32((float)v[0], (float)v[1]);
        } else {
            if (fn == 0) res = bq_tmh_f64(v[0], v[1], v[2], v[3], v[4]);
            else if (fn == 1) res = bq_tw_f64(v[0], v[1], v[2], v[3], v[4]);
            else res = bq_ratio_f64(v[0], v[1]);
        }
        out[j] = res;
    }
}

/* ------------------------------------------------------------------------ */
/* traces and single steps for the tests                                      */
/* ------------------------------------------------------------------------ */
int orc_keep_traces(orc_t* o, int on)
{
    if (!on || o->tr_alpha) return 0;
    const int B = o->B, G = o->G, K = o->K, N = o->N;
    const size_t P = (size_t)(o->n_iter > o->n_burnin ? o->n_iter - o->n_burnin : 0);
    o->n_pres = (int)P;
    o->tr_gamma = (double*)zal(8 * 2 * (size_t)B * P);
    o->tr_alpha = (double*)zal(8 * (size_t)G * P);
    o->tr_L = (int32_t*)zal(4 * (size_t)(K - 1) * G * P);
    o->tr_beta = (double*)zal(8 * (size_t)(K - 1) * G * P);
    o->tr_ptau = (double*)zal(8 * 2 * P);
    o->tr_nu = (double*)zal(8 * (size_t)(B - 1) * G * P);
    o->tr_delta = (double*)zal(8 * (size_t)N * P);
    o->tr_phi = (double*)zal(8 * (size_t)B * G * P);
    o->tr_W = (int32_t*)zal(4 * (size_t)N * P);
    o->tr_pi = (double*)zal(8 * (size_t)B * K * P);
    return o->tr_alpha && o->tr_W && o->tr_delta ? 0 : -1;
}

int orc_get_traces(const orc_t* o, orc_traces* t)
{
    if (!o->tr_alpha) return 0;
    const int B = o->B, G = o->G, K = o->K, N = o->N;
    const size_t P = (size_t)o->n_pres;
#define CPT(f, n, sz) if (t->f) memcpy(t->f, o->f, (sz) * (size_t)(n) * P)
    CPT(tr_gamma, 2 * B, 8); CPT(tr_alpha, G, 8); CPT(tr_L, (size_t)(K - 1) * G, 4);
    CPT(tr_beta, (size_t)(K - 1) * G, 8); CPT(tr_ptau, 2, 8); CPT(tr_nu, (size_t)(B - 1) * G, 8);
    CPT(tr_delta, N, 8); CPT(tr_phi, (size_t)B * G, 8); CPT(tr_W, N, 4); CPT(tr_pi, (size_t)B * K, 8);
#undef CPT
    return o->n_pres;
}

/* one step of sweep index t on the current state: 0 the pi draws, 1 the p and
 * tau0 draws, 2 the L draws */
void orc_step(orc_t* o, int which, int t)
{
    if (which == 0) update_pi(o, t);
    else if (which == 1) update_ptau(o, t);
    else if (which == 2) update_L(o, t);
}
