/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.  CPU restatement of the BUSseq Gibbs
 * sampler of the reference (BUSseq_gpu.cu) used as the parity checker for the
 * gfx950 path and as the CPU baseline timer.  Nothing in the product links or
 * calls this; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * The oracle follows the reference's update rules (cited per function in
 * oracle.c) in the sampler's own scan order, driven by the same Philox streams
 * and the same numerics (include/busseq/numerics.h, model.h) as the GPU, so
 * integer draws match bit for bit.  Sums are exact fixed-point sums, so the
 * result does not depend on the number of OpenMP threads.
 */
#ifndef BUSSEQ_ORACLE_H
#define BUSSEQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* precision: 32 = single-precision per-element terms, 64 = double */
orc_t* orc_create(int B, const int* n_b, int G, int K, const int32_t* Y,
                  uint32_t seed, uint32_t chain, int precision,
                  int n_iter, int n_burnin, int n_unchanged);
void orc_destroy(orc_t* o);
int orc_init(orc_t* o);                 /* 0 ok, <0 invalid input (zero library size) */
void orc_sweeps(orc_t* o, int n);       /* n sweeps; records after burn-in */
int orc_iter(const orc_t* o);
void orc_set_tile_rows(orc_t* o, int tile_rows);   /* cells per reduction tile, default 64 */
void orc_set_threads(int n);

/* state views: caller-allocated arrays of the documented sizes */
typedef struct {
    int32_t* S;        /* N*G encoded X/Z: Y>0 -> X=Y (Z=0); Y=0 -> 0 (X=0,Z=0) or ~X (Z=1) */
    int32_t* W;        /* N   */
    double* alpha;     /* G   */
    double* beta;      /* K*G, row 0 zero */
    int32_t* L;        /* K*G, row 0 zero */
    double* nu;        /* B*G, row 0 zero */
    double* delta;     /* N   */
    double* phi;       /* B*G */
    double* gamma;     /* 2*B: gamma0 then gamma1 */
    double* pi;        /* B*K */
    double* ptau;      /* 2: p, tau0 */
} orc_state;

void orc_get_state(const orc_t* o, orc_state* s);
void orc_set_state(orc_t* o, const orc_state* s);
void orc_get_priors(const orc_t* o, double* mu_alpha, double* mu_nu, double* mu_delta);

/* per-sweep debug record of the last sweep (proposals and exact sums) */
typedef struct {
    double* alpha_prop; double* alpha_sum;     /* G */
    double* beta_prop;  double* beta_sum;      /* K*G (row 0 unused) */
    double* nu_prop;    double* nu_sum;        /* B*G (row 0 unused); nu_sum uses the accepted beta branch */
    double* delta_prop; double* delta_sum;     /* N */
    double* phi_prop;   double* phi_sum;       /* B*G */
    double* gamma_prop; double* gamma_sum;     /* 2*B: S0 then S1 */
    double* w_logit;                           /* N*K: sum term (k=0 is 0), without log pi */
} orc_debug;
void orc_get_debug(const orc_t* o, orc_debug* d);

/* posterior summaries after the recorded sweeps (BUSseq_gpu.cu:1557-1608) */
typedef struct {
    double* gamma; double* alpha; double* L; double* beta; double* ptau;
    double* nu; double* delta; double* phi; double* pi; int32_t* W;
} orc_post;
void orc_posterior(const orc_t* o, orc_post* p);
double orc_bic(const orc_t* o, const orc_post* p);   /* BUSseq_gpu.cu:967-1024,1815-1833 */

/* traces of the recorded sweeps in the reference's layout [param * n_pres + iter]
 * (BUSseq_gpu.cu:892-936): orc_keep_traces(o, 1) before the first sweep;
 * orc_get_traces copies them into caller arrays (NULL = skip) and returns n_pres */
typedef struct {
    double* tr_gamma; double* tr_alpha; int32_t* tr_L; double* tr_beta; double* tr_ptau;
    double* tr_nu; double* tr_delta; double* tr_phi; int32_t* tr_W; double* tr_pi;
} orc_traces;
int orc_keep_traces(orc_t* o, int on);
int orc_get_traces(const orc_t* o, orc_traces* t);

/* one step of sweep index t on the current state, for the tests that pin the
 * conditional draws: 0 pi (cu:856-890), 1 p and tau0 (cu:650-668), 2 L (cu:634-648) */
void orc_step(orc_t* o, int which, int t);

/* planted mistakes (tests only; 0 = none): 1 pi prior xi = 1, 2 p ~ Beta(., 1),
 * 3 mu_nu sign, 4 tau0 shape without the /2, 5 W mode ties to the highest k,
 * 6 BIC penalty without N - B, 7 BIC mixture without pi, 8 raw means over the
 * batch size, 9 delta0 from the batch's last cell, 10 posterior delta over n + 1 */
void orc_set_mutation(int m);

/* numerics hooks for the tests (the host build of include/busseq/numerics.h) */
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
void orc_stream(uint32_t seed, uint32_t chain, uint32_t a, uint32_t b, uint32_t sweep,
                uint32_t phase, int kind, double p1, double p2, int n, double* out);
void orc_math(int fn, const double* x, const double* y, double* out, int n);
void orc_nb_inv(const double* u, const double* mu, const double* phi, double* out, int n);
void orc_terms(int fn, int prec, const double* args, int nargs, double* out, int n);
void orc_xz_chain(double mu, double phi, double g0, double g1, uint32_t seed, int n, int nx, int64_t* counts);

#ifdef __cplusplus
}
#endif
#endif
