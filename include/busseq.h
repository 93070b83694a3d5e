/*
 * busseq.h -- C ABI of the MI355X-native BUSseq Gibbs sampler (libbusseq.so).
 *
 * The reference (BUSseq_gpu.cu) has no library API: its boundary is the
 * executable `BUSseq_gpu -B -N -G -K -s -c -i -b -u -p -o` and its text
 * outputs (BUSseq_gpu.cu:1026-2021, README.md:17-49).  This header is the
 * library form of that boundary: plain pointers and sizes, status codes
 * instead of exit(), one sampler handle per host thread.  The bundled
 * executable BUSseq_gpu is a thin CLI over these calls.
 *
 * Layout conventions follow the reference (BUSseq_gpu.cu:1148, 1177-1189):
 *   counts Y, X, Z   N x G int32, row-major [i*G + g], cells grouped by batch
 *   beta, L          K x G  [k*G + g]   (row k = 0 is the reference subtype)
 *   nu, phi          B x G  [b*G + g]   (nu row b = 0 is zero)
 *   pi               B x K  [b*K + k]
 *   gamma            2 x B  [0*B + b] = gamma0, [1*B + b] = gamma1
 *   ptau             2      p, tau0
 */
#ifndef BUSSEQ_H
#define BUSSEQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    BUSSEQ_OK = 0,
    BUSSEQ_ERR_ARG = 1,      /* invalid configuration (cu:1120-1134 checks and more) */
    BUSSEQ_ERR_IO = 2,       /* file could not be opened or written */
    BUSSEQ_ERR_FORMAT = 3,   /* count or dimension file malformed or short */
    BUSSEQ_ERR_DATA = 4,     /* data the model cannot use (e.g. a cell with zero counts) */
    BUSSEQ_ERR_DEVICE = 5,   /* HIP runtime error */
    BUSSEQ_ERR_NOMEM = 6     /* host or device allocation failed */
} busseq_status;

typedef struct busseq_config {
    int B;                /* batches (-B) */
    const int* n_b;       /* cells per batch, B entries (-N file) */
    int G;                /* genes (-G) */
    int K;                /* subtypes (-K), 2 <= K <= 64, with B * K <= 2048 */
    uint32_t seed;        /* -s */
    int n_iter;           /* -i */
    int n_burnin;         /* -b */
    int n_unchanged;      /* -u: p and tau0 are updated from this iteration on */
    int keep_traces;      /* -p: keep every recorded iteration */
    int precision;        /* 32: single-precision per-element terms (default); 64: double */
    int chains;           /* independent chains on the same data in one launch (default 1) */
    int chain0;           /* Philox key word of the first chain (default 0) */
    int tile_rows;        /* cells per reduction tile, 1..256 (0, the default: busseq_default_tile_rows) */
    int device;           /* HIP device ordinal (default 0) */
} busseq_config;

/* fills the reference defaults of BUSseq_gpu.cu:1028-1038 (n_b left NULL) */
void busseq_config_default(busseq_config* cfg);

/* one chain's state; caller-allocated arrays of the documented sizes, NULL = skip */
typedef struct busseq_state {
    int32_t* X; int32_t* Z; int32_t* W;
    double* alpha; double* beta; int32_t* L; double* nu; double* delta; double* phi;
    double* gamma; double* pi; double* ptau;
} busseq_state;

/* posterior summaries; library-allocated, released by busseq_result_free */
typedef struct busseq_result {
    int B, N, G, K, n_pres;
    int* n_b;
    int32_t* X; int32_t* Z;   /* last iteration (cu:1614-1645) */
    int32_t* W;               /* posterior mode, lowest k on ties (cu:944-959) */
    double* gamma; double* alpha; double* L; double* beta; double* ptau;
    double* nu; double* delta; double* phi; double* pi;   /* posterior means; L = P(L=1) */
    double bic;               /* cu:967-1024, 1815-1833 */
    /* every recorded iteration when keep_traces: [param * n_pres + iter] (cu:892-936) */
    double* tr_gamma; double* tr_alpha; int32_t* tr_L; double* tr_beta; double* tr_ptau;
    double* tr_nu; double* tr_delta; double* tr_phi; int32_t* tr_W; double* tr_pi;
} busseq_result;

typedef struct busseq_sampler busseq_sampler;

/* Whole run: init + n_iter sweeps + posterior + BIC.  Replaces main() of
 * BUSseq_gpu.cu:1163-1843 (everything after argument parsing and file input). */
busseq_status busseq_run(const busseq_config* cfg, const int32_t* Y, busseq_result* out);

/* Text outputs <prefix>_post_*.txt, _BIC.txt and, if print_preserved,
 * _preserved_*.txt in the reference's formats (BUSseq_gpu.cu:1610-2018). */
busseq_status busseq_write_outputs(const busseq_result* r, const char* prefix, int print_preserved);
void busseq_result_free(busseq_result* r);

/* Stepping interface (same semantics as busseq_run, split for tests and timing).
 * create = allocation + initialisation (BUSseq_gpu.cu:1163-1244). */
busseq_status busseq_create(const busseq_config* cfg, const int32_t* Y, busseq_sampler** out);
/* n Gibbs sweeps (the loops of BUSseq_gpu.cu:1267-1555), replayed from one hipGraph */
busseq_status busseq_sweeps(busseq_sampler* s, int n);
/* n sweeps timed with HIP events on the sampler's stream; *ms = elapsed milliseconds */
busseq_status busseq_time_sweeps(busseq_sampler* s, int n, double* ms);
int busseq_iteration(const busseq_sampler* s);
busseq_status busseq_get_state(busseq_sampler* s, int chain, busseq_state* st);
busseq_status busseq_set_state(busseq_sampler* s, int chain, const busseq_state* st);
/* posterior + BIC of one chain after the run (BUSseq_gpu.cu:1557-1608, 1815-1833) */
busseq_status busseq_finish(busseq_sampler* s, int chain, busseq_result* out);
/* BIC of one chain's counts at the given parameter values (host arrays in the
 * layouts above: alpha G, beta K*G, nu B*G, delta N, phi B*G, gamma 2*B, pi B*K):
 * the per-count likelihood with dropout marginalised, the pi-weighted mixture
 * over subtypes and the penalty of BUSseq_gpu.cu:967-1024, 1815-1833 (which
 * evaluates it at the posterior means; busseq_finish calls this).  Allocates
 * (N*K + G*(K+2B+1) + N + 2B + B*K + 1) doubles of device memory per call. */
busseq_status busseq_bic(busseq_sampler* s, int chain, const double* alpha, const double* beta,
                         const double* nu, const double* delta, const double* phi, const double* gamma,
                         const double* pi, double* bic);
/* busseq_bic without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under stream capture,
 * serving loops, harnesses that forbid hipMalloc.  The same kernels and the
 * same result bit for bit, on a 32 MiB array that each device holds from the
 * library's load to the end of the process.  Rejects, with BUSSEQ_ERR_ARG, a
 * chain whose N*K + G*(K+2B+1) + N + 2B + B*K + 1 doubles exceed 4194304.
 * Calls on one device share that array: calls that overlap in time on one
 * device (from several threads or samplers) race. */
busseq_status busseq_bic_noalloc(busseq_sampler* s, int chain, const double* alpha, const double* beta,
                                 const double* nu, const double* delta, const double* phi,
                                 const double* gamma, const double* pi, double* bic);
void busseq_destroy(busseq_sampler* s);

/* Batched form (no counterpart in the reference, which runs one chain per
 * process): several problems -- different count matrices, K values, seeds --
 * each with cfg->chains chains, advanced together: every sweep of every chain
 * runs in the same kernel launches of one hipGraph.  Chains are numbered in
 * problem order; chain j of a problem uses Philox key word cfg->chain0 + j and
 * reproduces the plain run with that chain0 bit for bit.  Problems must share
 * precision, tile_rows and device. */
typedef struct busseq_problem {
    const busseq_config* cfg;
    const int32_t* Y;
} busseq_problem;
busseq_status busseq_create_batch(const busseq_problem* probs, int n_probs, busseq_sampler** out);
int busseq_chain_count(const busseq_sampler* s);

/* Input files.  -N file: B whitespace-separated ints (cu:1094-1098).
 * -c file: N*G whitespace-separated ints, cells as rows (cu:1145-1159). */
busseq_status busseq_read_dims(const char* path, int B, int* n_b);
/* Replaces the fscanf loop of cu:1145-1159.  Text is parsed by all OpenMP
 * threads; the error reported is the one a serial reader meets first (a bad
 * token among the first N*G, a short file, a negative count).  A file written
 * by busseq_write_counts_binary is recognised by its header and read as is. */
busseq_status busseq_read_counts(const char* path, int N, int G, int32_t* Y);
/* Binary count file (no counterpart in the reference): a 32-byte header
 * ("BQCOUNT1", int64 N, int64 G, int64 0) and N*G little-endian int32, cells
 * as rows; written to path.tmp and renamed, so readers never see a partial file. */
busseq_status busseq_write_counts_binary(const char* path, int N, int G, const int32_t* Y);

/* Synthetic counts from the BUSseq generative model (no counterpart in the
 * reference): alpha_g ~ N(1.5, 0.5^2); beta_gk = 0 for k = 0 and N(+-2, 0.5^2)
 * on 10% of the genes for each k >= 1; nu_bg = 0 for b = 0, else N(0, 0.5^2);
 * delta_i ~ N(0, 0.2^2), 0 for each batch's first cell; phi_bg ~ Gamma(5, rate
 * 0.5); pi_b ~ Dirichlet(1); X ~ NB(exp(alpha + beta + nu + delta), phi);
 * gamma0 ~ U(g0_lo, g0_hi), gamma1 ~ U(-0.3, -0.1); Z ~ Bernoulli(logistic(
 * gamma0 + gamma1 X)); Y = X (1 - Z).  Streams keyed by data_seed.
 * Outputs that are NULL are skipped: Y (N*G), true W (N), X (N*G). */
busseq_status busseq_simulate(int B, const int* n_b, int G, int K, uint32_t data_seed,
                              double g0_lo, double g0_hi, int32_t* Y, int32_t* W_true,
                              int32_t* X_true);

/* Cells per reduction tile for a configuration with tile_rows = 0 (no
 * counterpart in the reference): 16 below 2560 cells over all chains of the
 * sampler, 32 below 65536, else 64.  The tile size sets the grouping of the
 * per-gene sums (their last-bit rounding), not the model. */
int busseq_default_tile_rows(long long cells);

/* Tile autotuner (no counterpart in the reference): times `sweeps` sweeps of
 * the given problem(s) with each tile size in {16, 32, 64, 128} cells, twice in
 * turn, and returns the size with the fastest run in *tile_rows (and its
 * sweeps/s).  The tile size sets
 * the grouping of the per-gene sums (their last-bit rounding) and not the
 * model; a run is reproducible given the tile_rows it used.  The problems are
 * not modified; the timing chains are discarded. */
busseq_status busseq_autotune(const busseq_config* cfg, const int32_t* Y, int sweeps,
                              int* tile_rows, double* sweeps_per_s);
busseq_status busseq_autotune_batch(const busseq_problem* problems, int n_problems, int sweeps,
                                    int* tile_rows, double* sweeps_per_s);

/* Measurement helpers (no counterpart in the reference, which has no timers).
 * busseq_kernel_times runs n sweeps as individual launches with HIP events
 * between them, back to back on the sampler's stream with one wait at the end,
 * and returns the average milliseconds of each of the busseq_kernel_count()
 * kernels (names from busseq_kernel_name). */
int busseq_kernel_count(void);
const char* busseq_kernel_name(int j);
busseq_status busseq_kernel_times(busseq_sampler* s, int n, double* ms);
/* streaming copies of `bytes` bytes on the device: one warm-up launch, then *ms = the
 * average of ten launches (profiler counter calibration, achievable bandwidth) */
busseq_status busseq_copy_probe(size_t bytes, int device, double* ms);
/* Numerics self-check (no counterpart in the reference): *mismatches = the
 * number of normal floats z in [2^-126, 2^126) (252 x 2^23 values)
 * for which the device's reciprocal-plus-Newton form differs from the IEEE
 * quotient 1.0f / z of z held to [2^-126, 2^125], the host's value.  Single-precision samplers depend on that number being
 * 0; busseq_create/busseq_create_batch run the same check once per device and
 * process and fail with BUSSEQ_ERR_DEVICE otherwise. */
busseq_status busseq_rcp_check(int device, uint64_t* mismatches);

/* message of the last failing call on this thread */
const char* busseq_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
