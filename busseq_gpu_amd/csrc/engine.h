// engine.h -- internal layout shared by the gfx950 kernels and the host driver.
//
// One ChainDev describes one Markov chain: its dimensions, its Philox key and
// the device addresses of its state.  Kernels index an array of them with
// blockIdx.z, so the plain form (one chain) and the batched form (many
// chains, possibly of different K, or of different data sets) run the same
// kernels.
#pragma once
#include <stdint.h>
#include <vector>

#define BQ_KMAX 64     // largest K (the subtype draw keeps subtype k's logit in lane k)
#define BQ_KREG 16     // subtypes per register group: the K <= 16 subtype-draw instantiations,
                       // the K > 16 draw's groups, and the chunks of the cells' partition
#define BQ_BMAX 128    // largest B (the reference's __constant__ limit, cu:9-10)
#define BQ_BKMAX 2048  // largest B * K (the partition's tables in shared memory)
#define BQ_WAVE 64
#define BQ_CHUNK 1024 // genes per row-sum chunk (part of the reduction contract)
#define BQ_TR_MAX 256  // largest tile (cells per tile-pass workgroup)
#define BQ_HLO 64      // histogram bins kept in striped copies (the frequent small values)
#define BQ_STRIPES 32 // copies of each per-sweep counter, one per group of workgroups, so that
                      // the atomics of thousands of cells do not queue on one cache line
#define BQ_PHI_HV 128  // counts 1 <= Y < BQ_PHI_HV of each (batch, gene) are kept as a histogram for
                      // the phi step's lgamma parts; larger counts and imputed X are summed one by one

// Fields of one chain.  Q qualifies the pointee of every device address:
// empty on the host, the global address space in kernels (so that loads,
// stores and atomics through them are global, not flat, instructions).
#define BQ_CHAIN_FIELDS(Q) \
    /* dimensions and schedule */ \
    int N, G, Gp, K, B; \
    uint32_t seed, key;  /* Philox key words: (seed, chain id) */ \
    int n_burnin, n_unchanged, n_pres, traces; \
    int tr_rows;  /* rows of the trace ring: sweep t is written to row t mod tr_rows */ \
    int TR, T_max;  /* rows per tile, tile capacity */ \
    int prec;  /* 32 or 64 */ \
    /* static data */ \
    const Q int32_t* batch;  /* [N] batch of each cell */ \
    const Q int32_t* boff;  /* [B+1] first cell of each batch */ \
    const Q int64_t* hist_s;  /* [B*H] static histogram of Y>0 entries with Y<H */ \
    const Q int32_t* zrow;  /* [N+1] CSR row offsets of the Y = 0 entries */ \
    const Q uint16_t* zcol;  /* [nz]  their genes */ \
    Q int32_t* zs;  /* [nz]  their current S values, kept equal to S at those entries */ \
    const Q int32_t* big_off;  /* [B+1] static counts Y >= H, per batch */ \
    const Q int32_t* big;  /* [nbig] their values */ \
    const Q uint64_t* hlist;  /* [n_hlist] the nonzero entries of the per-(batch, gene) histogram of the counts \
                               * 1 <= Y < BQ_PHI_HV: (j << 32) | (n << 7) | v for n cells with Y = v in item \
                               * j = b G + g, the entries of an item together */ \
    long long n_hlist; \
    /* state */ \
    Q int* iter;  /* [1] global iteration (sweep) index */ \
    Q int32_t* S;  /* [N*Gp] encoded X/Z (see numerics contract) */ \
    Q int32_t* W;  /* [N] */ \
    Q int32_t* L;  /* [K*G] */ \
    Q double *alpha, *beta, *nu, *delta, *phi, *gamma, *pi, *ptau; \
    Q double *mu_alpha, *mu_nu, *mu_delta; \
    /* precision-T copies of the parameters */ \
    Q void *ea, *eb, *bt, *en, *ed, *ph;  /* bt: [G][bt_stride(K)][2] (beta_k, expm1 beta_k clamped as bq_tw_em_f32 in single precision), k = 1..K-1 */ \
    /* proposals and their precision-T constants */ \
    Q double *alpha_p, *beta_p, *nu_p, *phi_p, *gamma_p; \
    Q double *delta_p, *delta_lu;  /* [N] delta proposal and log of its accept uniform (k_prep) */ \
    Q void *dd, *cd;  /* [N] delta' - delta and expm1 of it in precision T */ \
    Q void *da, *ca, *db, *cb, *dn, *cn; \
    Q void* pc;  /* [6][B*G]: phi', d, d/phi, 1/phi', log(phi'/phi), lgamma diff */ \
    Q void* tm;  /* [K*B*G] (e^alpha e^beta_k) e^nu_b in precision T, the X/Z step's mean without e^delta */ \
    /* plan of the sweep */ \
    Q int32_t* cnt;  /* [BQ_STRIPES][B*K] cells per (batch, subtype), striped */ \
    Q int32_t* perm;  /* [N] cells sorted stably by (batch, subtype, index) */ \
    Q int32_t* tiles;  /* [T_max*4] (b, k, start, count) */ \
    Q int32_t* tile_off;  /* [B*K+1] first tile of each (batch, subtype) */ \
    Q int32_t* ntiles;  /* [1] */ \
    Q uint64_t* gsum;  /* exact per-item sums the tile passes add their tile values to (wrapping \
                        * 64-bit fixed point, so any order gives the same bits); read and cleared \
                        * by the decision kernels: alpha [G] | beta [K*G] | nu without and with \
                        * the accepted beta of the group's subtype [2][B*K*G] | phi [B*G] */ \
    /* gamma statistics */ \
    Q uint32_t* hist_d;  /* [B*H] dynamic histogram (entries with Y = 0), values >= BQ_HLO */ \
    Q uint32_t* hist_lo;  /* [BQ_STRIPES][B*BQ_HLO] the same for values < BQ_HLO, striped */ \
    Q uint64_t* gstat;  /* [BQ_STRIPES][B*5] n_z1, sum x z, over0, over1_rej, over1_acc, striped */ \
    Q uint64_t* lstat;  /* [2] sum of L and exact sum of beta^2 over L = 0 (k >= 1), for p and tau0 */ \
    Q double* odds_tab;  /* [B*H] bq_xz_odds(gamma0, gamma1, x) = 1 + e^{-(gamma0 + gamma1 x)}, current gamma */ \
    Q int32_t* xz_tail;  /* [nz] per cell, at the cell's zero-entry offset: entries left to the gamma-Poisson draw */ \
    Q int32_t* xz_ntail;  /* [N] their number per cell this sweep */ \
    Q int32_t* beta_acc;  /* [K*G] 1 where this sweep accepted beta_kg */ \
    /* posterior accumulators */ \
    Q int* n_rec; \
    Q int* iter_rec;  /* [1] the sweep index k_record uses, copied from iter by k_bn_dec */ \
    Q double *s_gamma, *s_alpha, *s_beta, *s_nu, *s_delta, *s_phi, *s_pi, *s_ptau; \
    Q int32_t *c_L, *c_W; \
    /* traces: a ring of tr_rows recorded sweeps, iteration-major ([row * n_params + param]), \
     * streamed to host memory in blocks while the chain runs */ \
    Q double *tr_gamma, *tr_alpha, *tr_beta, *tr_ptau, *tr_nu, *tr_delta, *tr_phi, *tr_pi; \
    Q int32_t *tr_L, *tr_W; \
    /* end */

struct ChainDev { BQ_CHAIN_FIELDS() };

#ifdef __HIPCC__
#define BQ_GLOBAL __attribute__((address_space(1)))
// kernel view of ChainDev: the same layout, pointers into the global address space
struct ChainG { BQ_CHAIN_FIELDS(BQ_GLOBAL) };
static_assert(sizeof(ChainG) == sizeof(ChainDev), "ChainG mirrors ChainDev");
#endif

// Host-side initial parameters of one chain (BUSseq_gpu.cu:1163-1244).
struct HostInit {
    std::vector<int32_t> W, L;
    std::vector<double> alpha, beta, nu, delta, phi, gamma, pi, ptau;
    std::vector<double> mu_alpha, mu_nu, mu_delta;
};

// host.cpp: the closed forms of the initial state.  Stage 1 (pi0, W0, delta0)
// needs the library sizes; stage 2 (the rest) the exact raw-mean sums rs.
void bq_init_cells(int B, const int* n_b, int K, const int64_t* lib, uint32_t seed, uint32_t chain, HostInit& o);
void bq_init_params(int B, const int* n_b, int G, int K, const uint64_t* rs, uint32_t seed, uint32_t chain,
                    HostInit& o);

// host.cpp: rows of a streamed trace block (rows x np, row-major) into the
// parameter-major host traces dst (np x P): row q to column col[q], skipped
// where col[q] < 0; esz is 4 or 8 bytes
void bq_trace_transpose(const void* stage, int esz, size_t np, int rows, const int* col, size_t P, void* dst);

// init.hip: the passes over the counts (hipError_t values as int)
// rows: S (Y > 0 -> Y, else 0, padding 0), per-row counts of zero entries and
// of entries >= BQ_HIST, library sizes, static histogram hist[b*BQ_HIST + y]
// of the entries 0 < Y < BQ_HIST (added to, so it starts cleared)
int bq_dev_init_rows(const int32_t* Yd, int N, int G, int Gp, const int32_t* batch, int32_t* S,
                     int32_t* nz_row, int32_t* nbig_row, int64_t* lib, int64_t* hist, void* stream);
// lists: genes of the zero entries of row i at zrow[i], values >= BQ_HIST at brow[i], in gene order
int bq_dev_init_lists(const int32_t* S, int N, int G, int Gp, const int32_t* zrow, const int32_t* brow,
                      uint16_t* zcol, int32_t* big, void* stream);
// per-(batch, gene) histogram hist_bg[(b*G + g)*BQ_PHI_HV + v] += cells with S_ig = v,
// 1 <= v < BQ_PHI_HV (added to, so it starts cleared)
int bq_dev_init_hist_bg(const int32_t* S, int N, int G, int Gp, const int32_t* batch, uint32_t* hist_bg,
                        void* stream);
// the histogram as a list: cnt[j] = distinct counts of item j, then its entries
// (j << 32) | (n << 7) | v from list[off[j]] (off: exclusive prefix of cnt)
int bq_dev_hist_distinct(const uint32_t* hist_bg, long long BG, uint32_t* cnt, void* stream);
int bq_dev_hist_list(const uint32_t* hist_bg, long long BG, const uint64_t* off, uint64_t* list, void* stream);
// raw-mean sums rs[(b*K + k)*G + g] += fix(log1p(S_ig ie_i)) over the cells of each
// chunk (b, first cell, count) with W_i = k (rs starts cleared)
int bq_dev_init_raw(const int32_t* S, int G, int Gp, int K, const int32_t* W, const double* ie,
                    const int32_t* chunks, int n_chunks, uint64_t* rs, void* stream);
