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

#define BQ_KMAX 16     // largest K supported (register arrays in the row kernels)
#define BQ_BMAX 128    // largest B (the reference's __constant__ limit, cu:9-10)
#define BQ_WAVE 64
#define BQ_CHUNK 1024 // genes per row-sum chunk (part of the reduction contract)

struct ChainDev {
    // dimensions and schedule
    int N, G, Gp, K, B;
    uint32_t seed, key;            // Philox key words: (seed, chain id)
    int n_burnin, n_unchanged, n_pres, traces;
    int TR, T_max;                 // rows per tile, tile capacity
    int prec;                      // 32 or 64

    // static data
    const int32_t* batch;          // [N] batch of each cell
    const int32_t* boff;           // [B+1] first cell of each batch
    const int64_t* hist_s;         // [B*H] static histogram of Y>0 entries with Y<H
    const int32_t* zrow;           // [N+1] CSR row offsets of the Y = 0 entries
    const uint16_t* zcol;          // [nz]  their genes
    const int32_t* big_off;        // [B+1] static counts Y >= H, per batch
    const int32_t* big;            // [nbig] their values

    // state
    int* iter;                     // [1] global iteration (sweep) index
    int32_t* S;                    // [N*Gp] encoded X/Z (see numerics contract)
    int32_t* W;                    // [N]
    int32_t* L;                    // [K*G]
    double *alpha, *beta, *nu, *delta, *phi, *gamma, *pi, *ptau;
    double *mu_alpha, *mu_nu, *mu_delta;

    // precision-T copies of the parameters
    void *ea, *eb, *em, *bt, *en, *ed, *ph;

    // proposals and their precision-T constants
    double *alpha_p, *beta_p, *nu_p, *phi_p, *gamma_p;
    void *da, *ca, *db, *cb, *dn, *cn;
    void* pc;                      // [6][B*G]: phi', d, d/phi, 1/phi', log(phi'/phi), lgamma diff

    // plan of the sweep
    int32_t* cnt;                  // [B*K] cells per (batch, subtype)
    int32_t* perm;                 // [N] cells sorted stably by (batch, subtype, index)
    int32_t* tiles;                // [T_max*4] (b, k, start, count)
    int32_t* tile_off;             // [B*K+1] first tile of each (batch, subtype)
    int32_t* ntiles;               // [1]
    uint64_t* part;                // [T_max][3][Gp] exact tile partial sums

    // gamma statistics
    uint32_t* hist_d;              // [B*H] dynamic histogram (entries with Y = 0)
    uint64_t* gstat;               // [B*5] n_z1, sum x z, over0, over1_rej, over1_acc
    double* sp_tab;                // [B*H] log(1 + e^{-(gamma0 + gamma1 x)}) for the current gamma

    int32_t* beta_acc;             // [K*G] 1 where this sweep accepted beta_kg

    // posterior accumulators
    int* n_rec;
    double *s_gamma, *s_alpha, *s_beta, *s_nu, *s_delta, *s_phi, *s_pi, *s_ptau;
    int32_t *c_L, *c_W;
    // traces [param * n_pres + r]
    double *tr_gamma, *tr_alpha, *tr_beta, *tr_ptau, *tr_nu, *tr_delta, *tr_phi, *tr_pi;
    int32_t *tr_L, *tr_W;
};

// Host-side initial state of one chain (BUSseq_gpu.cu:1163-1244).
struct HostInit {
    std::vector<int32_t> S;        // N*Gp
    std::vector<int32_t> W, L;
    std::vector<double> alpha, beta, nu, delta, phi, gamma, pi, ptau;
    std::vector<double> mu_alpha, mu_nu, mu_delta;
    std::vector<int64_t> hist_s;
    std::vector<int32_t> zrow;     // CSR of the Y = 0 entries
    std::vector<uint16_t> zcol;
    std::vector<int32_t> big_off, big;   // static counts >= the histogram size, per batch
};

// host.cpp
int bq_host_init(int B, const int* n_b, int G, int Gp, int K, const int32_t* Y,
                 uint32_t seed, uint32_t chain, HostInit& out);
