// engine.hip -- gfx950 kernels of one BUSseq Gibbs sweep and the host driver
// that allocates the chains, captures the sweep into a hipGraph and replays it.
//
// Sweep (one hipGraph of ten launches in one stream, chains on blockIdx.z):
//   k_w          W_i: categorical full conditional, one wave per cell     [row pass]
//   k_xz         leading blocks: stable sort of cells by (b, W) and tiles (one per
//                batch); L draws, beta, nu, phi proposals.  Then the X/Z imputation
//                of the Y = 0 entries and the gamma histogram             [row pass]
//   k_xz_tail    the X/Z entries with a gamma-Poisson proposal
//   k_alpha      alpha partial sums                                      [tile pass]
//   k_alpha_dec  alpha accept
//   k_beta_nu    leading blocks: gamma0/gamma1 accept per batch, pi_b Dirichlet
//                draw, p and tau0 draw.  Then beta and nu partial sums (nu under
//                both beta outcomes)                                     [tile pass]
//   k_bn_dec     beta accept per (gene, subtype), then nu accept with the
//                accepted beta branch per (gene, batch); a block owns 64 genes
//   k_delta      delta row sum and accept, one wave per cell             [row pass]
//   k_phi        leading blocks: the next sweep's alpha, gamma and delta
//                proposals, X/Z tables and cleared counters.  Then phi partial
//                sums with the new delta                                 [tile pass]
//   k_record     phi accept; posterior accumulators and traces after burn-in;
//                subtype counts cleared; the sweep index advanced (one thread)
// Work that the launch it joins neither reads nor writes for runs in that
// launch's leading blocks, not on a second graph branch: a branch's fork and
// join each cost a cross-queue wait of about 10 us.
//
// Row passes: genes are cut into chunks of BQ_CHUNK = 1024; lane l of the wave
// owning cell i sums the genes g = l (mod 64) of each chunk in double
// precision and the (chunk, lane) sums are added exactly in fixed point.
// Tile passes: a workgroup owns one tile (<= TR cells of one (batch, subtype)
// group, in stable sorted order) and 256 genes; each lane sums its gene over
// the tile in double precision and adds the exact fixed-point value to the
// item's sum with a 64-bit atomic (wrapping integer adds: the order of the
// tiles does not change a bit); the decision kernels read and clear the sums.
// The host oracle uses the same grouping, so the sampler reproduces it bit
// for bit.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/busseq.h"
#include "../../include/busseq/model.h"
#include "engine.h"

#define H BQ_HIST

void bq_set_error(const std::string& m);

// ---------------------------------------------------------------------------
// precision-overloaded element terms
// ---------------------------------------------------------------------------
__device__ __forceinline__ float tmu(float a, float b, float c, float d) { return bq_mu_f32(a, b, c, d); }
__device__ __forceinline__ double tmu(double a, double b, double c, double d) { return bq_mu_f64(a, b, c, d); }
__device__ __forceinline__ float tratio(float m, float p) { return bq_ratio_f32(m, p); }
__device__ __forceinline__ double tratio(double m, double p) { return bq_ratio_f64(m, p); }
__device__ __forceinline__ float tlog1p(float x) { return bq_log1p_mh_f32(x); }
__device__ __forceinline__ double tlog1p(double x) { return bq_log1p(x); }
__device__ __forceinline__ float ttw(float x, float p, float r, float b, float e) { return bq_tw_f32(x, p, r, b, e); }
__device__ __forceinline__ double ttw(double x, double p, double r, double b, double e) { return bq_tw_f64(x, p, r, b, e); }
__device__ __forceinline__ float ttmh(float x, float p, float r, float d, float c) { return bq_tmh_f32(x, p, r, d, c); }
__device__ __forceinline__ double ttmh(double x, double p, double r, double d, double c) { return bq_tmh_f64(x, p, r, d, c); }
__device__ __forceinline__ float ttmh_l(float x, float p, float d, float l) { return bq_tmh_l_f32(x, p, d, l); }
__device__ __forceinline__ double ttmh_l(double x, double p, double d, double l) { return bq_tmh_l_f64(x, p, d, l); }
__device__ __forceinline__ float ttmh2(float x, float p, float r, float d, float c, float l) { return bq_tmh2_f32(x, p, r, d, c, l); }
__device__ __forceinline__ double ttmh2(double x, double p, double r, double d, double c, double l) { return bq_tmh2_f64(x, p, r, d, c, l); }
__device__ __forceinline__ float ttphi_nb(float x, float ph, float pp, float mu, float d, float dop, float ip, float lr)
{ return bq_tphi_nb_f32(x, ph, pp, mu, d, dop, ip, lr); }
__device__ __forceinline__ double ttphi_nb(double x, double ph, double pp, double mu, double d, double dop, double ip, double lr)
{ return bq_tphi_nb_f64(x, ph, pp, mu, d, dop, ip, lr); }
__device__ __forceinline__ float ttphi_lg(float x, float ph, float d, float dl) { return bq_tphi_lg_f32(x, ph, d, dl); }
__device__ __forceinline__ double ttphi_lg(double x, double ph, double d, double dl) { return bq_tphi_lg_f64(x, ph, d, dl); }

// sections of a chain's exact per-item sums (gsum)
__host__ __device__ __forceinline__ size_t gs_beta(int G) { return (size_t)G; }
__host__ __device__ __forceinline__ size_t gs_nu(int G, int K) { return (size_t)G + (size_t)K * G; }
__host__ __device__ __forceinline__ size_t gs_phi(int G, int K, int B) { return gs_nu(G, K) + 2 * (size_t)B * K * G; }
__host__ __device__ __forceinline__ size_t gs_size(int G, int K, int B) { return gs_phi(G, K, B) + (size_t)B * G; }

// a tile value added to an exact per-item sum (no return value: nothing waits on it)
__device__ __forceinline__ void gsum_add(BQ_GLOBAL uint64_t* p, uint64_t v)
{
    if (v) atomicAdd((unsigned long long*)p, (unsigned long long)v);
}

// an exact per-item sum, read and cleared for the next sweep
__device__ __forceinline__ uint64_t gsum_take(BQ_GLOBAL uint64_t* p)
{
    const uint64_t v = *p;
    *p = 0;
    return v;
}

// value of type V at byte offset o (32-bit) from a wave-uniform base: the form
// the compiler emits as a global load with a scalar base and a lane offset
template <typename V>
__device__ __forceinline__ V ld_at(const BQ_GLOBAL void* base, unsigned o)
{
    return *(const BQ_GLOBAL V*)((const BQ_GLOBAL char*)base + o);
}

// (beta_kg, expm1 beta_kg) pairs of the subtype draw (expm1 clamped in single
// precision, tw_em), gene-major: gene g holds
// the pairs of k = 1..K-1 at pair index g * bt_stride(K) + k - 1; the stride is
// K - 1 rounded up to even, so a gene's pairs start 16-byte aligned in single
// precision and are read with one lane offset and immediate offsets.
__host__ __device__ __forceinline__ int bt_stride(int K) { return K & ~1; }

// expm1(beta) as the subtype draw reads it: clamped in single precision (bq_tw_em_f32)
__host__ __device__ __forceinline__ float tw_em(double b, float) { return bq_tw_em_f32(b); }
__host__ __device__ __forceinline__ double tw_em(double b, double) { return bq_expm1(b); }

// value of lane k (a constant or wave-uniform index) in every lane
__device__ __forceinline__ double read_lane(double v, int k)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), k);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// exact wave-wide sum of a fixed-point value (wrapping uint64 adds are associative)
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)v, m, 64);
        uint32_t hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
        v += ((uint64_t)hi << 32) | (uint64_t)lo;
    }
    return v;
}

// exact block-wide sum (256 threads = 4 waves); result valid in every thread
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* sh)
{
    v = wave_sum_u64(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t r = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return r;
}

// block-wide sums of three values (blockDim.x == 256)
__device__ __forceinline__ void block_sum3_u64(uint64_t& a, uint64_t& b, uint64_t& d, uint64_t (*sh)[4])
{
    a = wave_sum_u64(a);
    b = wave_sum_u64(b);
    d = wave_sum_u64(d);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = a;
        sh[1][threadIdx.x >> 6] = b;
        sh[2][threadIdx.x >> 6] = d;
    }
    __syncthreads();
    a = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    b = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    d = (sh[2][0] + sh[2][1]) + (sh[2][2] + sh[2][3]);
}

// atomics on global-space addresses (the cast back to a generic pointer is
// undone by the compiler's address-space inference: the instruction is global)
template <typename V>
__device__ __forceinline__ V gatomic_add(BQ_GLOBAL V* p, V v) { return atomicAdd((V*)p, v); }

// sum of the BQ_STRIPES copies p[r * stride] of a striped counter, every load
// issued before the first add (integer adds: exact in any order)
template <typename V, typename P>
__device__ __forceinline__ V sum_stripes(const P* p, size_t stride)
{
    V v[BQ_STRIPES];
#pragma unroll
    for (int r = 0; r < BQ_STRIPES; ++r) v[r] = (V)p[(size_t)r * stride];
    V acc = 0;
#pragma unroll
    for (int r = 0; r < BQ_STRIPES; ++r) acc += v[r];
    return acc;
}

__device__ __forceinline__ double ph_uniform(const ChainG& c, uint32_t a, uint32_t b, int t, int phase)
{
    bq_rng r;
    bq_rng_init(&r, c.seed, c.key, a, b, (uint32_t)t, (uint32_t)phase);
    return bq_uniform(&r);
}

// ---------------------------------------------------------------------------
// prep_item: what sweep t draws and tabulates ahead of its passes, item j:
// alpha and gamma proposals (cu:514-522, 602-606), delta proposals, counters
// to zero, the table of the X/Z acceptance factors 1 + e^{-(gamma0 + gamma1 x)},
// x < H, and the X/Z step's per-(subtype, batch, gene) mean factors.  It reads
// only parameters final at the end of sweep t - 1, and nothing after k_delta
// in that sweep reads what it writes, so sweep t - 1's k_phi runs it in its
// leading blocks; k_prep runs it once for the first sweep after the chain's
// parameters were set.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void prep_item(const ChainG& c, int j, int t, bool with_tm)
{
    if (j < c.G) {
        bq_rng r;
        bq_rng_init(&r, c.seed, c.key, (uint32_t)j, 0u, (uint32_t)t, BQ_PH_ALPHA_P);
        double a = c.alpha[j];
        double ap = a + BQ_RW_SD * bq_normal(&r);
        c.alpha_p[j] = ap;
        double dl = ap - a;
        ((T*)c.da)[j] = (T)dl;
        ((T*)c.ca)[j] = (T)bq_expm1(dl);
    }
    if (j < c.B) {
        const int B = c.B;
        bq_rng r0;
        bq_rng_init(&r0, c.seed, c.key, (uint32_t)j, 0u, (uint32_t)t, BQ_PH_GAMMA0_P);
        c.gamma_p[j] = c.gamma[j] + BQ_RW_SD * bq_normal(&r0);
        bq_rng r1;
        bq_rng_init(&r1, c.seed, c.key, (uint32_t)j, 0u, (uint32_t)t, BQ_PH_GAMMA1_P);
        c.gamma_p[B + j] = -bq_rgamma(&r1, -10.0 * c.gamma[B + j], 0.1);
        for (int r = 0; r < BQ_STRIPES; ++r)
            for (int q = 0; q < 5; ++q) c.gstat[(size_t)r * B * 5 + j * 5 + q] = 0;
    }
    if (j < c.B * BQ_HLO * BQ_STRIPES) c.hist_lo[j] = 0u;
    if (j < c.B * H) {
        c.hist_d[j] = 0u;
        const int b = j / H, x = j - b * H;
        c.odds_tab[j] = bq_xz_odds(c.gamma[b], c.gamma[c.B + b], (double)x);
    }
    if (j < 2) c.lstat[j] = 0;
    if (j < c.N) {
        // delta proposal and the log of its accept uniform (cu:739-744, 774): one
        // thread per cell here instead of every lane of k_delta's wave; delta is
        // not changed before k_delta reads these
        bq_rng r;
        bq_rng_init(&r, c.seed, c.key, (uint32_t)j, 0u, (uint32_t)t, BQ_PH_DELTA_P);
        const double del = c.delta[j];
        const double dp = del + BQ_RW_SD * bq_normal(&r);
        const double dl = dp - del;
        c.delta_p[j] = dp;
        ((T*)c.dd)[j] = (T)dl;
        ((T*)c.cd)[j] = (T)bq_expm1(dl);
        c.delta_lu[j] = bq_log(ph_uniform(c, (uint32_t)j, 0u, t, BQ_PH_DELTA_A));
    }
    if (with_tm && j < c.K * c.B * c.G) {
        // ((e^alpha e^beta) e^nu): the first two products of the mean, in its order
        // (in the sweep, k_bn_dec writes them once the parameters are decided)
        const int G = c.G, g = j % G, kb = j / G, b = kb % c.B, k = kb / c.B;
        ((T*)c.tm)[j] = (((const T*)c.ea)[g] * ((const T*)c.eb)[(size_t)k * G + g]) * ((const T*)c.en)[(size_t)b * G + g];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_prep(const ChainDev* __restrict__ cds)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    prep_item<T>(c, blockIdx.x * 256 + threadIdx.x, *c.iter, true);
}

// ---------------------------------------------------------------------------
// k_w: subtype draw from the categorical full conditional (replaces the MH
// step of cu:820-854); logit_k = log pi_bk + sum_g [x beta_k - (phi+x) log1p(r0 expm1 beta_k)].
// One wave per cell; KT = K at compile time.
// ---------------------------------------------------------------------------
template <typename T, int KT>
__device__ __forceinline__ void w_cell(const ChainG& c, int i, int lane)
{
    typedef typename std::conditional<sizeof(T) == 4, float2, double2>::type T2;
    const int t = *c.iter;
    const int G = c.G, b = c.batch[i];
    // subtype 0's mean factors (e^alpha e^beta_0) e^nu_b with e^beta_0 = 1: the
    // first two products of tmu(e^alpha, 1, e^nu, e^delta), from the X/Z table
    const BQ_GLOBAL char* t0 = (const BQ_GLOBAL char*)c.tm + (size_t)b * G * sizeof(T);
    const BQ_GLOBAL char* ph = (const BQ_GLOBAL char*)c.ph + (size_t)b * G * sizeof(T);
    const BQ_GLOBAL char* be = (const BQ_GLOBAL char*)c.bt;   // (beta_kg, tw_em(beta_kg)) pairs
    const T edi = ((const T*)c.ed)[i];
    const BQ_GLOBAL char* row = (const BQ_GLOBAL char*)(c.S + (size_t)i * c.Gp);

    uint64_t acc[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) acc[k] = 0;
    for (int c0 = 0; c0 < G; c0 += BQ_CHUNK) {
        const unsigned c1 = (unsigned)(c0 + BQ_CHUNK < G ? c0 + BQ_CHUNK : G);
        T part[KT];   // the lane's terms of the chunk, summed in precision T
#pragma unroll
        for (int k = 0; k < KT; ++k) part[k] = (T)0;
        for (unsigned g = (unsigned)(c0 + lane); g < c1; g += 64) {
            const unsigned o = g * (unsigned)sizeof(T);
            int32_t s = *(const BQ_GLOBAL int32_t*)(row + g * 4u);
            int xi = s < 0 ? ~s : s;
            T x = (T)xi;
            T phi = *(const BQ_GLOBAL T*)(ph + o);
            T2 v[KT];
            const BQ_GLOBAL T* bg = (const BQ_GLOBAL T*)(be + g * (unsigned)(bt_stride(KT) * sizeof(T2)));
#pragma unroll
            for (int k = 1; k < KT; ++k) { v[k].x = bg[2 * k - 2]; v[k].y = bg[2 * k - 1]; }
            T mu0 = *(const BQ_GLOBAL T*)(t0 + o) * edi;   // tmu's last product
            T r0 = tratio(mu0, phi);
#pragma unroll
            for (int k = 1; k < KT; ++k) part[k] += ttw(x, phi, r0, v[k].x, v[k].y);
        }
#pragma unroll
        for (int k = 1; k < KT; ++k) acc[k] += bq_fix((double)part[k]);
    }
    // Subtype k's logit, log pi and exponential are evaluated once, in lane k,
    // instead of in every lane; the maximum, the total and the search then run
    // in subtype order on the values read back from the lanes, as a serial loop.
    double lpl = -BQ_INF;
    {
        uint64_t mine = 0;
#pragma unroll
        for (int k = 1; k < KT; ++k) {
            const uint64_t a = wave_sum_u64(acc[k]);
            if (lane == k) mine = a;
        }
        if (lane < KT) lpl = bq_unfix(mine) + bq_log(c.pi[b * KT + lane]);
    }
    double mx = -BQ_INF;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const double v = read_lane(lpl, k);
        if (v > mx) mx = v;
    }
    const double el = bq_exp(lpl - mx);
    double lp[KT];
    double tot = 0.0;
#pragma unroll
    for (int k = 0; k < KT; ++k) { lp[k] = read_lane(el, k); tot += lp[k]; }
    double tgt = ph_uniform(c, (uint32_t)i, 0u, t, BQ_PH_W) * tot, cum = 0.0;
    int kk = KT - 1, found = 0;
#pragma unroll
    for (int k = 0; k < KT; ++k)
        if (!found) { cum += lp[k]; if (tgt <= cum) { kk = k; found = 1; } }
    if (lane == 0) {
        c.W[i] = kk;
        gatomic_add(&c.cnt[(size_t)(blockIdx.x % BQ_STRIPES) * c.B * KT + b * KT + kk], 1);
    }
}

template <typename T, int KT>
__global__ __launch_bounds__(256) void k_w(const ChainDev* __restrict__ cds)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    // the wave's cell, made wave-uniform so that every row and parameter base
    // is a scalar and each load is base + 32-bit lane offset
    const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= c.N) return;
    w_cell<T, KT>(c, i, threadIdx.x & 63);
}

template <typename T, int KT, int KMAX>
__device__ __forceinline__ void w_dispatch(const ChainG& c, int i, int lane)
{
    if (c.K == KT) { w_cell<T, KT>(c, i, lane); return; }
    if constexpr (KT < KMAX) w_dispatch<T, KT + 1, KMAX>(c, i, lane);
}

// k_w_mixed: the subtype draw of a batch whose chains have different K (2..KMAX)
// in one launch: each block's chain (blockIdx.z) takes its own K's body, the
// same as k_w<T, K>.  Launched once per K, chains of a K-sweep would each
// fill only part of the GPU, one launch after another.
template <typename T, int KMAX>
__global__ __launch_bounds__(256) void k_w_mixed(const ChainDev* __restrict__ cds)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= c.N) return;
    w_dispatch<T, 2, KMAX>(c, i, threadIdx.x & 63);
}

// ---------------------------------------------------------------------------
// k_w_wide: the subtype draw for BQ_KREG < K <= BQ_KMAX.  The same logits
// and the same draw as k_w (each subtype's sum keeps its per-lane, per-chunk
// grouping), computed for BQ_KREG subtypes at a time, one pass over the cell's
// row per group; the logit of subtype k ends in lane k, as in k_w.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_w_wide(const ChainDev* __restrict__ cds)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    const int lane = threadIdx.x & 63;
    const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= c.N) return;
    const int t = *c.iter;
    const int G = c.G, K = c.K, b = c.batch[i];
    const BQ_GLOBAL char* t0 = (const BQ_GLOBAL char*)c.tm + (size_t)b * G * sizeof(T);
    const BQ_GLOBAL char* ph = (const BQ_GLOBAL char*)c.ph + (size_t)b * G * sizeof(T);
    const BQ_GLOBAL T* be = (const BQ_GLOBAL T*)c.bt;
    const unsigned pst = 2u * (unsigned)bt_stride(K);   // T values per gene in the pair table
    const T edi = ((const T*)c.ed)[i];
    const BQ_GLOBAL char* row = (const BQ_GLOBAL char*)(c.S + (size_t)i * c.Gp);
    uint64_t mine = 0;
    for (int k0 = 1; k0 < K; k0 += BQ_KREG) {
        const int nk = K - k0 < BQ_KREG ? K - k0 : BQ_KREG;   // subtypes k0 .. k0 + nk - 1
        uint64_t acc[BQ_KREG];
#pragma unroll
        for (int j = 0; j < BQ_KREG; ++j) acc[j] = 0;
        for (int c0 = 0; c0 < G; c0 += BQ_CHUNK) {
            const unsigned c1 = (unsigned)(c0 + BQ_CHUNK < G ? c0 + BQ_CHUNK : G);
            T part[BQ_KREG];
#pragma unroll
            for (int j = 0; j < BQ_KREG; ++j) part[j] = (T)0;
            for (unsigned g = (unsigned)(c0 + lane); g < c1; g += 64) {
                const unsigned o = g * (unsigned)sizeof(T);
                int32_t s = *(const BQ_GLOBAL int32_t*)(row + g * 4u);
                int xi = s < 0 ? ~s : s;
                T x = (T)xi;
                T phi = *(const BQ_GLOBAL T*)(ph + o);
                const BQ_GLOBAL T* bg = be + g * pst + 2u * (unsigned)(k0 - 1);
                T mu0 = *(const BQ_GLOBAL T*)(t0 + o) * edi;   // tmu's last product
                T r0 = tratio(mu0, phi);
#pragma unroll
                for (int j = 0; j < BQ_KREG; ++j)
                    if (j < nk) part[j] += ttw(x, phi, r0, bg[2 * j], bg[2 * j + 1]);
            }
#pragma unroll
            for (int j = 0; j < BQ_KREG; ++j)
                if (j < nk) acc[j] += bq_fix((double)part[j]);
        }
#pragma unroll
        for (int j = 0; j < BQ_KREG; ++j) {
            if (j < nk) {
                const uint64_t a = wave_sum_u64(acc[j]);
                if (lane == k0 + j) mine = a;
            }
        }
    }
    double lpl = -BQ_INF;
    if (lane < K) lpl = bq_unfix(mine) + bq_log(c.pi[b * K + lane]);
    double mx = -BQ_INF;
    for (int k = 0; k < K; ++k) {
        const double v = read_lane(lpl, k);
        if (v > mx) mx = v;
    }
    const double el = bq_exp(lpl - mx);
    double tot = 0.0;
    for (int k = 0; k < K; ++k) tot += read_lane(el, k);
    double tgt = ph_uniform(c, (uint32_t)i, 0u, t, BQ_PH_W) * tot, cum = 0.0;
    int kk = K - 1;
    for (int k = 0; k < K; ++k) {
        cum += read_lane(el, k);
        if (tgt <= cum) { kk = k; break; }
    }
    if (lane == 0) {
        c.W[i] = kk;
        gatomic_add(&c.cnt[(size_t)(blockIdx.x % BQ_STRIPES) * c.B * K + b * K + kk], 1);
    }
}

template <typename T>
void launch_w(int K, dim3 grid, dim3 blk, hipStream_t st, const ChainDev* cd)
{
    switch (K) {
#define BQ_W_CASE(KV) case KV: hipLaunchKernelGGL((k_w<T, KV>), grid, blk, 0, st, cd); break;
        BQ_W_CASE(2) BQ_W_CASE(3) BQ_W_CASE(4) BQ_W_CASE(5) BQ_W_CASE(6) BQ_W_CASE(7) BQ_W_CASE(8)
        BQ_W_CASE(9) BQ_W_CASE(10) BQ_W_CASE(11) BQ_W_CASE(12) BQ_W_CASE(13) BQ_W_CASE(14)
        BQ_W_CASE(15) BQ_W_CASE(16)
#undef BQ_W_CASE
    default: hipLaunchKernelGGL(k_w_wide<T>, grid, blk, 0, st, cd); break;
    }
}

// ---------------------------------------------------------------------------
// plan_block: stable partition of batch b's cells by subtype (the order every
// tile pass reads them in) and the tile list.  One block of PLAN_THREADS
// threads per batch (the leading blocks of k_xz, which reads neither);
// thread t owns a contiguous run of the batch's cells.  lds: PLAN_LDS_INTS
// ints of the caller's shared memory.
// ---------------------------------------------------------------------------
#define PLAN_THREADS 256
#define PLAN_LDS_INTS (2 * BQ_BKMAX + 1 + (PLAN_THREADS / 64) * BQ_KREG)
__device__ __forceinline__ void plan_block(const ChainG& c, int b, int* lds)
{
    if (b >= c.B) return;
    const int K = c.K, B = c.B, TR = c.TR, tid = threadIdx.x;
    int* s_toff = lds;                                   // BQ_BKMAX + 1
    int* s_cnt = lds + BQ_BKMAX + 1;                     // BQ_BKMAX
    int (*s_wsum)[BQ_KREG] = (int (*)[BQ_KREG])(s_cnt + BQ_BKMAX);
    for (int j = tid; j < B * K; j += PLAN_THREADS) {
        s_cnt[j] = sum_stripes<int>((const int32_t*)c.cnt + j, (size_t)B * K);
    }
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int j = 0; j < B * K; ++j) { s_toff[j] = acc; acc += (s_cnt[j] + TR - 1) / TR; }
        s_toff[B * K] = acc;
    }
    __syncthreads();
    if (b == 0) {
        for (int j = tid; j <= B * K; j += PLAN_THREADS) c.tile_off[j] = s_toff[j];
        if (tid == 0) *c.ntiles = s_toff[B * K];
    }
    // tiles of this batch's groups
    const int b0 = c.boff[b], nb = c.boff[b + 1] - b0;
    int gstart = b0;
    for (int k = 0; k < K; ++k) {
        int n = s_cnt[b * K + k], t0 = s_toff[b * K + k];
        for (int j = tid; j * TR < n; j += PLAN_THREADS) {
            int q = t0 + j;
            c.tiles[4 * q] = b;
            c.tiles[4 * q + 1] = k;
            c.tiles[4 * q + 2] = gstart + j * TR;
            c.tiles[4 * q + 3] = n - j * TR < TR ? n - j * TR : TR;
        }
        gstart += n;
    }
    // stable partition: thread tid owns cells [lo, hi) of the batch; the
    // subtypes are taken BQ_KREG at a time (one pass over the owned cells each)
    const int chunk = (nb + PLAN_THREADS - 1) / PLAN_THREADS;
    const int lo = b0 + tid * chunk;
    const int hi = lo + chunk < b0 + nb ? lo + chunk : b0 + nb;
    const int lane = tid & 63, wv = tid >> 6;
    int base = b0;   // first position of subtype kc's group in the batch's run
    for (int kc = 0; kc < K; kc += BQ_KREG) {
        int my[BQ_KREG];
#pragma unroll
        for (int k = 0; k < BQ_KREG; ++k) my[k] = 0;
        // subtypes of the owned cells, eight loads in flight
        for (int i0 = lo; i0 < hi; i0 += 8) {
            int wb[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) wb[u] = i0 + u < hi ? c.W[i0 + u] - kc : -1;
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int k = 0; k < BQ_KREG; ++k) my[k] += (wb[u] == k);
        }
        // exclusive prefix of the counts over the threads, per subtype: a scan
        // within each wave, then the totals of the waves before
        int incl[BQ_KREG];
#pragma unroll
        for (int k = 0; k < BQ_KREG; ++k) {
            int v = my[k];
            if (kc + k < K) {
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const int o = __shfl_up(v, d);
                    if (lane >= d) v += o;
                }
                if (lane == 63) s_wsum[wv][k] = v;
            }
            incl[k] = v;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < BQ_KREG; ++k) {
            if (kc + k < K) {
                int off = base;
                for (int w = 0; w < wv; ++w) off += s_wsum[w][k];
                my[k] = off + incl[k] - my[k];
                base += s_cnt[b * K + kc + k];
            }
        }
        __syncthreads();   // s_wsum is rewritten by the next chunk
        for (int i0 = lo; i0 < hi; i0 += 8) {
            int wb[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) wb[u] = i0 + u < hi ? c.W[i0 + u] - kc : -1;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (i0 + u >= hi) break;
                int pos = -1;
#pragma unroll
                for (int k = 0; k < BQ_KREG; ++k)
                    if (wb[u] == k) { pos = my[k]; my[k] += 1; }
                if (pos >= 0) c.perm[pos] = i0 + u;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_xz: X/Z imputation (cu:484-512) of the Y = 0 entries, one wave per cell,
// lanes over the cell's static list of zero entries; dynamic gamma histogram.
// The Z draw runs on every entry; entries with Z = 1 are queued in LDS with
// their random-stream state and their negative-binomial proposals are drawn
// 64 at a time, so the proposal path runs with full lanes.  Proposals with
// a small mean are drawn here by CDF inversion; the others (mean > 30 or a
// heavy tail) and entries whose current X is >= H are listed, in the cell's
// own segment of the tail list, for k_xz_tail, which redoes the whole update
// of such an entry with the gamma-Poisson draw.  Every value k_xz writes is
// below H.
// The update of every entry is bq_xz_update; an entry is written back only
// when its state changes.
// ---------------------------------------------------------------------------
template <typename T>
struct XzItem {
    uint32_t pg;           // index of the entry in its cell's zero-entry list (low 16 bits), gene << 16
    int32_t s;
    uint32_t b0, b1;       // the next two words of the entry's stream after its Z draw, if any
    T mu;                  // the proposal mean as computed, in precision T
    double phi;
};
static_assert(sizeof(XzItem<float>) == 32, "32-byte queue entries in single precision");

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// statistics of one updated entry with Z = 1 and value xi >= H for the gamma
// sums (the entries below H go to the histogram)
__device__ __forceinline__ void xz_big_terms(double g0, double g1, double g0p, double g1p, int xi,
                                             uint64_t& o0, uint64_t& o1r, uint64_t& o1a)
{
    const double xd = (double)xi;
    const double a0 = bq_sp_lin(g0, g1, xd), a1 = bq_sp_lin(g0p, g1, xd);
    const double a2 = bq_sp_lin(g0, g1p, xd), a3 = bq_sp_lin(g0p, g1p, xd);
    o0 += bq_fix(a0 - a1);
    o1r += bq_fix(a0 - a2);
    o1a += bq_fix(a1 - a3);
}

static_assert(BQ_HLO == 64, "k_xz keeps one low histogram bin per lane");
#define XZ_SPT 256   // acceptance-factor table entries staged in LDS (every inversion draw is below it)
#define XZ_QCAP 192  // queue entries per wave: two stacks of < 128 entries, fewer than 192 together
#define XZ_MU_SPLIT 6.0  // proposals with a larger mean go to the high stack

template <typename T>
__device__ void draws_block(const ChainG& c, int blk, uint64_t (*s_r3)[4]);

static_assert(sizeof(XzItem<float>) * 4 * XZ_QCAP >= sizeof(int) * PLAN_LDS_INTS,
              "the plan blocks of k_xz use its queue's shared memory");

// Leading blocks: nplan plan blocks (one per batch), then ndraws blocks of the
// draws ahead of the parameter passes; neither reads what the X/Z blocks
// write, nor the other way round.  The rest: four cells per block.
// five waves per SIMD: the proposal step's double-precision polynomials take
// their coefficients as scalar operands (constant tables), which leaves the
// registers of two rounds of gathers in flight within five waves' share
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_xz(
    const ChainDev* __restrict__ cds, int nplan, int ndraws)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    __shared__ uint32_t hist[4][BQ_HLO];
    __shared__ XzItem<T> queue[4][XZ_QCAP];
    __shared__ double s_odds[XZ_SPT];
    __shared__ double s_inv[BQ_INV_N];
    if ((int)blockIdx.x < nplan) {
        plan_block(c, (int)blockIdx.x, (int*)&queue[0][0]);
        return;
    }
    if ((int)blockIdx.x < nplan + ndraws) {
        draws_block<T>(c, (int)blockIdx.x - nplan, (uint64_t (*)[4])&queue[0][0]);
        return;
    }
    const int bx = (int)blockIdx.x - nplan - ndraws;
    const int lane = threadIdx.x & 63;
    // wave and cell are wave-uniform: everything derived from them stays in scalar registers
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = bx * 4 + wv;
    // the grid covers the largest chain of a batch: blocks past this chain's cells leave
    if (bx * 4 >= c.N) return;
    // the acceptance-factor table of the batch of the block's first cell; a wave whose cell
    // lies in another batch reads its own batch's table from memory
    const int bb = c.batch[bx * 4];
    for (int x = threadIdx.x; x < XZ_SPT; x += 256) s_odds[x] = c.odds_tab[(size_t)bb * H + x];
    for (int x = threadIdx.x; x < BQ_INV_N; x += 256) s_inv[x] = bq_inv_tab[x];
    __syncthreads();
    if (i >= c.N) return;
    uint32_t* hw = hist[wv];
    XzItem<T>* qw = queue[wv];
    hw[lane] = 0u;
    const int G = c.G, t = *c.iter;
    const int b = c.batch[i], k = c.W[i];
    const double g0 = c.gamma[b];
    const double qz = 1.0 / (1.0 + bq_exp(g0));
    const T edi = ((const T*)c.ed)[i];
    const T* tmk = (const T*)c.tm + ((size_t)k * c.B + b) * G;
    auto phi = c.phi + (size_t)b * G;
    auto odt = c.odds_tab + (size_t)b * H;
    const bool own = b == bb;
    auto row = c.S + (size_t)i * c.Gp;
    auto zs = c.zs;
    const int z0 = c.zrow[i], z1 = c.zrow[i + 1];
    const uint64_t below = (1ull << lane) - 1ull;
    // Z = 1 entries drawn here and the sum of their values: at most G entries per cell,
    // each below 256, so 32 bits hold the wave's totals
    uint32_t n1 = 0, sx1 = 0;
    int nt = 0;
    const int stripe = bx % BQ_STRIPES;
    auto tl = c.xz_tail + z0;
    // Queued proposals: a batch of 64 runs its CDF search up to the largest
    // draw among its lanes, so proposals are kept in two stacks by their mean:
    // low means grow from qw[0] up, high means from qw[XZ_QCAP - 1] down.
    int nlo = 0, nhi = 0;
    // proposals of the queue entries idx (one per lane < take)
    auto flush = [&](int take, int idx) {
        wave_lds_sync();
        bool tail = false;
        int tp = 0;
        if (lane < take) {
            const XzItem<T> it = qw[idx];
            const int g = (int)(it.pg >> 16), ip = z0 + (int)(it.pg & 0xFFFFu);
            const int xo = it.s < 0 ? ~it.s : 0;
            // the next two words of the entry's stream as bq_xz_update reads
            // them after its Z draw (or, without one, its first two): the
            // inversion's uniform and the acceptance uniform
            const double u = bq_uniform_of(it.b0);
            const double mu = (double)it.mu;
            if (xo < XZ_SPT && bq_nb_inv_ok(mu, it.phi)) {
                // the inversion returns x' <= 255: both acceptance factors are table entries
                const int xp = bq_rnbinom_inv_t(u, mu, it.phi, s_inv);
                const double eo = own ? s_odds[xo] : odt[xo];
                const double ep = own ? s_odds[xp] : odt[xp];
                const int xi = bq_uniform_of(it.b1) * ep <= eo ? xp : xo;
                const int32_t sn = ~xi;
                if (sn != it.s) { row[g] = sn; zs[ip] = sn; }
                n1 += 1;
                sx1 += (uint32_t)xi;
                if (xi < BQ_HLO) atomicAdd(&hw[xi], 1u);
                else gatomic_add(&c.hist_d[b * H + xi], 1u);
            } else {
                tail = true;
                tp = ip;
            }
        }
        const uint64_t tm = __ballot(tail);
        if (tail) tl[nt + __popcll(tm & below)] = tp;
        nt += __popcll(tm);
        wave_lds_sync();
    };
    // entries of the current round with their mean factor and phi, the next
    // round's entries, whose gathers are issued while this round runs, and the
    // entries of the round after it
    int gc = 0, gn = 0;
    int32_t sc = 0, sc1 = 0;
    T mc = (T)0;
    double ph = 0.0;
    if (z0 + lane < z1) { gc = c.zcol[z0 + lane]; sc = zs[z0 + lane]; mc = tmk[gc]; ph = phi[gc]; }
    if (z0 + 64 + lane < z1) { gn = c.zcol[z0 + 64 + lane]; sc1 = zs[z0 + 64 + lane]; }
    for (int base = z0;; base += 64) {
        const bool more = base < z1;
        int g2 = 0;
        int32_t s2 = 0;
        T mcn = (T)0;
        double phn = 0.0;
        if (more) {
            const int p2 = base + 128 + lane;
            if (p2 < z1) { g2 = c.zcol[p2]; s2 = zs[p2]; }
            if (base + 64 + lane < z1) { mcn = tmk[gn]; phn = phi[gn]; }
            const int p = base + lane;
            const bool act = p < z1;
            int z = 0;
            // the first block of the entry's stream, for every listed entry
            // alike (a lane that draws Z and one that does not run the same
            // instructions); the words after the Z draw go to the queue
            bq_u32x4 w = {0u, 0u, 0u, 0u};
            if (act) {
                w = bq_philox4x32_10((uint32_t)i, (uint32_t)gc, (uint32_t)t, (uint32_t)BQ_PH_XZ << 20,
                                     c.seed, c.key);
                if (sc >= -1) {   // X = 0 redraws Z whatever Z was (cu:493-496)
                    z = bq_uniform_of(w.v0) > qz;
                    w.v0 = w.v1; w.v1 = w.v2;
                } else {
                    z = 1;
                }
            }
            const bool zq = act && z;
            const T mu = mc * edi;   // the mean of tmu: ((e^alpha e^beta) e^nu) e^delta
            const bool hq = zq && (double)mu > XZ_MU_SPLIT;
            const uint64_t mlo = __ballot(zq && !hq), mhi = __ballot(hq);
            const uint64_t m0 = __ballot(act && !z);
            if (zq) {
                XzItem<T> it;
                it.pg = (uint32_t)(p - z0) | ((uint32_t)gc << 16);
                it.s = sc;
                it.b0 = w.v0; it.b1 = w.v1;
                it.mu = mu;
                it.phi = ph;
                qw[hq ? XZ_QCAP - 1 - (nhi + __popcll(mhi & below)) : nlo + __popcll(mlo & below)] = it;
            }
            if (act && !z && sc != 0) { row[gc] = 0; zs[p] = 0; }
            if (lane == 0 && m0) atomicAdd(&hw[0], (uint32_t)__popcll(m0));
            nlo += __popcll(mlo);
            nhi += __popcll(mhi);
        }
        // Full batches of either stack (the newest 64 entries of it); after the
        // last round what remains: one batch when it fits, else the high stack
        // first.  One call site, so that the proposal code is emitted once.
        for (;;) {
            int take, idx;
            if (nlo >= 64) { take = 64; idx = nlo - 64 + lane; nlo -= 64; }
            else if (nhi >= 64) { take = 64; idx = XZ_QCAP - nhi + lane; nhi -= 64; }
            else if (!more && nlo + nhi > 64) { take = nhi; idx = XZ_QCAP - nhi + lane; nhi = 0; }
            else if (!more && nlo + nhi > 0) {
                take = nlo + nhi;
                idx = lane < nlo ? lane : XZ_QCAP - nhi + (lane - nlo);
                nlo = nhi = 0;
            } else break;
            flush(take, idx);
        }
        if (!more) break;
        gc = gn; sc = sc1; mc = mcn; ph = phn;
        gn = g2; sc1 = s2;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        n1 += __shfl_xor(n1, m, 64);
        sx1 += __shfl_xor(sx1, m, 64);
    }
    wave_lds_sync();
    {
        const uint32_t v = hw[lane];
        if (v) gatomic_add(&c.hist_lo[(size_t)stripe * c.B * BQ_HLO + b * BQ_HLO + lane], v);
    }
    if (lane == 0) {
        c.xz_ntail[i] = nt;
        unsigned long long* gs = (unsigned long long*)(c.gstat + (size_t)stripe * c.B * 5 + b * 5);
        if (n1) atomicAdd(gs + 0, (unsigned long long)n1);
        if (sx1) atomicAdd(gs + 1, (unsigned long long)sx1);
    }
}

// ---------------------------------------------------------------------------
// k_xz_tail: the entries k_xz left to the gamma-Poisson proposal.  A block
// takes TAIL_CELLS consecutive cells and runs over the concatenation of their
// listed entries with full lanes, 64 entries per wave in turn; the complete
// update bq_xz_update runs from each entry's stream.  Counts and gamma sums
// are gathered per batch in LDS and added once per block.
// ---------------------------------------------------------------------------
#define TAIL_CELLS 32   // cells per block of k_xz_tail (a power of two, at most 64)
// batches a block's TAIL_CELLS consecutive cells can span (every batch has a cell)
#define TAIL_NB (TAIL_CELLS < BQ_BMAX ? TAIL_CELLS : BQ_BMAX)

// three waves per SIMD (a few registers spilled) instead of the two that the
// full update's registers would allow: the waves wait on their double-precision
// chains, and a third per SIMD hides more of that latency
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_xz_tail(const ChainDev* __restrict__ cds)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    __shared__ unsigned long long st[TAIL_NB * 5];
    __shared__ uint32_t hl[TAIL_NB * BQ_HLO];
    __shared__ int incl[TAIL_CELLS];
    const int N = c.N, G = c.G, B = c.B, t = *c.iter;
    const int cb0 = blockIdx.x * TAIL_CELLS;
    if (cb0 >= N) return;
    const int cb1 = (cb0 + TAIL_CELLS < N ? cb0 + TAIL_CELLS : N) - 1;
    const int bmin = c.batch[cb0], nbb = c.batch[cb1] - bmin + 1;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int stripe = blockIdx.x % BQ_STRIPES;
    for (int q = threadIdx.x; q < nbb * 5; q += 256) st[q] = 0ull;
    for (int q = threadIdx.x; q < nbb * BQ_HLO; q += 256) hl[q] = 0u;
    // inclusive prefix of the listed counts over the block's cells
    if (wv == 0) {
        int v = (lane < TAIL_CELLS && cb0 + lane < N) ? c.xz_ntail[cb0 + lane] : 0;
#pragma unroll
        for (int o = 1; o < TAIL_CELLS; o <<= 1) {
            const int u = __shfl_up(v, o, 64);
            if (lane >= o) v += u;
        }
        if (lane < TAIL_CELLS) incl[lane] = v;
    }
    __syncthreads();
    // the block's listed entries in batches of 64, dealt to its waves in turn,
    // so the waves of a block differ by at most one batch
    const int total = incl[TAIL_CELLS - 1];
    for (int j = wv * 64 + lane; j - lane < total; j += 256) {
        if (j >= total) continue;
        // m = the number of cells whose prefix ends at or before entry j: its cell
        int m = 0;
#pragma unroll
        for (int step = TAIL_CELLS / 2; step >= 1; step >>= 1)
            if (incl[m + step - 1] <= j) m += step;
        const int i = cb0 + m;
        const int l = j - (m ? incl[m - 1] : 0);
        const int pz = c.xz_tail[c.zrow[i] + l];
        const int g = c.zcol[pz];
        const int b = c.batch[i], k = c.W[i];
        auto e = c.S + (size_t)i * c.Gp + g;
        const int32_t s = *e;
        const double g0 = c.gamma[b], g1 = c.gamma[B + b];
        const double qz = 1.0 / (1.0 + bq_exp(g0));
        const T mu = tmu(((const T*)c.ea)[g], ((const T*)c.eb)[(size_t)k * G + g],
                         ((const T*)c.en)[(size_t)b * G + g], ((const T*)c.ed)[i]);
        bq_rng r;
        bq_rng_init(&r, c.seed, c.key, (uint32_t)i, (uint32_t)g, (uint32_t)t, BQ_PH_XZ);
        const int32_t sn = bq_xz_update(s, (double)mu, c.phi[(size_t)b * G + g], g0, g1, qz, &r);
        if (sn != s) { *e = sn; c.zs[pz] = sn; }
        const int xi = sn < 0 ? ~sn : 0;
        const int bl = b - bmin;
        if (sn < 0) {
            atomicAdd(&st[bl * 5 + 0], 1ull);
            atomicAdd(&st[bl * 5 + 1], (unsigned long long)xi);
        }
        if (xi < BQ_HLO) {
            atomicAdd(&hl[bl * BQ_HLO + xi], 1u);
        } else if (xi < H) {
            gatomic_add(&c.hist_d[b * H + xi], 1u);
        } else {
            uint64_t o0 = 0, o1r = 0, o1a = 0;
            xz_big_terms(g0, g1, c.gamma_p[b], c.gamma_p[B + b], xi, o0, o1r, o1a);
            atomicAdd(&st[bl * 5 + 2], (unsigned long long)o0);
            atomicAdd(&st[bl * 5 + 3], (unsigned long long)o1r);
            atomicAdd(&st[bl * 5 + 4], (unsigned long long)o1a);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nbb * 5; q += 256)
        if (st[q]) atomicAdd((unsigned long long*)c.gstat + (size_t)stripe * B * 5 + bmin * 5 + q, st[q]);
    for (int q = threadIdx.x; q < nbb * BQ_HLO; q += 256)
        if (hl[q]) gatomic_add(&c.hist_lo[(size_t)stripe * B * BQ_HLO + bmin * BQ_HLO + q], hl[q]);
}

// ---------------------------------------------------------------------------
// The cells of a tile, four at a time: their S entries at gene g and their
// precision-T e^delta are loaded together (four loads in flight), then f(s, e)
// runs for each cell in tile order, so every sum keeps its order.
// ---------------------------------------------------------------------------
template <typename T, typename F>
__device__ __forceinline__ void for_tile_cells(const ChainG& c, int st, int nr, int g, F&& f)
{
    const T* ed = (const T*)c.ed;
    const int32_t* S = (const int32_t*)c.S;
    const size_t Gp = c.Gp;
    int rr = 0;
    for (; rr + 4 <= nr; rr += 4) {
        int ii[4];
        int32_t ss[4];
        T ee[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) ii[u] = c.perm[st + rr + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            ss[u] = S[(size_t)ii[u] * Gp + g];
            ee[u] = ed[ii[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) f(ss[u], ee[u]);
    }
    for (; rr < nr; ++rr) {
        const int i = c.perm[st + rr];
        f(S[(size_t)i * Gp + g], ed[i]);
    }
}

// ---------------------------------------------------------------------------
// k_alpha: alpha partial sums (cu:608-623).  Block = 256 genes x one tile.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_alpha(const ChainDev* __restrict__ cds)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    const int tile = blockIdx.y;
    if (tile >= *c.ntiles || blockIdx.x * 256 >= c.G) return;
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= c.G) return;
    const int G = c.G;
    const int b = c.tiles[4 * tile], k = c.tiles[4 * tile + 1];
    const int st = c.tiles[4 * tile + 2], nr = c.tiles[4 * tile + 3];
    const size_t bg = (size_t)b * G + g;
    const T eag = ((const T*)c.ea)[g], ebg = ((const T*)c.eb)[(size_t)k * G + g];
    const T eng = ((const T*)c.en)[bg], phg = ((const T*)c.ph)[bg];
    const T dag = ((const T*)c.da)[g], cag = ((const T*)c.ca)[g];
    double acc = 0.0;
    for_tile_cells<T>(c, st, nr, g, [&](int32_t s, T edi) {
        const int xi = s < 0 ? ~s : s;
        const T mu = tmu(eag, ebg, eng, edi);
        const T rt = tratio(mu, phg);
        acc += (double)ttmh((T)xi, phg, rt, dag, cag);
    });
    gsum_add(c.gsum + g, bq_fix(acc));
}

// The decision kernels run one block per 64 consecutive items (genes) of an
// item class; its first wave takes the items' exact sums (gsum) and decides.

// nu and phi proposals of k_alpha_dec: item j of G(B-1) (g, b>=1) nu items
// followed by G B (g, b) phi items
template <typename T>
__device__ __forceinline__ void dec_proposals(const ChainG& c, int j, int t)
{
    const int G = c.G, B = c.B;
    if (j < G * (B - 1)) {
        const int b = 1 + j / G, g = j % G;
        const size_t q = (size_t)b * G + g;
        const double nv = c.nu[q];
        bq_rng r;
        bq_rng_init(&r, c.seed, c.key, (uint32_t)g, (uint32_t)b, (uint32_t)t, BQ_PH_NU_P);
        double np = nv + BQ_RW_SD * bq_normal(&r);
        c.nu_p[q] = np;
        double dl = np - nv;
        ((T*)c.dn)[q] = (T)dl;
        ((T*)c.cn)[q] = (T)bq_expm1(dl);
        return;
    }
    j -= G * (B - 1);
    if (j < G * B) {
        const int b = j / G, g = j % G;
        const size_t q = (size_t)b * G + g, BG = (size_t)B * G;
        T* pc = (T*)c.pc;
        bq_rng r;
        bq_rng_init(&r, c.seed, c.key, (uint32_t)g, (uint32_t)b, (uint32_t)t, BQ_PH_PHI_P);
        double f = c.phi[q];
        double fp = bq_rgamma(&r, f, 1.0);
        c.phi_p[q] = fp;
        if (!(fp > 0.0)) fp = f;
        double d = fp - f;
        pc[q] = (T)fp;
        pc[BG + q] = (T)d;
        pc[2 * BG + q] = (T)(d / f);
        pc[3 * BG + q] = (T)(1.0 / fp);
        pc[4 * BG + q] = (T)bq_log1p(d / f);
        pc[5 * BG + q] = (T)bq_lgamma_diff64(f, d);
    }
}

// ---------------------------------------------------------------------------
// k_alpha_dec: alpha accept (cu:625-632), block j deciding genes [64 j, 64 j + 64)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_alpha_dec(const ChainDev* __restrict__ cds)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = blockIdx.x * 64 + lane;
    if (wv != 0 || g >= c.G) return;
    const int t = *c.iter;
    double sum = bq_unfix(gsum_take(c.gsum + g));
    double al = c.alpha[g], ap = c.alpha_p[g];
    double lr = sum + bq_lr_normal(al, ap, c.mu_alpha[g], BQ_SIGMA2_ALPHA);
    if (bq_log(ph_uniform(c, (uint32_t)g, 0u, t, BQ_PH_ALPHA_A)) <= lr) {
        c.alpha[g] = ap;
        ((T*)c.ea)[g] = (T)bq_exp(ap);
    }
}

// ---------------------------------------------------------------------------
// draws_block: flat work items drawn ahead of their passes, in leading blocks
// of k_xz (they read beta, nu, phi, p and tau0, which no kernel changes before
// their first reader, k_beta_nu; the Philox counters are those of their step):
//   G(K-1)  (g, k>=1)    L Gibbs draw (cu:634-648), beta proposal (cu:670-674)
//   G(B-1)  (g, b>=1)    nu proposal (cu:709-713)
//   G B     (g, b)       phi proposal and its per-(b,g) constants (cu:780-787)
// Blocks holding L items add sum L and sum beta^2 over L = 0 for the p and tau0
// draws (cu:654-662) to lstat, which prep_item clears ahead of the sweep.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void draws_block(const ChainG& c, int blk, uint64_t (*s_r3)[4])
{
    const int G = c.G, K = c.K;
    const int t = *c.iter;
    int j = blk * 256 + threadIdx.x;
    // blocks holding L items add sum L and sum beta^2 over L = 0 for the p and
    // tau0 draws (cu:654-662), with the new L and the beta before its update
    const bool lblk = blk * 256 < G * (K - 1);
    uint64_t lsum = 0, bsq = 0;
    if (j < G * (K - 1)) {
        const int k = 1 + j / G, g = j % G;
        const size_t q = (size_t)k * G + g;
        const double be = c.beta[q];
        const int lv = bq_bern_logodds(ph_uniform(c, (uint32_t)g, (uint32_t)k, t, BQ_PH_L),
                                       bq_logodds_L(c.ptau[0], c.ptau[1], be));
        c.L[q] = lv;
        if (lv) lsum = 1;
        else bsq = bq_fix(be * be);
        bq_rng r;
        bq_rng_init(&r, c.seed, c.key, (uint32_t)g, (uint32_t)k, (uint32_t)t, BQ_PH_BETA_P);
        double bp = be + BQ_RW_SD * bq_normal(&r);
        c.beta_p[q] = bp;
        double dl = bp - be;
        ((T*)c.db)[q] = (T)dl;
        ((T*)c.cb)[q] = (T)bq_expm1(dl);
    } else {
        j -= G * (K - 1);
        dec_proposals<T>(c, j, t);
    }
    if (lblk) {
        uint64_t unused = 0;
        block_sum3_u64(lsum, bsq, unused, s_r3);
        if (threadIdx.x == 0) {
            atomicAdd((unsigned long long*)c.lstat, (unsigned long long)lsum);
            atomicAdd((unsigned long long*)c.lstat + 1, (unsigned long long)bsq);
        }
    }
}


// ---------------------------------------------------------------------------
// gamma_ptau_block: gamma0 then gamma1 accept per batch from the X histogram
// (cu:524-600); pi_b ~ Dirichlet(n_b. + xi) (cu:856-890, normalisation race
// fixed); p and tau0 (cu:650-668).  Block j of a chain's 2B + 1: j < B the
// gamma step of batch j, B <= j < 2B the pi draw of batch j - B, j = 2B p and
// tau0.  These blocks lead the grid of k_beta_nu, which reads none of gamma,
// pi, p and tau0 (a second graph branch for them cost two cross-queue waits,
// about as long as the blocks themselves); pi is read only by the next sweep's
// subtype draw and by the recorder.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void gamma_ptau_block(const ChainG& c, int j)
{
    __shared__ uint64_t s_red[3][4];
    __shared__ int s_cnt[BQ_KMAX];
    __shared__ int64_t s_n[H];        // histogram of X over the batch's entries (static + dynamic)
    __shared__ uint64_t s_gs[5];      // the batch's gamma statistics, stripes added
    const int tid = threadIdx.x, B = c.B, t = *c.iter;
    if (j > 2 * B) return;
    if (j == 2 * B) {
        if (t < c.n_unchanged || tid != 0) return;
        const int G = c.G, K = c.K;
        const double n = (double)G * (K - 1);
        const double cnt = (double)(int64_t)c.lstat[0];
        bq_rng r;
        bq_rng_init(&r, c.seed, c.key, 0u, 0u, (uint32_t)t, BQ_PH_PTAU);
        c.ptau[0] = bq_rbeta(&r, cnt + BQ_A_P, n - cnt + BQ_B_P);
        double a = BQ_A_TAU0 + (n - cnt) / 2.0;
        double bb = BQ_B_TAU0 + bq_unfix(c.lstat[1]) / 2.0;
        c.ptau[1] = 1.0 / bq_rgamma(&r, a, 1.0 / bb);
        return;
    }
    if (j >= B) {
        const int b = j - B, K = c.K;
        if (tid >= 64) return;
        if (tid < K) s_cnt[tid] = sum_stripes<int>((const int32_t*)c.cnt + b * K + tid, (size_t)B * K);
        wave_lds_sync();
        if (tid != 0) return;
        bq_rng r;
        bq_rng_init(&r, c.seed, c.key, (uint32_t)b, 0u, (uint32_t)t, BQ_PH_PI);
        // the K gamma draws go to pi first and are divided by their sum in place
        double tot = 0.0;
        for (int k = 0; k < K; ++k) {
            const double gk = bq_rgamma(&r, s_cnt[k] + BQ_XI_PI, 1.0);
            c.pi[b * K + k] = gk;
            tot += gk;
        }
        for (int k = 0; k < K; ++k) c.pi[b * K + k] = c.pi[b * K + k] / tot;
        return;
    }
    const int b = j;
    auto hs = c.hist_s + (size_t)b * H;
    for (int x = tid; x < H; x += 256) {
        int64_t n = hs[x];
        if (x < BQ_HLO) {
            n += sum_stripes<int64_t>((const uint32_t*)c.hist_lo + (size_t)b * BQ_HLO + x, (size_t)B * BQ_HLO);
        } else {
            n += (int64_t)c.hist_d[(size_t)b * H + x];
        }
        s_n[x] = n;
    }
    if (tid < 5) {
        s_gs[tid] = sum_stripes<uint64_t>((const uint64_t*)c.gstat + (size_t)b * 5 + tid, (size_t)B * 5);
    }
    __syncthreads();
    // the gamma0 sum and the gamma1 sums for both outcomes of the gamma0 step
    const double g0 = c.gamma[b], g1 = c.gamma[B + b];
    const double g0p = c.gamma_p[b], g1p = c.gamma_p[B + b];
    uint64_t a0 = 0, a1r = 0, a1a = 0;
    for (int x = tid; x < H; x += 256) {
        const int64_t n = s_n[x];
        if (n) {
            const double xd = (double)x, nd = (double)n;
            const double s00 = bq_sp_lin(g0, g1, xd), s10 = bq_sp_lin(g0p, g1, xd);
            const double s01 = bq_sp_lin(g0, g1p, xd), s11 = bq_sp_lin(g0p, g1p, xd);
            a0 += bq_fix(nd * (s00 - s10));
            a1r += bq_fix(nd * (s00 - s01));
            a1a += bq_fix(nd * (s10 - s11));
        }
    }
    const int bb0 = c.big_off[b], bb1 = c.big_off[b + 1];
    for (int q = bb0 + tid; q < bb1; q += 256) {
        const double xd = (double)c.big[q];
        const double s00 = bq_sp_lin(g0, g1, xd), s10 = bq_sp_lin(g0p, g1, xd);
        const double s01 = bq_sp_lin(g0, g1p, xd), s11 = bq_sp_lin(g0p, g1p, xd);
        a0 += bq_fix(s00 - s10);
        a1r += bq_fix(s00 - s01);
        a1a += bq_fix(s10 - s11);
    }
    block_sum3_u64(a0, a1r, a1a, s_red);
    if (tid == 0) {
        const uint64_t* gs = s_gs;
        a0 += bq_fix((g0p - g0) * (double)(int64_t)gs[0]);
        a0 += gs[2];
        double s0 = bq_unfix(a0);
        double lr0 = s0 + bq_lr_normal(g0, g0p, 0.0, BQ_SIGMA2_GAMMA0);
        const int acc0 = bq_log(ph_uniform(c, (uint32_t)b, 0u, t, BQ_PH_GAMMA0_A)) <= lr0;
        if (acc0) c.gamma[b] = g0p;
        uint64_t a1 = acc0 ? a1a : a1r;
        a1 += bq_fix((g1p - g1) * (double)(int64_t)gs[1]);
        a1 += acc0 ? gs[4] : gs[3];
        double s1 = bq_unfix(a1);
        double u1 = ph_uniform(c, (uint32_t)b, 0u, t, BQ_PH_GAMMA1_A);
        if (g1p < 0.0) {
            double lr1 = s1 + bq_lr_gamma1(-g1, -g1p);
            if (bq_log(u1) <= lr1) c.gamma[B + b] = g1p;
        }
    }
}

// ---------------------------------------------------------------------------
// k_beta_nu: beta (cu:676-694) and nu (cu:715-728) partial sums; the nu sum
// is kept for both outcomes of the beta step of the cell's subtype
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 5))) void k_beta_nu(
    const ChainDev* __restrict__ cds, int gy)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    // the first gy rows of blocks: the gamma, pi, p and tau0 blocks
    if ((int)blockIdx.y < gy) {
        gamma_ptau_block(c, (int)(blockIdx.y * gridDim.x + blockIdx.x));
        return;
    }
    const int tile = blockIdx.y - gy;
    if (tile >= *c.ntiles || blockIdx.x * 256 >= c.G) return;
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= c.G) return;
    const int G = c.G;
    const int b = c.tiles[4 * tile], k = c.tiles[4 * tile + 1];
    const int st = c.tiles[4 * tile + 2], nr = c.tiles[4 * tile + 3];
    const size_t bg = (size_t)b * G + g, kg = (size_t)k * G + g;
    const T eag = ((const T*)c.ea)[g], ebg = ((const T*)c.eb)[kg];
    const T eng = ((const T*)c.en)[bg], phg = ((const T*)c.ph)[bg];
    const T dbg = ((const T*)c.db)[kg], cbg = ((const T*)c.cb)[kg];
    const T dng = ((const T*)c.dn)[bg], cng = ((const T*)c.cn)[bg];
    const T cbn = (cbg + cng) + cbg * cng;
    double pb = 0.0, pn0 = 0.0, pn1 = 0.0;
    for_tile_cells<T>(c, st, nr, g, [&](int32_t s, T edi) {
        const int xi = s < 0 ? ~s : s;
        const T x = (T)xi;
        const T mu = tmu(eag, ebg, eng, edi);
        const T rt = tratio(mu, phg);
        T lb = (T)0;
        if (k >= 1) {
            lb = tlog1p(rt * cbg);
            pb += (double)ttmh_l(x, phg, dbg, lb);
        }
        if (b >= 1) {
            pn0 += (double)ttmh(x, phg, rt, dng, cng);
            if (k >= 1) pn1 += (double)ttmh2(x, phg, rt, dng, cbn, lb);
        }
    });
    const int K = c.K;
    const size_t bkg = ((size_t)b * K + k) * G + g, BKG = (size_t)c.B * K * G;
    if (k >= 1) gsum_add(c.gsum + gs_beta(G) + kg, bq_fix(pb));
    if (b >= 1) {
        gsum_add(c.gsum + gs_nu(G, K) + bkg, bq_fix(pn0));
        if (k >= 1) gsum_add(c.gsum + gs_nu(G, K) + BKG + bkg, bq_fix(pn1));
    }
}

// ---------------------------------------------------------------------------
// k_bn_dec: beta accept (cu:696-707) for the items (k >= 1, g), then nu accept
// (cu:730-737) for the items (b >= 1, g) with the nu sums of each subtype's
// accepted beta branch.  Block j owns genes [32 j, 32 j + 32): the nu step reads
// only its own genes' beta outcomes, written before the block barrier.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void beta_decide(const ChainG& c, int k, int g, int t)
{
    const int G = c.G, K = c.K;
    const size_t q = (size_t)k * G + g;
    double sum = bq_unfix(gsum_take(c.gsum + gs_beta(G) + q));
    double be = c.beta[q], bp = c.beta_p[q];
    double lr = sum + bq_lr_beta(be, bp, c.L[q], c.ptau[1]);
    int acc = bq_log(ph_uniform(c, (uint32_t)g, (uint32_t)k, t, BQ_PH_BETA_A)) <= lr;
    c.beta_acc[q] = acc;
    if (acc) {
        c.beta[q] = bp;
        ((T*)c.eb)[q] = (T)bq_exp(bp);
        const size_t pq = (size_t)g * bt_stride(K) + (k - 1);
        ((T*)c.bt)[2 * pq] = (T)bp;
        ((T*)c.bt)[2 * pq + 1] = tw_em(bp, T());
    }
}

template <typename T>
__device__ __forceinline__ void nu_decide(const ChainG& c, int b, int g, int t)
{
    const int G = c.G, K = c.K, B = c.B;
    // each subtype group's nu sum under its accepted beta branch; both cleared
    const size_t BKG = (size_t)B * K * G;
    BQ_GLOBAL uint64_t* ns = c.gsum + gs_nu(G, K) + (size_t)b * K * G + g;
    uint64_t a = 0;
    for (int k = 0; k < K; ++k) {
        const uint64_t a0 = gsum_take(ns + (size_t)k * G), a1 = gsum_take(ns + BKG + (size_t)k * G);
        a += (k >= 1 && c.beta_acc[(size_t)k * G + g]) ? a1 : a0;
    }
    const size_t q = (size_t)b * G + g;
    double sum = bq_unfix(a);
    double nv = c.nu[q], np = c.nu_p[q];
    double lr = sum + bq_lr_normal(nv, np, c.mu_nu[b], BQ_SIGMA2_NU);
    if (bq_log(ph_uniform(c, (uint32_t)g, (uint32_t)b, t, BQ_PH_NU_A)) <= lr) {
        c.nu[q] = np;
        ((T*)c.en)[q] = (T)bq_exp(np);
    }
}

#define BN_GENES 32   // genes per block of k_bn_dec: (K - 1) * 32 beta items fit one pass for K <= 9
template <typename T>
__global__ __launch_bounds__(256) void k_bn_dec(const ChainDev* __restrict__ cds)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    const int G = c.G, K = c.K, B = c.B;
    const int g0 = blockIdx.x * BN_GENES;
    if (g0 >= G) return;
    const int t = *c.iter;
    if (blockIdx.x == 0 && threadIdx.x == 0) *c.iter_rec = t;   // for k_record
    // thread i takes the items i, i + 256, ... of the block's (k, gene) and then
    // (b, gene) items, gene fastest
    for (int q = threadIdx.x; q < (K - 1) * BN_GENES; q += 256) {
        const int g = g0 + q % BN_GENES;
        if (g < G) beta_decide<T>(c, 1 + q / BN_GENES, g, t);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < (B - 1) * BN_GENES; q += 256) {
        const int g = g0 + q % BN_GENES;
        if (g < G) nu_decide<T>(c, 1 + q / BN_GENES, g, t);
    }
    __syncthreads();
    // the mean factors (e^alpha e^beta_k) e^nu_b of the block's genes from the
    // decided parameters, none of which changes again this sweep: k_delta's means
    // (times e^delta) and the next sweep's X/Z table
    const T* ea = (const T*)c.ea;
    const T* eb = (const T*)c.eb;
    const T* en = (const T*)c.en;
    for (int q = threadIdx.x; q < K * B * BN_GENES; q += 256) {
        const int g = g0 + q % BN_GENES, kb = q / BN_GENES, b = kb % B, k = kb / B;
        if (g < G) ((T*)c.tm)[(size_t)kb * G + g] = (ea[g] * eb[(size_t)k * G + g]) * en[(size_t)b * G + g];
    }
}

// ---------------------------------------------------------------------------
// k_delta: delta proposal (cu:739-744), row sum (cu:746-761), accept
// (cu:763-778); the first cell of each batch keeps delta = 0
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_delta(const ChainDev* __restrict__ cds)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    const int lane = threadIdx.x & 63;
    // wave-uniform cell: row and parameter bases are scalars, loads take 32-bit lane offsets
    const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= c.N) return;
    const int b = c.batch[i];
    if (i == c.boff[b]) return;
    const int G = c.G, k = c.W[i];
    // (e^alpha e^beta_k) e^nu_b of the cell's group, written by k_bn_dec
    const BQ_GLOBAL T* tm = (const BQ_GLOBAL T*)c.tm + ((size_t)k * c.B + b) * G;
    const BQ_GLOBAL T* ph = (const BQ_GLOBAL T*)c.ph + (size_t)b * G;
    BQ_GLOBAL T* ed = (BQ_GLOBAL T*)c.ed;
    const T edi = ed[i];
    const double del = c.delta[i], dp = c.delta_p[i];
    const T dd = ((const BQ_GLOBAL T*)c.dd)[i], cd = ((const BQ_GLOBAL T*)c.cd)[i];
    auto row = c.S + (size_t)i * c.Gp;
    uint64_t a = 0;
    for (int c0 = 0; c0 < G; c0 += BQ_CHUNK) {
        const int c1 = c0 + BQ_CHUNK < G ? c0 + BQ_CHUNK : G;
        double part = 0.0;
        // four genes of the lane class per step, loaded together, added in gene order
        for (unsigned g0 = (unsigned)(c0 + lane); g0 < (unsigned)c1; g0 += 256) {
            int32_t sv[4];
            T pv[4], mv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const unsigned g = g0 + 64u * u, o = g * (unsigned)sizeof(T);
                if (g < (unsigned)c1) {
                    sv[u] = ld_at<int32_t>(row, g * 4u);
                    pv[u] = ld_at<T>(ph, o); mv[u] = ld_at<T>(tm, o);
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (g0 + 64 * u < c1) {
                    const int xi = sv[u] < 0 ? ~sv[u] : sv[u];
                    const T mu = mv[u] * edi;   // tmu's last product
                    const T rt = tratio(mu, pv[u]);
                    part += (double)ttmh((T)xi, pv[u], rt, dd, cd);
                }
            }
        }
        a += bq_fix(part);
    }
    a = wave_sum_u64(a);
    double sum = bq_unfix(a);
    double lr = sum + bq_lr_normal(del, dp, c.mu_delta[i], BQ_SIGMA2_DELTA);
    if (c.delta_lu[i] <= lr) {
        if (lane == 0) {
            c.delta[i] = dp;
            ed[i] = (T)bq_exp(dp);
        }
    }
}

// ---------------------------------------------------------------------------
// k_phi: phi partial sums (cu:789-804) with the new delta.  Each cell's term is
// its NB part (bq_tphi_nb, summed per tile) plus, for a count x >= 1, its
// lgamma part (bq_tphi_lg), which depends on x, phi and phi' only and enters
// the item's exact fixed-point sum on its own.  Leading blocks: the first gy
// rows run prep_item for the next sweep (they read and write nothing the phi
// sums and k_record use); the next gyh rows add the lgamma parts of the counts
// 1 <= Y < BQ_PHI_HV from the static per-(batch, gene) histogram, n fix(D(v))
// for each of its nonzero entries (v, n), 64 entries per wave, whatever the
// items' sizes.  The tile blocks queue the other counts with x >= 1 (X of the Y = 0
// entries, and Y >= BQ_PHI_HV) and evaluate their lgamma parts 64 at a time.
// ---------------------------------------------------------------------------
#define PHI_QCAP 128   // queued counts per wave: fewer than 64 wait, each cell row adds at most 64

template <typename T>
// double precision at three waves per SIMD: its lgamma difference keeps more
// values live than five (or four) waves' registers hold without spills
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 3))) void k_phi(
    const ChainDev* __restrict__ cds, int gy, int gyh)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    __shared__ uint32_t s_q[4][PHI_QCAP];          // (x << 6) | lane of the queued counts
    __shared__ T s_par[4][3][64];                  // phi, d, lgamma(phi') - lgamma(phi) of the wave's genes
    __shared__ unsigned long long s_acc[4][64];    // fixed-point lgamma parts of the wave's genes
    if ((int)blockIdx.y < gy) {
        prep_item<T>(c, (int)(blockIdx.y * gridDim.x + blockIdx.x) * 256 + (int)threadIdx.x, *c.iter + 1, false);
        return;
    }
    const int G = c.G;
    const size_t BG = (size_t)c.B * G;
    const T* pc = (const T*)c.pc;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if ((int)blockIdx.y < gy + gyh) {
        // 64 consecutive entries (j, v, n) of the static list per wave: n fix(D(v))
        // each, summed over the wave's run of each item j and added once per run
        const long long e0 = ((long long)(blockIdx.y - gy) * gridDim.x + blockIdx.x) * 256 + wv * 64;
        if (e0 >= c.n_hlist) return;   // wave-uniform
        uint32_t j = 0xFFFFFFFFu;      // past the list: a run of its own that adds nothing
        uint64_t a = 0;
        if (e0 + lane < c.n_hlist) {
            const uint64_t w = c.hlist[e0 + lane];
            j = (uint32_t)(w >> 32);
            const int v = (int)(w & 127u);
            const uint32_t n = (uint32_t)(w >> 7) & 0x1FFFFFFu;
            a = bq_fix((double)ttphi_lg((T)v, ((const T*)c.ph)[j], pc[BG + j], pc[5 * BG + j])) * (uint64_t)n;
        }
        // inclusive sums within runs of equal j (a run's lanes are contiguous)
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t lo = __shfl_up((uint32_t)a, o, 64), hi = __shfl_up((uint32_t)(a >> 32), o, 64);
            const uint32_t jo = __shfl_up(j, o, 64);
            if (lane >= o && jo == j) a += ((uint64_t)hi << 32) | lo;
        }
        const uint32_t jn = __shfl_down(j, 1, 64);
        if (j != 0xFFFFFFFFu && (lane == 63 || jn != j)) gsum_add(c.gsum + gs_phi(G, c.K, c.B) + j, a);
        return;
    }
    const int tile = blockIdx.y - gy - gyh;
    // a wave of genes past G leaves; in the last wave of genes, the lanes past G
    // (live = false) read gene G - 1's parameters and the row's padding, add
    // nothing, and take part in the queue's drains
    if (tile >= *c.ntiles || blockIdx.x * 256 + wv * 64 >= G) return;
    const int g0 = blockIdx.x * 256 + threadIdx.x;
    const bool live = g0 < G;
    const int g = live ? g0 : G - 1;
    const int b = c.tiles[4 * tile], k = c.tiles[4 * tile + 1];
    const int st = c.tiles[4 * tile + 2], nr = c.tiles[4 * tile + 3];
    const size_t bg = (size_t)b * G + g;
    const T eag = ((const T*)c.ea)[g], ebg = ((const T*)c.eb)[(size_t)k * G + g];
    const T eng = ((const T*)c.en)[bg], phg = ((const T*)c.ph)[bg];
    const T fp = pc[bg], d = pc[BG + bg], dop = pc[2 * BG + bg], ip = pc[3 * BG + bg];
    const T lrat = pc[4 * BG + bg], dlg0 = pc[5 * BG + bg];
    uint32_t* qw = s_q[wv];
    T (*pw)[64] = s_par[wv];
    unsigned long long* aw = s_acc[wv];
    pw[0][lane] = phg; pw[1][lane] = d; pw[2][lane] = dlg0;
    aw[lane] = 0ull;
    const uint64_t below = (1ull << lane) - 1ull;
    int nq = 0;           // wave-uniform queue length
    uint64_t lgx = 0;     // lgamma parts of counts too large for the queue's 26 bits, added here
    // the newest `take` (<= 64) queued counts, one per lane
    auto drain = [&](int take) {
        wave_lds_sync();
        if (lane < take) {
            const uint32_t e = qw[nq - take + lane];
            const int gl = (int)(e & 63u);
            const T lg = ttphi_lg((T)(int)(e >> 6), pw[0][gl], pw[1][gl], pw[2][gl]);
            atomicAdd(&aw[gl], (unsigned long long)bq_fix((double)lg));
        }
        nq -= take;
        wave_lds_sync();
    };
    wave_lds_sync();
    double acc = 0.0;
    for_tile_cells<T>(c, st, nr, g, [&](int32_t s, T edi) {
        const int xi = s < 0 ? ~s : s;
        const T mu = tmu(eag, ebg, eng, edi);
        acc += (double)ttphi_nb((T)xi, phg, fp, mu, d, dop, ip, lrat);
        // counts the histogram does not hold: X of a Y = 0 entry (S < 0 with X >= 1), Y >= BQ_PHI_HV
        const bool other = live && (s < -1 || s >= BQ_PHI_HV);
        const bool q = other && xi < (1 << 26);
        if (other && !q) lgx += bq_fix((double)ttphi_lg((T)xi, phg, d, dlg0));
        const uint64_t m = __ballot(q);
        if (q) qw[nq + __popcll(m & below)] = ((uint32_t)xi << 6) | (uint32_t)lane;
        nq += __popcll(m);
        if (nq >= 64) drain(64);
    });
    if (nq > 0) drain(nq);
    if (live) gsum_add(c.gsum + gs_phi(G, c.K, c.B) + bg, bq_fix(acc) + (uint64_t)aw[lane] + lgx);
}

// ---------------------------------------------------------------------------
// k_record: phi accept, then posterior accumulators and traces (cu:892-936,
// 1547-1554)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void record_sweep(const ChainG& c, int t, int j, int G, int K, int B, int N, int P)
{
    // recorded sweep r = t - n_burnin (t >= n_burnin here) goes to ring row t mod tr_rows
    const bool tr = c.traces && t - c.n_burnin < P;
    const int r = tr ? t % c.tr_rows : 0;
    if (j == 0) *c.n_rec += 1;
    if (j < 2 * B) {
        c.s_gamma[j] += c.gamma[j];
        if (tr) c.tr_gamma[(size_t)r * (2 * B) + j] = c.gamma[j];
    }
    if (j < 2) {
        c.s_ptau[j] += c.ptau[j];
        if (tr) c.tr_ptau[(size_t)r * 2 + j] = c.ptau[j];
    }
    if (j < G) {
        c.s_alpha[j] += c.alpha[j];
        if (tr) c.tr_alpha[(size_t)r * G + j] = c.alpha[j];
    }
    if (j < K * G) {
        c.s_beta[j] += c.beta[j];
        c.c_L[j] += c.L[j];
        if (tr && j >= G) {
            c.tr_beta[(size_t)r * ((K - 1) * G) + (j - G)] = c.beta[j];
            c.tr_L[(size_t)r * ((K - 1) * G) + (j - G)] = c.L[j];
        }
    }
    if (j < B * G) {
        c.s_nu[j] += c.nu[j];
        c.s_phi[j] += c.phi[j];
        if (tr) {
            if (j >= G) c.tr_nu[(size_t)r * ((B - 1) * G) + (j - G)] = c.nu[j];
            c.tr_phi[(size_t)r * (B * G) + j] = c.phi[j];
        }
    }
    if (j < N) {
        c.s_delta[j] += c.delta[j];
        c.c_W[(size_t)j * K + c.W[j]] += 1;
        if (tr) {
            c.tr_delta[(size_t)r * N + j] = c.delta[j];
            c.tr_W[(size_t)r * N + j] = c.W[j];
        }
    }
    if (j < B * K) {
        c.s_pi[j] += c.pi[j];
        if (tr) c.tr_pi[(size_t)r * (B * K) + j] = c.pi[j];
    }
}



// phi accept (cu:806-818) of item j = b G + g
template <typename T>
__device__ __forceinline__ void phi_decide(const ChainG& c, int j, int t)
{
    const int G = c.G, B = c.B, K = c.K;
    const int b = j / G, g = j - b * G;
    double sum = bq_unfix(gsum_take(c.gsum + gs_phi(G, K, B) + j));
    double f = c.phi[j], fp = c.phi_p[j];
    double u = ph_uniform(c, (uint32_t)g, (uint32_t)b, t, BQ_PH_PHI_A);
    if (!(fp > 0.0)) return;
    double lr = sum + bq_lr_phi(f, fp);
    if (bq_log(u) <= lr) {
        c.phi[j] = fp;
        ((T*)c.ph)[j] = (T)fp;
    }
}

// Thread j decides phi item j (the last step of the sweep), then records; the
// subtype counts of this sweep (read by the plan and the pi draw) are cleared
// for the next one.  The sweep index comes from iter_rec, which k_bn_dec
// copied from iter, so one thread can advance iter without racing the other blocks'
// reads (a last-block count would need a device-wide fence in every block).
template <typename T>
__global__ __launch_bounds__(256) void k_record(const ChainDev* __restrict__ cds)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    const int t = *c.iter_rec;
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int G = c.G, K = c.K, B = c.B, N = c.N, P = c.n_pres;
    if (j < B * G) phi_decide<T>(c, j, t);
    if (j < B * K * BQ_STRIPES) c.cnt[j] = 0;
    if (t >= c.n_burnin) record_sweep(c, t, j, G, K, B, N, P);
    if (j == 0) *c.iter = t + 1;
}

// ---------------------------------------------------------------------------
// BIC: per (cell, subtype) log-likelihood row sums (cu:967-1007), one wave each
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bic(const ChainDev* __restrict__ cds, const double* __restrict__ pm,
                                             double* __restrict__ ll)
{
    // pm: posterior means packed as alpha[G] | beta[K*G] | nu[B*G] | delta[N] | phi[B*G] | gamma[2B]
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int G = c.G, K = c.K, B = c.B, N = c.N;
    if (row >= N * K) return;
    const int i = row / K, k = row % K, b = c.batch[i];
    const double* al = pm;
    const double* be = al + G;
    const double* nu = be + (size_t)K * G;
    const double* de = nu + (size_t)B * G;
    const double* phi = de + N;
    const double* ga = phi + (size_t)B * G;
    auto Srow = c.S + (size_t)i * c.Gp;
    uint64_t a = 0;
    for (int c0 = 0; c0 < G; c0 += BQ_CHUNK) {
        const int c1 = c0 + BQ_CHUNK < G ? c0 + BQ_CHUNK : G;
        double part = 0.0;
        for (int g = c0 + lane; g < c1; g += 64) {
            int32_t s = Srow[g];
            int y = s > 0 ? s : 0;
            part += bq_bic_elem(al[g], be[(size_t)k * G + g], nu[(size_t)b * G + g], de[i], ga[b], ga[B + b],
                                phi[(size_t)b * G + g], y);
        }
        a += bq_fix(part);
    }
    a = wave_sum_u64(a);
    if (lane == 0) ll[(size_t)i * K + k] = bq_unfix(a);
}

// Per cell: the subtypes' row sums mixed with the batch's pi (cu:1009-1024),
// log sum_k pi_bk e^{ll_k}, in place of the cell's first row sum.
__global__ __launch_bounds__(256) void k_bic_cells(const ChainDev* __restrict__ cds, const double* __restrict__ pi,
                                                   double* __restrict__ ll)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int K = c.K;
    if (i >= c.N) return;
    const int b = c.batch[i];
    double* li = ll + (size_t)i * K;
    double mx = -BQ_INF;
    for (int k = 0; k < K; ++k) if (li[k] > mx) mx = li[k];
    double sm = 0.0;
    for (int k = 0; k < K; ++k) sm += pi[(size_t)b * K + k] * bq_exp(li[k] - mx);
    li[0] = mx + bq_log(sm);
}

// The cells' values added in cell order (one wave: 64 loaded together, added
// one by one from the lanes) and the penalty of cu:1833; *out = BIC.
__global__ __launch_bounds__(64) void k_bic_total(const ChainDev* __restrict__ cds, const double* __restrict__ ll,
                                                  double* __restrict__ out)
{
    const ChainG& c = ((const ChainG*)cds)[blockIdx.z];
    const int lane = threadIdx.x;
    const int N = c.N, G = c.G, K = c.K, B = c.B;
    double tot = 0.0;
    for (int i0 = 0; i0 < N; i0 += 64) {
        const double v = i0 + lane < N ? ll[(size_t)(i0 + lane) * K] : 0.0;
        const int n = N - i0 < 64 ? N - i0 : 64;
        for (int j = 0; j < n; ++j) tot += read_lane(v, j);
    }
    const double npar = (double)(B + G) * K + 2.0 * B + (double)G * (B * 2 - 1) + N - B;
    if (lane == 0) *out = -2.0 * tot + bq_log((double)G * (double)N) * npar;
}
// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
#define HIPCHK(x)                                                                     \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            bq_set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
            return BUSSEQ_ERR_DEVICE;                                                 \
        }                                                                             \
    } while (0)

// host-side description of one chain
struct ChainHost {
    int problem;
    std::vector<int> n_b;
    int N, G, Gp, K, B;
    int T_alloc;                   // tiles the partial-sum buffer holds
};

// Streamed traces (-p): each chain's device ring holds 2 blocks of
// BQ_TRACE_BLOCK sweeps.  When the sweeps of block k are enqueued, its rows are
// copied (on the copy stream, after them) into pinned host memory, and the
// previous block, whose copy has finished by then, is transposed into the
// host traces while the GPU runs block k.
#define BQ_TRACE_BLOCK 32      // sweeps per streamed block (a multiple of the 4-sweep graph)
#define BQ_TRACE_ARRAYS 10     // gamma alpha L beta ptau nu delta phi W pi
struct TraceHost {
    int P = 0, nb = 0;                       // recorded sweeps kept, burn-in
    size_t np[BQ_TRACE_ARRAYS] = {};         // parameters per row
    int esz[BQ_TRACE_ARRAYS] = {};           // bytes per value
    void* dev[BQ_TRACE_ARRAYS] = {};         // device ring (2 blocks of rows)
    void* stage[BQ_TRACE_ARRAYS] = {};       // pinned copy of the ring
    void* host[BQ_TRACE_ARRAYS] = {};        // [param * P + iter]
};

struct busseq_sampler {
    int C = 0;                     // chains
    int prec = 32, TR = 64, dev = 0;
    int maxN = 0, maxG = 0, maxK = 0, maxB = 0, maxT = 0, maxPrep = 0, maxRec = 0;
    int maxDraws = 0;              // draws blocks (leading blocks of k_xz)
    std::vector<ChainHost> ch;
    std::vector<ChainDev> hd;
    ChainDev* d_cd = nullptr;
    std::vector<std::pair<int, int>> krange;   // runs of consecutive chains with equal K
    hipStream_t stream = nullptr;
    hipGraph_t graph[2] = {};      // one sweep, BQ_GRAPH_SWEEPS sweeps
    hipGraphExec_t gexec[2] = {};
    std::vector<void*> allocs;
    int iter = 0;                  // sweeps enqueued
    bool oom = false;              // some dalloc failed
    // trace streaming
    std::vector<TraceHost> th;     // per chain (P = 0: no traces)
    bool tracing = false;
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_sweeps = nullptr, ev_copy[2] = {};
    bool copy_recorded[2] = {};    // ev_copy[h] has been recorded (streams wait only on recorded events)
    int pend_block = -1, pend_rows = 0;   // block whose copy is in flight, not yet transposed

    template <typename V>
    V* dalloc(size_t n)
    {
        void* p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(V)) != hipSuccess) { oom = true; return nullptr; }
        allocs.push_back(p);
        if (hipMemset(p, 0, n * sizeof(V)) != hipSuccess) { oom = true; return nullptr; }
        return (V*)p;
    }
};

namespace {

static const char* const kKernelNames[] = {
    "k_w", "k_xz", "k_xz_tail", "k_alpha", "k_alpha_dec",
    "k_beta_nu", "k_bn_dec", "k_delta", "k_phi", "k_record"};
static const int kNumKernels = 10;

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

// one sweep of every chain; ev: optional kNumKernels + 1 events around the launches
template <typename T>
busseq_status enqueue_sweep(busseq_sampler* s, hipEvent_t* ev = nullptr)
{
    const dim3 blk(256);
    const int C = s->C;
    hipStream_t st = s->stream;
    const ChainDev* cd = s->d_cd;
    const dim3 g_rows((s->maxN + 3) / 4, 1, C);
    const dim3 g_tile(nblk(s->maxG), s->maxT, C);
    const dim3 g_adec((s->maxG + 63) / 64, 1, C);
    // k_xz: the plan blocks (one per batch) and the draws blocks lead the cells' blocks
    const dim3 g_xz(s->maxB + s->maxDraws + (s->maxN + 3) / 4, 1, C);
    // rows of k_beta_nu blocks that carry the 2B + 1 gamma, pi, p/tau0 blocks of each chain
    const int gy_gamma = (int)((2 * s->maxB + 1 + g_tile.x - 1) / g_tile.x);
    // rows of k_phi blocks that carry the next sweep's prep items
    const int gy_prep = (int)((nblk(s->maxPrep) + g_tile.x - 1) / g_tile.x);
    // rows of k_phi blocks that carry the entries of each chain's histogram list, 256 per block
    long long max_hl = 0;
    for (int c = 0; c < C; ++c) max_hl = std::max(max_hl, s->hd[c].n_hlist);
    const int gy_hist = (int)((max_hl + 256LL * g_tile.x - 1) / (256LL * g_tile.x));
    const dim3 g_rec(nblk(s->maxRec), 1, C);
    int q = 0;
    // One stream, no branches: a second branch of the captured graph costs a
    // cross-queue wait of about 10 us at each fork and join, more than the short
    // kernels it would hide, so their work runs as leading blocks of the passes
    // beside which it may run.  With ev (per-kernel timing) events bracket each.
#define EV() do { if (ev) HIPCHK(hipEventRecord(ev[q], st)); ++q; } while (0)
    EV();
    if (s->krange.size() > 1 && s->maxK <= BQ_KREG) {
        // chains of several K, all within the specialised bodies: one launch
        if (s->maxK <= 10) hipLaunchKernelGGL((k_w_mixed<T, 10>), g_rows, blk, 0, st, cd);
        else hipLaunchKernelGGL((k_w_mixed<T, BQ_KREG>), g_rows, blk, 0, st, cd);
    } else {
        for (const auto& kr : s->krange) {
            int maxn = 0;
            for (int c = kr.first; c < kr.first + kr.second; ++c) maxn = s->ch[c].N > maxn ? s->ch[c].N : maxn;
            launch_w<T>(s->ch[kr.first].K, dim3((maxn + 3) / 4, 1, kr.second), blk, st, cd + kr.first);
        }
    }
    EV(); hipLaunchKernelGGL(k_xz<T>, g_xz, blk, 0, st, cd, s->maxB, s->maxDraws);
    EV(); hipLaunchKernelGGL(k_xz_tail<T>, dim3((s->maxN + TAIL_CELLS - 1) / TAIL_CELLS, 1, C), blk, 0, st, cd);
    EV(); hipLaunchKernelGGL(k_alpha<T>, g_tile, blk, 0, st, cd);
    EV(); hipLaunchKernelGGL(k_alpha_dec<T>, g_adec, blk, 0, st, cd);
    EV(); hipLaunchKernelGGL(k_beta_nu<T>, dim3(g_tile.x, g_tile.y + gy_gamma, C), blk, 0, st, cd, gy_gamma);
    EV(); hipLaunchKernelGGL(k_bn_dec<T>, dim3((s->maxG + BN_GENES - 1) / BN_GENES, 1, C), blk, 0, st, cd);
    EV(); hipLaunchKernelGGL(k_delta<T>, g_rows, blk, 0, st, cd);
    EV(); hipLaunchKernelGGL(k_phi<T>, dim3(g_tile.x, g_tile.y + gy_prep + gy_hist, C), blk, 0, st, cd, gy_prep,
                             gy_hist);
    EV(); hipLaunchKernelGGL(k_record<T>, g_rec, blk, 0, st, cd);
    EV();
#undef EV
    HIPCHK(hipGetLastError());
    return BUSSEQ_OK;
}

// the draws and tables of the current sweep (prep_item) from the parameters on
// the device: at creation and after busseq_set_state; each sweep's k_phi
// prepares the one after it
busseq_status refresh_prep(busseq_sampler* s)
{
    const dim3 g(nblk(s->maxPrep), 1, s->C);
    if (s->prec == 32) hipLaunchKernelGGL(k_prep<float>, g, dim3(256), 0, s->stream, s->d_cd);
    else hipLaunchKernelGGL(k_prep<double>, g, dim3(256), 0, s->stream, s->d_cd);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return BUSSEQ_OK;
}

// Graphs of one sweep and of BQ_GRAPH_SWEEPS sweeps: consecutive graph
// launches are a few microseconds apart on the device, consecutive kernels of
// one graph are not, so long runs replay the longer graph.
#define BQ_GRAPH_SWEEPS 4
void drop_graphs(busseq_sampler* s);

// both graphs, or neither (a failure part way drops what was built)
busseq_status build_graph_once(busseq_sampler* s)
{
    for (int v = 0; v < 2; ++v) {
        HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        busseq_status rc = BUSSEQ_OK;
        for (int r = 0; r < (v ? BQ_GRAPH_SWEEPS : 1) && rc == BUSSEQ_OK; ++r)
            rc = s->prec == 32 ? enqueue_sweep<float>(s) : enqueue_sweep<double>(s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s->stream, &g);
        if (rc != BUSSEQ_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) { bq_set_error(std::string("graph capture: ") + hipGetErrorString(e)); return BUSSEQ_ERR_DEVICE; }
        s->graph[v] = g;
        HIPCHK(hipGraphInstantiate(&s->gexec[v], g, nullptr, nullptr, 0));
    }
    return BUSSEQ_OK;
}

busseq_status build_graph(busseq_sampler* s)
{
    if (s->gexec[0] && s->gexec[1]) return BUSSEQ_OK;
    drop_graphs(s);
    busseq_status rc = build_graph_once(s);
    if (rc != BUSSEQ_OK) drop_graphs(s);
    return rc;
}

// ---------------------------------------------------------------------------
// trace streaming (see TraceHost)
// ---------------------------------------------------------------------------
// The first `rows` rows of block k (ring half k & 1) of every traced chain into
// its pinned copy, on the copy stream after the sweeps enqueued so far.
busseq_status trace_copy(busseq_sampler* s, int k, int rows)
{
    const int half = k & 1;
    HIPCHK(hipEventRecord(s->ev_sweeps, s->stream));
    HIPCHK(hipStreamWaitEvent(s->copy_stream, s->ev_sweeps, 0));
    for (int c = 0; c < s->C; ++c) {
        const TraceHost& t = s->th[c];
        if (!t.P) continue;
        for (int a = 0; a < BQ_TRACE_ARRAYS; ++a) {
            const size_t row = t.np[a] * (size_t)t.esz[a];
            if (!row) continue;
            const size_t off = (size_t)half * BQ_TRACE_BLOCK * row;
            HIPCHK(hipMemcpyAsync((char*)t.stage[a] + off, (const char*)t.dev[a] + off, (size_t)rows * row,
                                  hipMemcpyDeviceToHost, s->copy_stream));
        }
    }
    HIPCHK(hipEventRecord(s->ev_copy[half], s->copy_stream));
    s->copy_recorded[half] = true;
    s->pend_block = k;
    s->pend_rows = rows;
    return BUSSEQ_OK;
}

// The block whose copy is in flight, once it has arrived, into every traced
// chain's host traces (rows of sweeps before burn-in or past n_pres skipped).
busseq_status trace_drain(busseq_sampler* s)
{
    if (s->pend_block < 0) return BUSSEQ_OK;
    const int k = s->pend_block, half = k & 1, rows = s->pend_rows;
    s->pend_block = -1;
    HIPCHK(hipEventSynchronize(s->ev_copy[half]));
    int col[BQ_TRACE_BLOCK];
    for (int c = 0; c < s->C; ++c) {
        TraceHost& t = s->th[c];
        if (!t.P) continue;
        bool any = false;
        for (int q = 0; q < rows; ++q) {
            const int r = k * BQ_TRACE_BLOCK + q - t.nb;
            col[q] = r >= 0 && r < t.P ? r : -1;
            any = any || col[q] >= 0;
        }
        if (!any) continue;
        for (int a = 0; a < BQ_TRACE_ARRAYS; ++a) {
            if (!t.np[a]) continue;
            if (!t.host[a]) {
                t.host[a] = std::calloc(t.np[a] * (size_t)t.P, (size_t)t.esz[a]);
                if (!t.host[a]) { bq_set_error("host allocation of traces failed"); return BUSSEQ_ERR_NOMEM; }
            }
            const size_t row = t.np[a] * (size_t)t.esz[a];
            bq_trace_transpose((const char*)t.stage[a] + (size_t)half * BQ_TRACE_BLOCK * row, t.esz[a], t.np[a],
                               rows, col, (size_t)t.P, t.host[a]);
        }
    }
    return BUSSEQ_OK;
}

// whether some chain records a sweep of block k
bool trace_block_used(const busseq_sampler* s, int k)
{
    const int t0 = k * BQ_TRACE_BLOCK, t1 = t0 + BQ_TRACE_BLOCK;
    for (const TraceHost& t : s->th)
        if (t.P && t1 > t.nb && t0 < t.nb + t.P) return true;
    return false;
}

// After the sweeps of block k are enqueued: the previous block (its copy ran
// behind its sweeps) into the host traces, this block's copy enqueued, and the
// next block's sweeps held until the copy out of the ring half they overwrite
// has finished.
busseq_status trace_boundary(busseq_sampler* s, int k)
{
    if (trace_block_used(s, k)) {
        busseq_status rc = trace_drain(s);
        if (rc != BUSSEQ_OK) return rc;
        rc = trace_copy(s, k, BQ_TRACE_BLOCK);
        if (rc != BUSSEQ_OK) return rc;
    }
    if (s->copy_recorded[(k + 1) & 1]) HIPCHK(hipStreamWaitEvent(s->stream, s->ev_copy[(k + 1) & 1], 0));
    return BUSSEQ_OK;
}

// every recorded sweep so far into the host traces (the sweeps since the last
// block boundary too)
busseq_status trace_flush(busseq_sampler* s)
{
    if (!s->tracing) return BUSSEQ_OK;
    HIPCHK(hipStreamSynchronize(s->stream));
    busseq_status rc = trace_drain(s);
    if (rc != BUSSEQ_OK) return rc;
    const int k = s->iter / BQ_TRACE_BLOCK, rows = s->iter % BQ_TRACE_BLOCK;
    if (rows > 0 && trace_block_used(s, k)) {
        rc = trace_copy(s, k, rows);
        if (rc == BUSSEQ_OK) rc = trace_drain(s);
    }
    return rc;
}

// one sweep enqueued outside the graphs (busseq_kernel_times)
busseq_status count_sweep(busseq_sampler* s)
{
    s->iter += 1;
    if (s->tracing && s->iter % BQ_TRACE_BLOCK == 0) return trace_boundary(s, s->iter / BQ_TRACE_BLOCK - 1);
    return BUSSEQ_OK;
}

// n sweeps of every chain, enqueued on the sampler's stream; with traces, in
// runs that end at the streamed blocks' boundaries
busseq_status launch_sweeps(busseq_sampler* s, int n)
{
    while (n > 0) {
        int m = n;
        if (s->tracing) {
            const int left = BQ_TRACE_BLOCK - s->iter % BQ_TRACE_BLOCK;
            if (left < m) m = left;
        }
        n -= m;
        s->iter += m;
        for (; m >= BQ_GRAPH_SWEEPS; m -= BQ_GRAPH_SWEEPS) HIPCHK(hipGraphLaunch(s->gexec[1], s->stream));
        for (; m > 0; --m) HIPCHK(hipGraphLaunch(s->gexec[0], s->stream));
        if (s->tracing && s->iter % BQ_TRACE_BLOCK == 0) {
            busseq_status rc = trace_boundary(s, s->iter / BQ_TRACE_BLOCK - 1);
            if (rc != BUSSEQ_OK) return rc;
        }
    }
    return BUSSEQ_OK;
}

void drop_graphs(busseq_sampler* s)
{
    for (int v = 0; v < 2; ++v) {
        if (s->gexec[v]) { (void)hipGraphExecDestroy(s->gexec[v]); s->gexec[v] = nullptr; }
        if (s->graph[v]) { (void)hipGraphDestroy(s->graph[v]); s->graph[v] = nullptr; }
    }
}

template <typename T>
void to_t(std::vector<char>& out, const std::vector<double>& v, double (*f)(double))
{
    out.resize(v.size() * sizeof(T));
    T* p = (T*)out.data();
    for (size_t j = 0; j < v.size(); ++j) p[j] = (T)(f ? f(v[j]) : v[j]);
}
double f_exp(double x) { return bq_exp(x); }

// (beta, expm1(beta)) pairs in precision T, the layout k_w reads
template <typename T>
void to_pairs(std::vector<char>& out, const std::vector<double>& beta, int K, int G)
{
    const int st = bt_stride(K);
    out.assign(2 * (size_t)G * st * sizeof(T), 0);
    T* p = (T*)out.data();
    for (int g = 0; g < G; ++g)
        for (int k = 1; k < K; ++k) {
            const double v = beta[(size_t)k * G + g];
            const size_t q = (size_t)g * st + (k - 1);
            p[2 * q] = (T)v;
            p[2 * q + 1] = tw_em(v, T());
        }
}

// precision-T copies of a chain's parameters, computed on the host with the
// same functions the kernels use
busseq_status upload_tcopies(busseq_sampler* s, int ch, const std::vector<double>& alpha,
                             const std::vector<double>& beta, const std::vector<double>& nu,
                             const std::vector<double>& delta, const std::vector<double>& phi)
{
    const ChainDev& c = s->hd[ch];
    std::vector<char> buf;
    auto up = [&](void* dst) -> hipError_t { return hipMemcpy(dst, buf.data(), buf.size(), hipMemcpyHostToDevice); };
    if (s->prec == 32) {
        to_t<float>(buf, alpha, f_exp); HIPCHK(up(c.ea));
        to_t<float>(buf, beta, f_exp); HIPCHK(up(c.eb));
        to_pairs<float>(buf, beta, c.K, c.G); HIPCHK(up(c.bt));
        to_t<float>(buf, nu, f_exp); HIPCHK(up(c.en));
        to_t<float>(buf, delta, f_exp); HIPCHK(up(c.ed));
        to_t<float>(buf, phi, nullptr); HIPCHK(up(c.ph));
    } else {
        to_t<double>(buf, alpha, f_exp); HIPCHK(up(c.ea));
        to_t<double>(buf, beta, f_exp); HIPCHK(up(c.eb));
        to_pairs<double>(buf, beta, c.K, c.G); HIPCHK(up(c.bt));
        to_t<double>(buf, nu, f_exp); HIPCHK(up(c.en));
        to_t<double>(buf, delta, f_exp); HIPCHK(up(c.ed));
        to_t<double>(buf, phi, nullptr); HIPCHK(up(c.ph));
    }
    return BUSSEQ_OK;
}

template <typename V>
hipError_t h2d(V* dst, const std::vector<V>& v)
{
    return v.empty() ? hipSuccess : hipMemcpy(dst, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice);
}

template <typename V>
hipError_t d2h(std::vector<V>& v, const V* src, size_t n)
{
    v.resize(n);
    return hipMemcpy(v.data(), src, n * sizeof(V), hipMemcpyDeviceToHost);
}

busseq_status validate(const busseq_config* cfg)
{
    if (!cfg) { bq_set_error("null config"); return BUSSEQ_ERR_ARG; }
    // equality is rejected too: the run would record nothing, and the reference
    // (which accepts it, cu:1121) would write posterior means of zero iterations
    if (cfg->n_burnin >= cfg->n_iter) {
        bq_set_error("Burn-in iterations must be less than total number of iterations.");
        return BUSSEQ_ERR_ARG;
    }
    if (cfg->B < 2) { bq_set_error("The batch number must be greater than one."); return BUSSEQ_ERR_ARG; }
    if (cfg->B > BQ_BMAX) { bq_set_error("at most 128 batches are supported"); return BUSSEQ_ERR_ARG; }
    if (!cfg->n_b) { bq_set_error("batch sizes missing"); return BUSSEQ_ERR_ARG; }
    for (int b = 0; b < cfg->B; ++b)
        if (cfg->n_b[b] < cfg->K) {
            bq_set_error("The sample size in any batch must be greater than the assumed cell type number.");
            return BUSSEQ_ERR_ARG;
        }
    if (cfg->K < 2 || cfg->K > BQ_KMAX) { bq_set_error("K must be between 2 and 64"); return BUSSEQ_ERR_ARG; }
    if ((long long)cfg->B * cfg->K > BQ_BKMAX) { bq_set_error("B * K must be at most 2048"); return BUSSEQ_ERR_ARG; }
    if (cfg->G < 1 || cfg->G > 65535) { bq_set_error("G must be between 1 and 65535"); return BUSSEQ_ERR_ARG; }
    if (cfg->n_iter < 0 || cfg->n_burnin < 0) { bq_set_error("negative iteration count"); return BUSSEQ_ERR_ARG; }
    if (cfg->precision != 32 && cfg->precision != 64) { bq_set_error("precision must be 32 or 64"); return BUSSEQ_ERR_ARG; }
    if (cfg->chains < 1) { bq_set_error("chains must be >= 1"); return BUSSEQ_ERR_ARG; }
    if (cfg->tile_rows < 0 || cfg->tile_rows > BQ_TR_MAX) { bq_set_error("tile_rows must be between 1 and 256 (0: default)"); return BUSSEQ_ERR_ARG; }
    long long N = 0;
    for (int b = 0; b < cfg->B; ++b) N += cfg->n_b[b];
    if (N * (long long)((cfg->G + 63) & ~63) > 2147483647LL) { bq_set_error("N x G must stay below 2^31"); return BUSSEQ_ERR_ARG; }
    return BUSSEQ_OK;
}

// static data of one problem, shared by its chains
struct ProblemDev {
    int32_t *batch, *boff, *zrow, *big_off, *big;
    int64_t* hist;
    uint64_t* hlist;
    long long n_hlist;
    uint16_t* zcol;
};

// device buffers that live only while a sampler is created
struct DevTmp {
    std::vector<void*> p;
    bool oom = false;
    ~DevTmp() { for (void* q : p) (void)hipFree(q); }
    template <typename V>
    V* alloc(size_t n)
    {
        void* q = nullptr;
        if (hipMalloc(&q, (n ? n : 1) * sizeof(V)) != hipSuccess) { oom = true; return nullptr; }
        p.push_back(q);
        return (V*)q;
    }
};

#define INITCHK(x) HIPCHK((hipError_t)(x))
static const int kRawRows = 64;   // cells per chunk of the raw-mean pass

// Static data of one problem and its initial state matrix S0 (BUSseq_gpu.cu:1188-1189,
// 1204-1205, 1226-1228), built on the device from one upload of the counts.
// Returns the library sizes, the number of zero entries and the chunks of the
// raw-mean pass.
busseq_status init_problem(busseq_sampler* s, const int32_t* Y, int B, int N, int G, int Gp,
                           const std::vector<int32_t>& batch, const std::vector<int32_t>& boff, ProblemDev& pd,
                           int32_t*& S0, std::vector<int64_t>& lib, long& nz, std::vector<int32_t>& chunks)
{
    DevTmp tmp;
    int32_t* Yd = tmp.alloc<int32_t>((size_t)N * G);
    int32_t* nzr = tmp.alloc<int32_t>(N);
    int32_t* nbr = tmp.alloc<int32_t>(N);
    int64_t* libd = tmp.alloc<int64_t>(N);
    int32_t* browd = tmp.alloc<int32_t>((size_t)N + 1);
    pd.batch = s->dalloc<int32_t>(N); pd.boff = s->dalloc<int32_t>(B + 1);
    pd.hist = s->dalloc<int64_t>((size_t)B * H); pd.zrow = s->dalloc<int32_t>((size_t)N + 1);
    pd.big_off = s->dalloc<int32_t>((size_t)B + 1);
    uint32_t* hist_bg = tmp.alloc<uint32_t>((size_t)B * G * BQ_PHI_HV);
    uint32_t* hcnt = tmp.alloc<uint32_t>((size_t)B * G);
    S0 = s->dalloc<int32_t>((size_t)N * Gp);
    if (tmp.oom || s->oom) { bq_set_error("device allocation failed"); return BUSSEQ_ERR_NOMEM; }
    HIPCHK(hipMemset(hist_bg, 0, (size_t)B * G * BQ_PHI_HV * sizeof(uint32_t)));
    HIPCHK(hipDeviceSynchronize());   // the clears of dalloc precede the passes on the sampler's stream
    HIPCHK(hipMemcpy(Yd, Y, (size_t)N * G * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(h2d(pd.batch, batch));
    HIPCHK(h2d(pd.boff, boff));
    INITCHK(bq_dev_init_rows(Yd, N, G, Gp, pd.batch, S0, nzr, nbr, libd, pd.hist, s->stream));
    INITCHK(bq_dev_init_hist_bg(S0, N, G, Gp, pd.batch, hist_bg, s->stream));
    INITCHK(bq_dev_hist_distinct(hist_bg, (long long)B * G, hcnt, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    {
        // the histogram's nonzero entries as a list, each item's together
        std::vector<uint32_t> cnt;
        HIPCHK(d2h(cnt, hcnt, (size_t)B * G));
        std::vector<uint64_t> off((size_t)B * G);
        uint64_t tot = 0;
        for (size_t j = 0; j < off.size(); ++j) { off[j] = tot; tot += cnt[j]; }
        pd.n_hlist = (long long)tot;
        pd.hlist = s->dalloc<uint64_t>(tot);
        uint64_t* offd = tmp.alloc<uint64_t>(off.size());
        if (tmp.oom || s->oom) { bq_set_error("device allocation failed"); return BUSSEQ_ERR_NOMEM; }
        HIPCHK(h2d(offd, off));
        INITCHK(bq_dev_hist_list(hist_bg, (long long)B * G, offd, pd.hlist, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    std::vector<int32_t> nzh, nbh;
    HIPCHK(d2h(nzh, nzr, N));
    HIPCHK(d2h(nbh, nbr, N));
    HIPCHK(d2h(lib, libd, N));
    for (int i = 0; i < N; ++i)
        if (lib[i] <= 0) {
            bq_set_error("cell " + std::to_string(i) + " has no counts (zero library size)");
            return BUSSEQ_ERR_DATA;
        }
    std::vector<int32_t> zrow((size_t)N + 1, 0), brow((size_t)N + 1, 0), big_off((size_t)B + 1, 0);
    for (int i = 0; i < N; ++i) {
        zrow[(size_t)i + 1] = zrow[i] + nzh[i];
        brow[(size_t)i + 1] = brow[i] + nbh[i];
    }
    for (int b = 0; b < B; ++b) big_off[(size_t)b + 1] = brow[boff[b + 1]];
    nz = zrow[N];
    pd.zcol = s->dalloc<uint16_t>((size_t)nz);
    pd.big = s->dalloc<int32_t>((size_t)brow[N]);
    if (s->oom) { bq_set_error("device allocation failed"); return BUSSEQ_ERR_NOMEM; }
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(h2d(pd.zrow, zrow));
    HIPCHK(h2d(pd.big_off, big_off));
    HIPCHK(hipMemcpy(browd, brow.data(), brow.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    INITCHK(bq_dev_init_lists(S0, N, G, Gp, pd.zrow, browd, pd.zcol, pd.big, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    chunks.clear();
    for (int b = 0; b < B; ++b)
        for (int st = boff[b]; st < boff[b + 1]; st += kRawRows) {
            chunks.push_back(b);
            chunks.push_back(st);
            chunks.push_back(boff[b + 1] - st < kRawRows ? boff[b + 1] - st : kRawRows);
        }
    return BUSSEQ_OK;
}

// Raw-mean sums of one chain (cu:369-382, 1210-1221) from its W0 (already in
// Wd) and delta0: rs[(b*K + k)*G + g], exact fixed point.
busseq_status init_raw_sums(busseq_sampler* s, const int32_t* S0, int B, int N, int G, int Gp, int K,
                            const int32_t* Wd, const std::vector<double>& delta, const std::vector<int32_t>& chunks,
                            std::vector<uint64_t>& rs)
{
    DevTmp tmp;
    const size_t nrs = (size_t)B * K * G;
    uint64_t* rsd = tmp.alloc<uint64_t>(nrs);
    double* ied = tmp.alloc<double>(N);
    int32_t* chd = tmp.alloc<int32_t>(chunks.size());
    if (tmp.oom) { bq_set_error("device allocation failed"); return BUSSEQ_ERR_NOMEM; }
    std::vector<double> ie(N);
    for (int i = 0; i < N; ++i) ie[i] = bq_exp(-delta[i]);
    HIPCHK(h2d(ied, ie));
    HIPCHK(h2d(chd, chunks));
    HIPCHK(hipMemsetAsync(rsd, 0, nrs * sizeof(uint64_t), s->stream));
    INITCHK(bq_dev_init_raw(S0, G, Gp, K, Wd, ied, chd, (int)(chunks.size() / 3), rsd, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(d2h(rs, rsd, nrs));
    return BUSSEQ_OK;
}

} // namespace

// ---------------------------------------------------------------------------
// bq_rcpf's precondition: on this device it returns what the host returns, the
// IEEE quotient 1.0f / z of its argument held to [2^-126, 2^125], for every
// normal positive z below 2^126 (the reciprocal and its Newton step for each
// held value, and the hold itself).  One thread per float: exponent field
// 1 + blockIdx.y (1..252), mantissa blockIdx.x * 256 + threadIdx.x (< 2^23).
// ---------------------------------------------------------------------------
#define BQ_RCP_GRID_Y 252
__global__ __launch_bounds__(256) void k_rcp_check(unsigned long long* bad)
{
    const uint32_t e = 1u + blockIdx.y;
    const float z = bq_bitsf((e << 23) | (blockIdx.x * 256u + threadIdx.x));
    const float held = __builtin_fminf(__builtin_fmaxf(z, BQ_RCP_LO), BQ_RCP_HI);
    const bool differ = bq_fbits(bq_rcpf(z)) != bq_fbits(1.0f / held);
    const uint64_t m = __ballot(differ);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(bad, (unsigned long long)__popcll(m));
}

// number of the 252 x 2^23 floats for which bq_rcpf(z) differs from the host's
// value on the current device
static busseq_status rcp_count(unsigned long long* bad)
{
    unsigned long long* d = nullptr;
    if (hipMalloc((void**)&d, sizeof *d) != hipSuccess) {
        bq_set_error("hipMalloc failed in the reciprocal check");
        return BUSSEQ_ERR_NOMEM;
    }
    hipError_t e = hipMemset(d, 0, sizeof *d);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_rcp_check, dim3(1u << 15, BQ_RCP_GRID_Y), dim3(256), 0, 0, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(bad, d, sizeof *bad, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) {
        bq_set_error(std::string("HIP error in the reciprocal check: ") + hipGetErrorString(e));
        return BUSSEQ_ERR_DEVICE;
    }
    return BUSSEQ_OK;
}

// once per device and process, before its first sampler: the single-precision
// terms rely on bq_rcpf returning the host's quotient
#define BQ_RCP_DEVICES 64
static std::mutex g_rcp_mu;
static int g_rcp_state[BQ_RCP_DEVICES];   // 0 not checked, 1 equal everywhere, -1 differs

static busseq_status rcp_ready(int dev)
{
    if (dev < 0 || dev >= BQ_RCP_DEVICES) { bq_set_error("device index out of range"); return BUSSEQ_ERR_ARG; }
    std::lock_guard<std::mutex> lock(g_rcp_mu);
    if (g_rcp_state[dev] == 0) {
        unsigned long long bad = 0;
        busseq_status rc = rcp_count(&bad);
        if (rc != BUSSEQ_OK) return rc;
        g_rcp_state[dev] = bad == 0 ? 1 : -1;
    }
    if (g_rcp_state[dev] < 0) {
        bq_set_error("this device's v_rcp_f32 with one Newton step differs from the IEEE quotient 1/z for "
                     "some normal z in [2^-126, 2^125]; single-precision sweeps would not match the host's numerics");
        return BUSSEQ_ERR_DEVICE;
    }
    return BUSSEQ_OK;
}

extern "C" busseq_status busseq_rcp_check(int device, uint64_t* mismatches)
{
    if (!mismatches) { bq_set_error("null output of busseq_rcp_check"); return BUSSEQ_ERR_ARG; }
    if (hipSetDevice(device) != hipSuccess) { bq_set_error("hipSetDevice failed (no GPU?)"); return BUSSEQ_ERR_DEVICE; }
    unsigned long long bad = 0;
    busseq_status rc = rcp_count(&bad);
    if (rc == BUSSEQ_OK) *mismatches = bad;
    return rc;
}

// Cells per tile when the configuration leaves it at 0: 16 below 2560 cells in
// all chains together, 32 below 65536, 64 from there on.  Measured on one
// MI355X over c1, c2, c3 and six of c5's data sets (16..256 each), this is
// within 1.5% of the fastest size on each, where 64 lost up to 12% on the
// small ones; the autotuner measures instead.
extern "C" int busseq_default_tile_rows(long long cells)
{
    return cells < 2560 ? 16 : cells < 65536 ? 32 : 64;
}

extern "C" busseq_status busseq_create_batch(const busseq_problem* probs, int n_probs, busseq_sampler** out)
{
    if (!out || !probs || n_probs < 1) { bq_set_error("bad arguments to busseq_create_batch"); return BUSSEQ_ERR_ARG; }
    *out = nullptr;
    for (int p = 0; p < n_probs; ++p) {
        busseq_status rc = validate(probs[p].cfg);
        if (rc != BUSSEQ_OK) return rc;
        if (!probs[p].Y) { bq_set_error("null count matrix"); return BUSSEQ_ERR_ARG; }
        if (probs[p].cfg->precision != probs[0].cfg->precision || probs[p].cfg->tile_rows != probs[0].cfg->tile_rows ||
            probs[p].cfg->device != probs[0].cfg->device) {
            bq_set_error("all problems of a batch must share precision, tile_rows and device");
            return BUSSEQ_ERR_ARG;
        }
    }
    auto* s = new busseq_sampler();
    auto fail = [&](busseq_status r) { busseq_destroy(s); return r; };
    const busseq_config* c0 = probs[0].cfg;
    s->prec = c0->precision;
    long long cells = 0;   // every chain's cells: what the tile passes' grids cover
    for (int p = 0; p < n_probs; ++p) {
        long long n = 0;
        for (int b = 0; b < probs[p].cfg->B; ++b) n += probs[p].cfg->n_b[b];
        cells += n * probs[p].cfg->chains;
    }
    s->TR = c0->tile_rows > 0 ? c0->tile_rows : busseq_default_tile_rows(cells);
    s->dev = c0->device;
    if (hipSetDevice(s->dev) != hipSuccess) { bq_set_error("hipSetDevice failed (no GPU?)"); return fail(BUSSEQ_ERR_DEVICE); }
    if (s->prec == 32) {
        busseq_status rc = rcp_ready(s->dev);
        if (rc != BUSSEQ_OK) return fail(rc);
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        bq_set_error("hipStreamCreate failed");
        return fail(BUSSEQ_ERR_DEVICE);
    }
    const size_t ts = s->prec == 32 ? 4 : 8;
    const int TR = s->TR;
    for (int p = 0; p < n_probs; ++p) {
        const busseq_config* cfg = probs[p].cfg;
        const int32_t* Y = probs[p].Y;
        const int B = cfg->B, G = cfg->G, K = cfg->K, Gp = (G + 63) & ~63;
        std::vector<int> n_b(cfg->n_b, cfg->n_b + B);
        int N = 0;
        for (int b = 0; b < B; ++b) N += n_b[b];
        const int T_max = (N + TR - 1) / TR + B * K;
        const int P = cfg->n_iter - cfg->n_burnin;
        std::vector<int32_t> batch(N), boff(B + 1, 0);
        for (int b = 0; b < B; ++b) boff[b + 1] = boff[b] + n_b[b];
        for (int b = 0; b < B; ++b) for (int i = boff[b]; i < boff[b + 1]; ++i) batch[i] = b;
        ProblemDev pd{};
        int32_t* S0 = nullptr;   // initial state matrix: chain 0's S, copied to the others
        std::vector<int64_t> lib;
        std::vector<int32_t> chunks;
        long nz = 0;             // entries with Y = 0
        {
            busseq_status rc = init_problem(s, Y, B, N, G, Gp, batch, boff, pd, S0, lib, nz, chunks);
            if (rc != BUSSEQ_OK) return fail(rc);
        }
        for (int k = 0; k < cfg->chains; ++k) {
            HostInit hi;
            uint32_t key = (uint32_t)(cfg->chain0 + k);
            bq_init_cells(B, n_b.data(), K, lib.data(), cfg->seed, key, hi);
            ChainDev c{};
            c.N = N; c.G = G; c.Gp = Gp; c.K = K; c.B = B;
            c.seed = cfg->seed; c.key = key;
            c.n_burnin = cfg->n_burnin; c.n_unchanged = cfg->n_unchanged;
            c.n_pres = P; c.traces = cfg->keep_traces && P > 0 ? 1 : 0;
            c.tr_rows = c.traces ? 2 * BQ_TRACE_BLOCK : 1;
            c.TR = TR; c.T_max = T_max; c.prec = s->prec;
            c.batch = pd.batch; c.boff = pd.boff; c.hist_s = pd.hist; c.hlist = pd.hlist; c.n_hlist = pd.n_hlist;
            c.zrow = pd.zrow; c.zcol = pd.zcol; c.big_off = pd.big_off; c.big = pd.big;
            c.iter = s->dalloc<int>(1);
            c.S = k == 0 ? S0 : s->dalloc<int32_t>((size_t)N * Gp);
            c.W = s->dalloc<int32_t>(N);
            c.L = s->dalloc<int32_t>((size_t)K * G);
            c.alpha = s->dalloc<double>(G); c.beta = s->dalloc<double>((size_t)K * G);
            c.nu = s->dalloc<double>((size_t)B * G); c.delta = s->dalloc<double>(N);
            c.phi = s->dalloc<double>((size_t)B * G); c.gamma = s->dalloc<double>(2 * (size_t)B);
            c.pi = s->dalloc<double>((size_t)B * K); c.ptau = s->dalloc<double>(2);
            c.mu_alpha = s->dalloc<double>(G); c.mu_nu = s->dalloc<double>(B); c.mu_delta = s->dalloc<double>(N);
            c.ea = s->dalloc<char>(ts * G); c.eb = s->dalloc<char>(ts * K * G);
            c.bt = s->dalloc<char>(2 * ts * (size_t)bt_stride(K) * G); c.en = s->dalloc<char>(ts * B * G); c.ed = s->dalloc<char>(ts * N);
            c.ph = s->dalloc<char>(ts * B * G);
            c.alpha_p = s->dalloc<double>(G); c.beta_p = s->dalloc<double>((size_t)K * G);
            c.nu_p = s->dalloc<double>((size_t)B * G); c.phi_p = s->dalloc<double>((size_t)B * G);
            c.gamma_p = s->dalloc<double>(2 * (size_t)B);
            c.delta_p = s->dalloc<double>(N); c.delta_lu = s->dalloc<double>(N);
            c.dd = s->dalloc<char>(ts * N); c.cd = s->dalloc<char>(ts * N);
            c.da = s->dalloc<char>(ts * G); c.ca = s->dalloc<char>(ts * G);
            c.db = s->dalloc<char>(ts * K * G); c.cb = s->dalloc<char>(ts * K * G);
            c.dn = s->dalloc<char>(ts * B * G); c.cn = s->dalloc<char>(ts * B * G);
            c.pc = s->dalloc<char>(ts * 6 * B * G);
            c.tm = s->dalloc<char>(ts * K * B * G);
            c.cnt = s->dalloc<int32_t>((size_t)B * K * BQ_STRIPES);
            c.perm = s->dalloc<int32_t>(N);
            c.tiles = s->dalloc<int32_t>((size_t)T_max * 4);
            c.tile_off = s->dalloc<int32_t>((size_t)B * K + 1);
            c.ntiles = s->dalloc<int32_t>(1);
            c.gsum = s->dalloc<uint64_t>(gs_size(G, K, B));
            c.hist_d = s->dalloc<uint32_t>((size_t)B * H);
            c.hist_lo = s->dalloc<uint32_t>((size_t)B * BQ_HLO * BQ_STRIPES);
            c.gstat = s->dalloc<uint64_t>((size_t)B * 5 * BQ_STRIPES);
            c.lstat = s->dalloc<uint64_t>(2);
            c.xz_tail = s->dalloc<int32_t>(nz > 0 ? (size_t)nz : 1);
            c.zs = s->dalloc<int32_t>(nz > 0 ? (size_t)nz : 1);   // every Y = 0 entry starts at X = 0, Z = 0
            c.xz_ntail = s->dalloc<int32_t>(N);
            c.odds_tab = s->dalloc<double>((size_t)B * H);
            c.beta_acc = s->dalloc<int32_t>((size_t)K * G);
            c.n_rec = s->dalloc<int>(1);
            c.iter_rec = s->dalloc<int>(1);
            c.s_gamma = s->dalloc<double>(2 * (size_t)B); c.s_alpha = s->dalloc<double>(G);
            c.s_beta = s->dalloc<double>((size_t)K * G); c.s_nu = s->dalloc<double>((size_t)B * G);
            c.s_delta = s->dalloc<double>(N); c.s_phi = s->dalloc<double>((size_t)B * G);
            c.s_pi = s->dalloc<double>((size_t)B * K); c.s_ptau = s->dalloc<double>(2);
            c.c_L = s->dalloc<int32_t>((size_t)K * G); c.c_W = s->dalloc<int32_t>((size_t)N * K);
            TraceHost th;
            if (c.traces) {
                // rings of 2 streamed blocks of rows; their pinned copies
                const size_t R2 = (size_t)c.tr_rows;
                const size_t np[BQ_TRACE_ARRAYS] = {2 * (size_t)B, (size_t)G, (size_t)(K - 1) * G, (size_t)(K - 1) * G, 2,
                                                    (size_t)(B - 1) * G, (size_t)N, (size_t)B * G, (size_t)N, (size_t)B * K};
                const int esz[BQ_TRACE_ARRAYS] = {8, 8, 4, 8, 8, 8, 8, 8, 4, 8};
                void** dev[BQ_TRACE_ARRAYS] = {(void**)&c.tr_gamma, (void**)&c.tr_alpha, (void**)&c.tr_L, (void**)&c.tr_beta,
                                               (void**)&c.tr_ptau, (void**)&c.tr_nu, (void**)&c.tr_delta, (void**)&c.tr_phi,
                                               (void**)&c.tr_W, (void**)&c.tr_pi};
                th.P = P;
                th.nb = cfg->n_burnin;
                for (int a = 0; a < BQ_TRACE_ARRAYS; ++a) {
                    th.np[a] = np[a];
                    th.esz[a] = esz[a];
                    *dev[a] = s->dalloc<char>(R2 * np[a] * esz[a]);
                    th.dev[a] = *dev[a];
                    if (hipHostMalloc(&th.stage[a], R2 * np[a] * esz[a] + 1, hipHostMallocDefault) != hipSuccess) {
                        th.stage[a] = nullptr;
                        s->th.push_back(th);
                        bq_set_error("pinned host allocation of the trace blocks failed");
                        return fail(BUSSEQ_ERR_NOMEM);
                    }
                }
                s->tracing = true;
            }
            s->th.push_back(th);
            if (s->oom) {
                bq_set_error("device allocation failed");
                return fail(BUSSEQ_ERR_NOMEM);
            }
            if (hipDeviceSynchronize() != hipSuccess ||
                (k > 0 && hipMemcpy(c.S, S0, (size_t)N * Gp * sizeof(int32_t), hipMemcpyDeviceToDevice) != hipSuccess) ||
                h2d(c.W, hi.W) != hipSuccess) {
                bq_set_error("upload of the initial state failed");
                return fail(BUSSEQ_ERR_DEVICE);
            }
            {
                std::vector<uint64_t> rs;
                busseq_status rc = init_raw_sums(s, S0, B, N, G, Gp, K, c.W, hi.delta, chunks, rs);
                if (rc != BUSSEQ_OK) return fail(rc);
                bq_init_params(B, n_b.data(), G, K, rs.data(), cfg->seed, key, hi);
            }
            if (h2d(c.L, hi.L) != hipSuccess ||
                h2d(c.alpha, hi.alpha) != hipSuccess || h2d(c.beta, hi.beta) != hipSuccess ||
                h2d(c.nu, hi.nu) != hipSuccess || h2d(c.delta, hi.delta) != hipSuccess ||
                h2d(c.phi, hi.phi) != hipSuccess || h2d(c.gamma, hi.gamma) != hipSuccess ||
                h2d(c.pi, hi.pi) != hipSuccess || h2d(c.ptau, hi.ptau) != hipSuccess ||
                h2d(c.mu_alpha, hi.mu_alpha) != hipSuccess || h2d(c.mu_nu, hi.mu_nu) != hipSuccess ||
                h2d(c.mu_delta, hi.mu_delta) != hipSuccess) {
                bq_set_error("upload of the initial state failed");
                return fail(BUSSEQ_ERR_DEVICE);
            }
            s->hd.push_back(c);
            s->ch.push_back(ChainHost{p, n_b, N, G, Gp, K, B, T_max});
            busseq_status rc = upload_tcopies(s, (int)s->hd.size() - 1, hi.alpha, hi.beta, hi.nu, hi.delta, hi.phi);
            if (rc != BUSSEQ_OK) return fail(rc);
            auto mx = [](int& a, long v) { if (v > a) a = (int)v; };
            mx(s->maxN, N); mx(s->maxG, G); mx(s->maxK, K); mx(s->maxB, B); mx(s->maxT, T_max);
            mx(s->maxPrep, G); mx(s->maxPrep, N); mx(s->maxPrep, (long)K * B * G); mx(s->maxPrep, (long)B * H); mx(s->maxPrep, (long)B * BQ_HLO * BQ_STRIPES); mx(s->maxPrep, (long)B * K * BQ_STRIPES);
            mx(s->maxDraws, nblk((long)G * ((K - 1) + (B - 1) + B)));
            mx(s->maxRec, N); mx(s->maxRec, (long)K * G); mx(s->maxRec, (long)B * G); mx(s->maxRec, (long)B * K);
            mx(s->maxRec, 2L * B); mx(s->maxRec, (long)B * K * BQ_STRIPES);
        }
    }
    s->C = (int)s->hd.size();
    for (int c = 0; c < s->C; ++c) {
        if (!s->krange.empty() && s->ch[s->krange.back().first].K == s->ch[c].K) s->krange.back().second += 1;
        else s->krange.push_back({c, 1});
    }
    if (s->tracing) {
        if (hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_sweeps, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_copy[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_copy[1], hipEventDisableTiming) != hipSuccess) {
            bq_set_error("creating the trace copy stream failed");
            return fail(BUSSEQ_ERR_DEVICE);
        }
    }
    if (hipMalloc(&s->d_cd, sizeof(ChainDev) * s->C) != hipSuccess) { bq_set_error("device allocation failed"); return fail(BUSSEQ_ERR_NOMEM); }
    if (hipMemcpy(s->d_cd, s->hd.data(), sizeof(ChainDev) * s->C, hipMemcpyHostToDevice) != hipSuccess) {
        bq_set_error("upload failed");
        return fail(BUSSEQ_ERR_DEVICE);
    }
    if (hipDeviceSynchronize() != hipSuccess) { bq_set_error("device sync failed"); return fail(BUSSEQ_ERR_DEVICE); }
    if (refresh_prep(s) != BUSSEQ_OK) return fail(BUSSEQ_ERR_DEVICE);
    *out = s;
    return BUSSEQ_OK;
}

extern "C" busseq_status busseq_create(const busseq_config* cfg, const int32_t* Y, busseq_sampler** out)
{
    busseq_problem p;
    p.cfg = cfg;
    p.Y = Y;
    return busseq_create_batch(&p, 1, out);
}

extern "C" void busseq_destroy(busseq_sampler* s)
{
    if (!s) return;
    (void)hipSetDevice(s->dev);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->copy_stream) (void)hipStreamSynchronize(s->copy_stream);
    drop_graphs(s);
    for (TraceHost& t : s->th)
        for (int a = 0; a < BQ_TRACE_ARRAYS; ++a) {
            if (t.stage[a]) (void)hipHostFree(t.stage[a]);
            std::free(t.host[a]);
        }
    if (s->ev_sweeps) (void)hipEventDestroy(s->ev_sweeps);
    for (hipEvent_t e : s->ev_copy) if (e) (void)hipEventDestroy(e);
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    for (void* p : s->allocs) (void)hipFree(p);
    if (s->d_cd) (void)hipFree(s->d_cd);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

extern "C" int busseq_iteration(const busseq_sampler* s) { return s ? s->iter : -1; }
extern "C" int busseq_chain_count(const busseq_sampler* s) { return s ? s->C : 0; }

extern "C" busseq_status busseq_sweeps(busseq_sampler* s, int n)
{
    if (!s || n < 0) { bq_set_error("bad arguments to busseq_sweeps"); return BUSSEQ_ERR_ARG; }
    HIPCHK(hipSetDevice(s->dev));
    busseq_status rc = build_graph(s);
    if (rc != BUSSEQ_OK) return rc;
    rc = launch_sweeps(s, n);
    if (rc != BUSSEQ_OK) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    return BUSSEQ_OK;
}

extern "C" busseq_status busseq_time_sweeps(busseq_sampler* s, int n, double* ms)
{
    if (!s || n < 0 || !ms) { bq_set_error("bad arguments to busseq_time_sweeps"); return BUSSEQ_ERR_ARG; }
    HIPCHK(hipSetDevice(s->dev));
    busseq_status rc = build_graph(s);
    if (rc != BUSSEQ_OK) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipEventRecord(e0, s->stream));
    rc = launch_sweeps(s, n);
    if (rc != BUSSEQ_OK) return rc;
    HIPCHK(hipEventRecord(e1, s->stream));
    HIPCHK(hipEventSynchronize(e1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = f;
    return BUSSEQ_OK;
}

// Tile rows of every chain of a sampler: the tile list k_plan builds, the
// tile capacity and the tile-pass grid follow; the sweep graph is rebuilt on
// the next launch.  Only sizes whose tiles fit the partial-sum buffer.
static busseq_status set_tile_rows(busseq_sampler* s, int tr)
{
    if (tr < 1 || tr > BQ_TR_MAX) { bq_set_error("tile_rows must be between 1 and 256"); return BUSSEQ_ERR_ARG; }
    int maxT = 0;
    for (int c = 0; c < s->C; ++c) {
        const ChainHost& h = s->ch[c];
        const int T = (h.N + tr - 1) / tr + h.B * h.K;
        if (T > h.T_alloc) { bq_set_error("tile_rows below the size the sampler was created for"); return BUSSEQ_ERR_ARG; }
        maxT = T > maxT ? T : maxT;
    }
    HIPCHK(hipSetDevice(s->dev));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int c = 0; c < s->C; ++c) {
        s->hd[c].TR = tr;
        s->hd[c].T_max = (s->ch[c].N + tr - 1) / tr + s->ch[c].B * s->ch[c].K;
    }
    s->TR = tr;
    s->maxT = maxT;
    HIPCHK(hipMemcpy(s->d_cd, s->hd.data(), sizeof(ChainDev) * s->C, hipMemcpyHostToDevice));
    drop_graphs(s);
    return BUSSEQ_OK;
}

extern "C" busseq_status busseq_autotune_batch(const busseq_problem* probs, int n_probs, int sweeps,
                                               int* tile_rows, double* sweeps_per_s)
{
    static const int cand[] = {16, 32, 64, 128};
    if (!probs || n_probs < 1 || n_probs > 4096 || sweeps < 1 || !tile_rows) {
        bq_set_error("bad arguments to busseq_autotune");
        return BUSSEQ_ERR_ARG;
    }
    std::vector<busseq_config> cfgs(n_probs);
    std::vector<busseq_problem> ps(n_probs);
    for (int p = 0; p < n_probs; ++p) {
        if (!probs[p].cfg) { bq_set_error("null config"); return BUSSEQ_ERR_ARG; }
        cfgs[p] = *probs[p].cfg;
        cfgs[p].tile_rows = cand[0];
        cfgs[p].keep_traces = 0;
        ps[p].cfg = &cfgs[p];
        ps[p].Y = probs[p].Y;
    }
    busseq_sampler* s = nullptr;
    busseq_status rc = busseq_create_batch(ps.data(), n_probs, &s);
    if (rc != BUSSEQ_OK) return rc;
    // two rounds over the candidates, each keeping its faster run: a drift of
    // the clock or a one-off delay during one timing does not pick the size
    const int nc = (int)(sizeof cand / sizeof cand[0]);
    double sps_of[sizeof cand / sizeof cand[0]] = {};
    for (int round = 0; round < 2; ++round) {
        for (int j = 0; j < nc; ++j) {
            double ms = 0.0;
            rc = set_tile_rows(s, cand[j]);
            if (rc == BUSSEQ_OK) rc = busseq_sweeps(s, 2);
            if (rc == BUSSEQ_OK) rc = busseq_time_sweeps(s, sweeps, &ms);
            if (rc != BUSSEQ_OK) { busseq_destroy(s); return rc; }
            const double sps = sweeps / (ms * 1e-3);
            if (sps > sps_of[j]) sps_of[j] = sps;
        }
    }
    int best = 64;
    double best_sps = -1.0;
    for (int j = 0; j < nc; ++j)
        if (sps_of[j] > best_sps) { best_sps = sps_of[j]; best = cand[j]; }
    busseq_destroy(s);
    *tile_rows = best;
    if (sweeps_per_s) *sweeps_per_s = best_sps;
    return BUSSEQ_OK;
}

extern "C" busseq_status busseq_autotune(const busseq_config* cfg, const int32_t* Y, int sweeps,
                                         int* tile_rows, double* sweeps_per_s)
{
    busseq_problem p;
    p.cfg = cfg;
    p.Y = Y;
    return busseq_autotune_batch(&p, 1, sweeps, tile_rows, sweeps_per_s);
}

extern "C" int busseq_kernel_count(void) { return kNumKernels; }

extern "C" const char* busseq_kernel_name(int j)
{
    return (j >= 0 && j < kNumKernels) ? kKernelNames[j] : "";
}

extern "C" busseq_status busseq_kernel_times(busseq_sampler* s, int n, double* ms)
{
    if (!s || n < 1 || !ms) { bq_set_error("bad arguments to busseq_kernel_times"); return BUSSEQ_ERR_ARG; }
    HIPCHK(hipSetDevice(s->dev));
    // the n sweeps run back to back, as in the graphs, with one wait at the end
    std::vector<hipEvent_t> ev((size_t)n * (kNumKernels + 1), nullptr);
    busseq_status rc = BUSSEQ_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) { e = nullptr; rc = BUSSEQ_ERR_DEVICE; break; }
    if (rc != BUSSEQ_OK) bq_set_error("hipEventCreate failed in busseq_kernel_times");
    for (int j = 0; j < kNumKernels; ++j) ms[j] = 0.0;
    for (int it = 0; it < n && rc == BUSSEQ_OK; ++it) {
        hipEvent_t* e = ev.data() + (size_t)it * (kNumKernels + 1);
        rc = s->prec == 32 ? enqueue_sweep<float>(s, e) : enqueue_sweep<double>(s, e);
        if (rc == BUSSEQ_OK) rc = count_sweep(s);
    }
    if (rc == BUSSEQ_OK && hipStreamSynchronize(s->stream) != hipSuccess) {
        bq_set_error("hipStreamSynchronize failed in busseq_kernel_times");
        rc = BUSSEQ_ERR_DEVICE;
    }
    for (int it = 0; it < n && rc == BUSSEQ_OK; ++it) {
        hipEvent_t* e = ev.data() + (size_t)it * (kNumKernels + 1);
        for (int j = 0; j < kNumKernels; ++j) {
            float f = 0.f;
            if (hipEventElapsedTime(&f, e[j], e[j + 1]) != hipSuccess) {
                bq_set_error("hipEventElapsedTime failed in busseq_kernel_times");
                rc = BUSSEQ_ERR_DEVICE;
                break;
            }
            ms[j] += f;
        }
    }
    for (int j = 0; j < kNumKernels; ++j) ms[j] /= n;
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

// streaming copy of known size; calibrates the HBM byte counters of a profiler
__global__ __launch_bounds__(256) void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    for (size_t j = (size_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (size_t)gridDim.x * 256) dst[j] = src[j];
}

extern "C" busseq_status busseq_copy_probe(size_t bytes, int device, double* ms)
{
    if (bytes < 16 || !ms) { bq_set_error("bad arguments to busseq_copy_probe"); return BUSSEQ_ERR_ARG; }
    HIPCHK(hipSetDevice(device));
    size_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n * 16));
    HIPCHK(hipMalloc(&b, n * 16));
    HIPCHK(hipMemset(a, 1, n * 16));
    HIPCHK(hipMemset(b, 0, n * 16));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // one warm-up launch, then the average of ten
    hipLaunchKernelGGL(k_copy_probe, dim3(8192), dim3(256), 0, nullptr, (const uint4*)a, (uint4*)b, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipEventRecord(e0, nullptr));
    for (int r = 0; r < 10; ++r)
        hipLaunchKernelGGL(k_copy_probe, dim3(8192), dim3(256), 0, nullptr, (const uint4*)a, (uint4*)b, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, e0, e1));
    *ms = f / 10.0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipFree(a));
    HIPCHK(hipFree(b));
    return BUSSEQ_OK;
}

extern "C" busseq_status busseq_get_state(busseq_sampler* s, int ch, busseq_state* st)
{
    if (!s || !st || ch < 0 || ch >= s->C) { bq_set_error("bad arguments to busseq_get_state"); return BUSSEQ_ERR_ARG; }
    HIPCHK(hipSetDevice(s->dev));
    HIPCHK(hipStreamSynchronize(s->stream));
    const ChainDev& c = s->hd[ch];
    const int N = c.N, G = c.G, Gp = c.Gp, K = c.K, B = c.B;
    if (st->X || st->Z) {
        std::vector<int32_t> S;
        HIPCHK(d2h(S, c.S, (size_t)N * Gp));
        for (int i = 0; i < N; ++i)
            for (int g = 0; g < G; ++g) {
                int32_t v = S[(size_t)i * Gp + g];
                if (st->X) st->X[(size_t)i * G + g] = v < 0 ? ~v : v;
                if (st->Z) st->Z[(size_t)i * G + g] = v < 0 ? 1 : 0;
            }
    }
    if (st->W) HIPCHK(hipMemcpy(st->W, c.W, 4 * (size_t)N, hipMemcpyDeviceToHost));
    if (st->alpha) HIPCHK(hipMemcpy(st->alpha, c.alpha, 8 * (size_t)G, hipMemcpyDeviceToHost));
    if (st->beta) HIPCHK(hipMemcpy(st->beta, c.beta, 8 * (size_t)K * G, hipMemcpyDeviceToHost));
    if (st->L) HIPCHK(hipMemcpy(st->L, c.L, 4 * (size_t)K * G, hipMemcpyDeviceToHost));
    if (st->nu) HIPCHK(hipMemcpy(st->nu, c.nu, 8 * (size_t)B * G, hipMemcpyDeviceToHost));
    if (st->delta) HIPCHK(hipMemcpy(st->delta, c.delta, 8 * (size_t)N, hipMemcpyDeviceToHost));
    if (st->phi) HIPCHK(hipMemcpy(st->phi, c.phi, 8 * (size_t)B * G, hipMemcpyDeviceToHost));
    if (st->gamma) HIPCHK(hipMemcpy(st->gamma, c.gamma, 16 * (size_t)B, hipMemcpyDeviceToHost));
    if (st->pi) HIPCHK(hipMemcpy(st->pi, c.pi, 8 * (size_t)B * K, hipMemcpyDeviceToHost));
    if (st->ptau) HIPCHK(hipMemcpy(st->ptau, c.ptau, 16, hipMemcpyDeviceToHost));
    return BUSSEQ_OK;
}

extern "C" busseq_status busseq_set_state(busseq_sampler* s, int ch, const busseq_state* st)
{
    if (!s || !st || ch < 0 || ch >= s->C) { bq_set_error("bad arguments to busseq_set_state"); return BUSSEQ_ERR_ARG; }
    HIPCHK(hipSetDevice(s->dev));
    HIPCHK(hipStreamSynchronize(s->stream));
    const ChainDev& c = s->hd[ch];
    const int N = c.N, G = c.G, Gp = c.Gp, K = c.K, B = c.B;
    if (st->X && st->Z) {
        std::vector<int32_t> S;
        HIPCHK(d2h(S, c.S, (size_t)N * Gp));
        for (int i = 0; i < N; ++i)
            for (int g = 0; g < G; ++g) {
                int32_t& v = S[(size_t)i * Gp + g];
                if (v > 0) continue;   // observed count: X = Y, Z = 0 stay fixed
                int32_t x = st->X[(size_t)i * G + g], z = st->Z[(size_t)i * G + g];
                v = z ? ~x : 0;
            }
        HIPCHK(hipMemcpy(c.S, S.data(), S.size() * 4, hipMemcpyHostToDevice));
        // the compact copy of the Y = 0 entries follows S
        std::vector<int32_t> zrow(N + 1);
        HIPCHK(hipMemcpy(zrow.data(), c.zrow, 4 * ((size_t)N + 1), hipMemcpyDeviceToHost));
        const size_t nz = (size_t)zrow[N];
        if (nz > 0) {
            std::vector<uint16_t> zcol(nz);
            std::vector<int32_t> zs(nz);
            HIPCHK(hipMemcpy(zcol.data(), c.zcol, 2 * nz, hipMemcpyDeviceToHost));
            for (int i = 0; i < N; ++i)
                for (int q = zrow[i]; q < zrow[i + 1]; ++q) zs[q] = S[(size_t)i * Gp + zcol[q]];
            HIPCHK(hipMemcpy(c.zs, zs.data(), 4 * nz, hipMemcpyHostToDevice));
        }
    }
    auto put = [&](void* dst, const void* src, size_t n) -> hipError_t {
        return src ? hipMemcpy(dst, src, n, hipMemcpyHostToDevice) : hipSuccess;
    };
    HIPCHK(put(c.W, st->W, 4 * (size_t)N));
    HIPCHK(put(c.alpha, st->alpha, 8 * (size_t)G));
    HIPCHK(put(c.beta, st->beta, 8 * (size_t)K * G));
    HIPCHK(put(c.L, st->L, 4 * (size_t)K * G));
    HIPCHK(put(c.nu, st->nu, 8 * (size_t)B * G));
    HIPCHK(put(c.delta, st->delta, 8 * (size_t)N));
    HIPCHK(put(c.phi, st->phi, 8 * (size_t)B * G));
    HIPCHK(put(c.gamma, st->gamma, 16 * (size_t)B));
    HIPCHK(put(c.pi, st->pi, 8 * (size_t)B * K));
    HIPCHK(put(c.ptau, st->ptau, 16));
    std::vector<double> al, be, nu, de, ph;
    HIPCHK(d2h(al, c.alpha, G));
    HIPCHK(d2h(be, c.beta, (size_t)K * G));
    HIPCHK(d2h(nu, c.nu, (size_t)B * G));
    HIPCHK(d2h(de, c.delta, N));
    HIPCHK(d2h(ph, c.phi, (size_t)B * G));
    busseq_status rc = upload_tcopies(s, ch, al, be, nu, de, ph);
    if (rc != BUSSEQ_OK) return rc;
    return refresh_prep(s);
}

namespace {

// The BIC's device buffers: the parameters packed as alpha[G] | beta[K*G] |
// nu[B*G] | delta[N] | phi[B*G] | gamma[2B] | pi[B*K], then N*K row sums, then
// the result.  Sizes in doubles.
struct BicLayout {
    size_t a, be, nu, de, ph, ga, pi, np, ll, out, total;
    explicit BicLayout(const ChainDev& c)
    {
        const size_t G = c.G, K = c.K, B = c.B, N = c.N;
        a = 0; be = a + G; nu = be + K * G; de = nu + B * G; ph = de + N; ga = ph + B * G;
        pi = ga + 2 * B; np = pi + B * K;
        ll = np; out = ll + N * K; total = out + 1;
    }
};

// the BIC of chain ch at the given host parameters into buf (bl.total doubles
// of device memory), result copied to *bic; every word of buf that the kernels
// read is written first, so buf need not be cleared
busseq_status bic_into(busseq_sampler* s, int ch, const double* alpha, const double* beta, const double* nu,
                       const double* delta, const double* phi, const double* gamma, const double* pi,
                       double* buf, double* bic)
{
    const ChainDev& c = s->hd[ch];
    const BicLayout bl(c);
    const size_t G = c.G, K = c.K, B = c.B, N = c.N;
    hipStream_t st = s->stream;
    HIPCHK(hipMemcpyAsync(buf + bl.a, alpha, 8 * G, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(buf + bl.be, beta, 8 * K * G, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(buf + bl.nu, nu, 8 * B * G, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(buf + bl.de, delta, 8 * N, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(buf + bl.ph, phi, 8 * B * G, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(buf + bl.ga, gamma, 8 * 2 * B, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(buf + bl.pi, pi, 8 * B * K, hipMemcpyHostToDevice, st));
    const ChainDev* cd = s->d_cd + ch;
    hipLaunchKernelGGL(k_bic, dim3((unsigned)((N * K + 3) / 4)), dim3(256), 0, st, cd, buf, buf + bl.ll);
    hipLaunchKernelGGL(k_bic_cells, dim3(nblk((long)N)), dim3(256), 0, st, cd, buf + bl.pi, buf + bl.ll);
    hipLaunchKernelGGL(k_bic_total, dim3(1), dim3(64), 0, st, cd, buf + bl.ll, buf + bl.out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(bic, buf + bl.out, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return BUSSEQ_OK;
}

busseq_status bic_args(busseq_sampler* s, int ch, const double* alpha, const double* beta, const double* nu,
                       const double* delta, const double* phi, const double* gamma, const double* pi, double* bic)
{
    if (!s || ch < 0 || ch >= s->C || !alpha || !beta || !nu || !delta || !phi || !gamma || !pi || !bic) {
        bq_set_error("bad arguments to busseq_bic");
        return BUSSEQ_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->dev));
    HIPCHK(hipStreamSynchronize(s->stream));
    return BUSSEQ_OK;
}

} // namespace

extern "C" busseq_status busseq_bic(busseq_sampler* s, int ch, const double* alpha, const double* beta,
                                    const double* nu, const double* delta, const double* phi,
                                    const double* gamma, const double* pi, double* bic)
{
    busseq_status rc = bic_args(s, ch, alpha, beta, nu, delta, phi, gamma, pi, bic);
    if (rc != BUSSEQ_OK) return rc;
    const BicLayout bl(s->hd[ch]);
    double* buf = nullptr;
    HIPCHK(hipMalloc(&buf, bl.total * sizeof(double)));
    rc = bic_into(s, ch, alpha, beta, nu, delta, phi, gamma, pi, buf, bic);
    (void)hipStreamSynchronize(s->stream);
    (void)hipFree(buf);
    return rc;
}

// ---------------------------------------------------------------------------
// busseq_bic_noalloc: the kernels of busseq_bic on a file-scope device array
// instead of a buffer allocated per call.  Each device holds its own copy of
// the array from the library's load to the end of the process; its address on
// a device is taken once, with that device current, and kept per device.
// ---------------------------------------------------------------------------
#define BQ_SCRATCH_DOUBLES (4u << 20)   // 32 MiB per device
#define BQ_SCRATCH_DEVICES 64
static __device__ __align__(256) double scratch[BQ_SCRATCH_DOUBLES];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[BQ_SCRATCH_DEVICES];

// device address of this device's scratch (the current device is dev)
static busseq_status scratch_of(int dev, double** out)
{
    if (dev < 0 || dev >= BQ_SCRATCH_DEVICES) {
        bq_set_error("busseq_bic_noalloc: device ordinal beyond " + std::to_string(BQ_SCRATCH_DEVICES - 1));
        return BUSSEQ_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return BUSSEQ_OK;
}

extern "C" busseq_status busseq_bic_noalloc(busseq_sampler* s, int ch, const double* alpha, const double* beta,
                                            const double* nu, const double* delta, const double* phi,
                                            const double* gamma, const double* pi, double* bic)
{
    busseq_status rc = bic_args(s, ch, alpha, beta, nu, delta, phi, gamma, pi, bic);
    if (rc != BUSSEQ_OK) return rc;
    const BicLayout bl(s->hd[ch]);
    if (bl.total > BQ_SCRATCH_DOUBLES) {
        bq_set_error("busseq_bic_noalloc: this chain needs " + std::to_string(bl.total) +
                     " doubles of scratch (N*K + G*(K+2B+1) + N + 2B + B*K + 1), more than its " +
                     std::to_string(BQ_SCRATCH_DOUBLES) + " (32 MiB); use busseq_bic");
        return BUSSEQ_ERR_ARG;
    }
    double* buf = nullptr;
    rc = scratch_of(s->dev, &buf);
    if (rc != BUSSEQ_OK) return rc;
    return bic_into(s, ch, alpha, beta, nu, delta, phi, gamma, pi, buf, bic);
}

extern "C" busseq_status busseq_finish(busseq_sampler* s, int ch, busseq_result* r)
{
    if (!s || !r || ch < 0 || ch >= s->C) { bq_set_error("bad arguments to busseq_finish"); return BUSSEQ_ERR_ARG; }
    HIPCHK(hipSetDevice(s->dev));
    HIPCHK(hipStreamSynchronize(s->stream));
    std::memset(r, 0, sizeof(*r));
    const ChainDev& c = s->hd[ch];
    const ChainHost& h = s->ch[ch];
    const int N = c.N, G = c.G, K = c.K, B = c.B;
    int nrec = 0;
    HIPCHK(hipMemcpy(&nrec, c.n_rec, 4, hipMemcpyDeviceToHost));
    if (nrec <= 0) { bq_set_error("no recorded iterations (n_iter must exceed n_burnin)"); return BUSSEQ_ERR_ARG; }
    const double n = nrec;
    auto md = [](size_t k) { return (double*)std::calloc(k ? k : 1, 8); };
    auto mi = [](size_t k) { return (int32_t*)std::calloc(k ? k : 1, 4); };
    r->B = B; r->N = N; r->G = G; r->K = K;
    r->n_pres = c.traces ? (c.n_pres < nrec ? c.n_pres : nrec) : 0;
    r->n_b = (int*)std::calloc(B, sizeof(int));
    for (int b = 0; b < B; ++b) r->n_b[b] = h.n_b[b];
    r->X = mi((size_t)N * G); r->Z = mi((size_t)N * G); r->W = mi(N);
    r->gamma = md(2 * (size_t)B); r->alpha = md(G); r->L = md((size_t)K * G); r->beta = md((size_t)K * G);
    r->ptau = md(2); r->nu = md((size_t)B * G); r->delta = md(N); r->phi = md((size_t)B * G); r->pi = md((size_t)B * K);
    if (!r->X || !r->Z || !r->W) { bq_set_error("host allocation failed"); return BUSSEQ_ERR_NOMEM; }
    busseq_state st{};
    st.X = r->X; st.Z = r->Z;
    busseq_status rc = busseq_get_state(s, ch, &st);
    if (rc != BUSSEQ_OK) return rc;
    std::vector<double> v;
    std::vector<int32_t> vi;
    HIPCHK(d2h(v, c.s_gamma, 2 * (size_t)B)); for (int j = 0; j < 2 * B; ++j) r->gamma[j] = v[j] / n;
    HIPCHK(d2h(v, c.s_alpha, G)); for (int j = 0; j < G; ++j) r->alpha[j] = v[j] / n;
    HIPCHK(d2h(v, c.s_beta, (size_t)K * G)); for (size_t j = 0; j < (size_t)K * G; ++j) r->beta[j] = v[j] / n;
    HIPCHK(d2h(vi, c.c_L, (size_t)K * G)); for (size_t j = 0; j < (size_t)K * G; ++j) r->L[j] = (double)vi[j] / n;
    HIPCHK(d2h(v, c.s_nu, (size_t)B * G)); for (size_t j = 0; j < (size_t)B * G; ++j) r->nu[j] = v[j] / n;
    HIPCHK(d2h(v, c.s_phi, (size_t)B * G)); for (size_t j = 0; j < (size_t)B * G; ++j) r->phi[j] = v[j] / n;
    HIPCHK(d2h(v, c.s_delta, N)); for (int j = 0; j < N; ++j) r->delta[j] = v[j] / n;
    HIPCHK(d2h(v, c.s_pi, (size_t)B * K)); for (size_t j = 0; j < (size_t)B * K; ++j) r->pi[j] = v[j] / n;
    HIPCHK(d2h(v, c.s_ptau, 2)); r->ptau[0] = v[0] / n; r->ptau[1] = v[1] / n;
    HIPCHK(d2h(vi, c.c_W, (size_t)N * K));
    for (int i = 0; i < N; ++i) {
        int best = 0, bv = vi[(size_t)i * K];
        for (int k = 1; k < K; ++k) if (vi[(size_t)i * K + k] > bv) { bv = vi[(size_t)i * K + k]; best = k; }
        r->W[i] = best;
    }
    // BIC at the posterior means (cu:1815-1833)
    rc = busseq_bic(s, ch, r->alpha, r->beta, r->nu, r->delta, r->phi, r->gamma, r->pi, &r->bic);
    if (rc != BUSSEQ_OK) return rc;
    if (c.traces && r->n_pres > 0) {
        // the streamed host traces ([param * n_pres + iter]) with the stride the
        // result reports: after fewer sweeps than n_iter only nrec were recorded
        rc = trace_flush(s);
        if (rc != BUSSEQ_OK) return rc;
        const TraceHost& t = s->th[ch];
        const size_t P = (size_t)r->n_pres, Pd = (size_t)t.P;
        void** dst[BQ_TRACE_ARRAYS] = {(void**)&r->tr_gamma, (void**)&r->tr_alpha, (void**)&r->tr_L, (void**)&r->tr_beta,
                                       (void**)&r->tr_ptau, (void**)&r->tr_nu, (void**)&r->tr_delta, (void**)&r->tr_phi,
                                       (void**)&r->tr_W, (void**)&r->tr_pi};
        for (int a = 0; a < BQ_TRACE_ARRAYS; ++a) {
            const size_t es = (size_t)t.esz[a];
            char* d = (char*)std::calloc(t.np[a] * P + 1, es);
            *dst[a] = d;
            if (!d) { bq_set_error("host allocation failed"); return BUSSEQ_ERR_NOMEM; }
            const char* src = (const char*)t.host[a];
            if (!src) continue;
            if (P == Pd) std::memcpy(d, src, t.np[a] * P * es);
            else
                for (size_t j = 0; j < t.np[a]; ++j) std::memcpy(d + j * P * es, src + j * Pd * es, P * es);
        }
    }
    return BUSSEQ_OK;
}

extern "C" busseq_status busseq_run(const busseq_config* cfg, const int32_t* Y, busseq_result* out)
{
    busseq_sampler* s = nullptr;
    busseq_status rc = busseq_create(cfg, Y, &s);
    if (rc != BUSSEQ_OK) return rc;
    rc = busseq_sweeps(s, cfg->n_iter);
    if (rc == BUSSEQ_OK) rc = busseq_finish(s, 0, out);
    busseq_destroy(s);
    return rc;
}
