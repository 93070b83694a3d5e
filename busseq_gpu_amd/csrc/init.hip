// init.hip -- the passes of the initial state over the N x G counts
// (BUSseq_gpu.cu:1163-1244): the state matrix S, the lists of zero entries and
// of large counts, library sizes, the static histogram of the gamma step and
// the raw-mean sums of alpha0/beta0/nu0.  The reference runs these as about
// 272k launches of one-block means (cu:1204-1244); here they are three
// streaming passes.  Every sum is exact (integers, or fixed-point values added
// as wrapping 64-bit integers), so the result does not depend on the launch
// shape or the order of the atomics and equals the CPU oracle's bit for bit.
#include <hip/hip_runtime.h>

#include "../../include/busseq/model.h"
#include "engine.h"

#define H BQ_HIST

namespace {

// One wave per cell.  Each wave keeps its own LDS histogram of the row and
// adds the bins it used to the batch's global histogram.
__global__ __launch_bounds__(256) void k_init_rows(const int32_t* __restrict__ Y, int N, int G, int Gp,
                                                   const int32_t* __restrict__ batch, int32_t* __restrict__ S,
                                                   int32_t* __restrict__ nz_row, int32_t* __restrict__ nbig_row,
                                                   int64_t* __restrict__ lib, unsigned long long* __restrict__ hist)
{
    __shared__ uint32_t h[4][H];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wv;
    uint32_t* hw = h[wv];
    for (int q = lane; q < H; q += 64) hw[q] = 0;
    __syncthreads();
    const bool live = i < N;
    int64_t s = 0;
    int nz = 0, nb = 0;
    if (live) {
        const int32_t* yr = Y + (size_t)i * G;
        int32_t* sr = S + (size_t)i * Gp;
        for (int g = lane; g < Gp; g += 64) {
            int32_t v = 0;
            if (g < G) {
                const int32_t y = yr[g];
                v = y > 0 ? y : 0;
                nz += v == 0;
                nb += v >= H;
                s += v;
                if (v > 0 && v < H) atomicAdd(&hw[v], 1u);
            }
            sr[g] = v;
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        nz += __shfl_xor(nz, m, 64);
        nb += __shfl_xor(nb, m, 64);
        s += __shfl_xor(s, m, 64);
    }
    __syncthreads();
    if (!live) return;
    if (lane == 0) {
        nz_row[i] = nz;
        nbig_row[i] = nb;
        lib[i] = s;
    }
    const size_t hb = (size_t)batch[i] * H;
    for (int q = lane; q < H; q += 64)
        if (hw[q]) atomicAdd(&hist[hb + q], (unsigned long long)hw[q]);
}

// One wave per cell: a ballot per 64 genes places each listed entry at its
// rank, so the lists are in gene order within the row.
__global__ __launch_bounds__(256) void k_init_lists(const int32_t* __restrict__ S, int N, int G, int Gp,
                                                    const int32_t* __restrict__ zrow, const int32_t* __restrict__ brow,
                                                    uint16_t* __restrict__ zcol, int32_t* __restrict__ big)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= N) return;
    const int32_t* sr = S + (size_t)i * Gp;
    const uint64_t below = lane ? ~0ull >> (64 - lane) : 0ull;
    int zb = zrow[i], bb = brow[i];
    for (int g0 = 0; g0 < G; g0 += 64) {
        const int g = g0 + lane;
        const int32_t v = g < G ? sr[g] : 1;
        const bool z = g < G && v == 0, l = g < G && v >= H;
        const uint64_t mz = __ballot(z), ml = __ballot(l);
        if (z) zcol[zb + __popcll(mz & below)] = (uint16_t)g;
        if (l) big[bb + __popcll(ml & below)] = v;
        zb += __popcll(mz);
        bb += __popcll(ml);
    }
}

// Block = 256 genes x one chunk of cells of one batch.  Thread g keeps the
// chunk's subtype sums of gene g in LDS (its own column: no atomics), BQ_KREG
// subtypes per pass over the chunk's cells, and adds the nonzero ones to rs.
__global__ __launch_bounds__(256) void k_init_raw(const int32_t* __restrict__ S, int G, int Gp, int K,
                                                  const int32_t* __restrict__ W, const double* __restrict__ ie,
                                                  const int32_t* __restrict__ chunks,
                                                  unsigned long long* __restrict__ rs)
{
    __shared__ uint64_t acc[BQ_KREG][256];
    const int t = threadIdx.x;
    const int g = blockIdx.y * 256 + t;
    const int b = chunks[3 * blockIdx.x], st = chunks[3 * blockIdx.x + 1], n = chunks[3 * blockIdx.x + 2];
    if (g >= G) return;
    for (int kc = 0; kc < K; kc += BQ_KREG) {
        for (int k = 0; k < BQ_KREG; ++k) acc[k][t] = 0;
        for (int r = st; r < st + n; ++r) {
            const int32_t y = S[(size_t)r * Gp + g];
            const int k = W[r] - kc;
            if (y > 0 && k >= 0 && k < BQ_KREG) acc[k][t] += bq_fix(bq_log1p((double)y * ie[r]));
        }
        for (int k = 0; k < BQ_KREG && kc + k < K; ++k) {
            const uint64_t v = acc[k][t];
            if (v) atomicAdd(&rs[((size_t)b * K + kc + k) * G + g], (unsigned long long)v);
        }
    }
}

// Per-(batch, gene) histogram of the counts 1 <= Y < BQ_PHI_HV (S holds Y there).
// A thread takes one gene over a run of BQ_HB_CELLS cells and counts into LDS bins
// of its own (bins[v * 256 + thread]: the lanes of a wave stay on distinct banks
// whatever their values), so the counting needs neither atomics nor barriers.  At
// the end of the run, and where the batch changes inside it, the thread adds its
// non-zero bins to h with integer atomics, so any order gives the same counts.  A
// bin holds at most BQ_HB_CELLS < 2^16 cells between two flushes.
#define BQ_HB_CELLS 1024

__device__ __forceinline__ void hist_flush(uint16_t* my, uint32_t* __restrict__ h, size_t item)
{
    for (int v = 1; v < BQ_PHI_HV; ++v) {
        const uint32_t n = my[v * 256];
        if (n) {
            atomicAdd(&h[item * BQ_PHI_HV + v], n);
            my[v * 256] = 0;
        }
    }
}

__global__ __launch_bounds__(256) void k_init_hist_bg(const int32_t* __restrict__ S, int N, int G, int Gp,
                                                      const int32_t* __restrict__ batch, uint32_t* __restrict__ h)
{
    __shared__ uint16_t bins[BQ_PHI_HV * 256];
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= G) return;
    uint16_t* my = bins + threadIdx.x;
    for (int v = 0; v < BQ_PHI_HV; ++v) my[v * 256] = 0;
    for (int i0 = blockIdx.y * BQ_HB_CELLS; i0 < N; i0 += gridDim.y * BQ_HB_CELLS) {
        const int i1 = i0 + BQ_HB_CELLS < N ? i0 + BQ_HB_CELLS : N;
        int b = batch[i0];
        for (int i = i0; i < i1; ++i) {
            const int bi = batch[i];
            if (bi != b) {
                hist_flush(my, h, (size_t)b * G + g);
                b = bi;
            }
            const int32_t v = S[(size_t)i * Gp + g];
            if (v >= 1 && v < BQ_PHI_HV) my[v * 256] += 1;
        }
        hist_flush(my, h, (size_t)b * G + g);
    }
}

// The static histogram as a list: the number of distinct counts 1 <= v < BQ_PHI_HV
// of each item j = b G + g (one thread per item), then the items' entries
// (j << 32) | (n << 7) | v at off[j], v ascending (n, a number of cells, is below 2^25)
__global__ __launch_bounds__(256) void k_hist_distinct(const uint32_t* __restrict__ h, long long BG,
                                                       uint32_t* __restrict__ cnt)
{
    const long long j = (long long)blockIdx.x * 256 + threadIdx.x;
    if (j >= BG) return;
    uint32_t m = 0;
    for (int v = 1; v < BQ_PHI_HV; ++v) m += h[j * BQ_PHI_HV + v] != 0u;
    cnt[j] = m;
}

__global__ __launch_bounds__(256) void k_hist_list(const uint32_t* __restrict__ h, long long BG,
                                                   const uint64_t* __restrict__ off, uint64_t* __restrict__ list)
{
    const long long j = (long long)blockIdx.x * 256 + threadIdx.x;
    if (j >= BG) return;
    uint64_t o = off[j];
    for (int v = 1; v < BQ_PHI_HV; ++v) {
        const uint32_t n = h[j * BQ_PHI_HV + v];
        if (n) list[o++] = ((uint64_t)j << 32) | ((uint64_t)n << 7) | (uint64_t)v;
    }
}

} // namespace

int bq_dev_hist_distinct(const uint32_t* hist_bg, long long BG, uint32_t* cnt, void* stream)
{
    if (BG <= 0) return (int)hipSuccess;
    hipLaunchKernelGGL(k_hist_distinct, dim3((unsigned)((BG + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       hist_bg, BG, cnt);
    return (int)hipGetLastError();
}

int bq_dev_hist_list(const uint32_t* hist_bg, long long BG, const uint64_t* off, uint64_t* list, void* stream)
{
    if (BG <= 0) return (int)hipSuccess;
    hipLaunchKernelGGL(k_hist_list, dim3((unsigned)((BG + 255) / 256)), dim3(256), 0, (hipStream_t)stream, hist_bg,
                       BG, off, list);
    return (int)hipGetLastError();
}

int bq_dev_init_rows(const int32_t* Yd, int N, int G, int Gp, const int32_t* batch, int32_t* S,
                     int32_t* nz_row, int32_t* nbig_row, int64_t* lib, int64_t* hist, void* stream)
{
    if (N <= 0) return (int)hipSuccess;
    hipLaunchKernelGGL(k_init_rows, dim3((N + 3) / 4), dim3(256), 0, (hipStream_t)stream, Yd, N, G, Gp, batch, S,
                       nz_row, nbig_row, lib, (unsigned long long*)hist);
    return (int)hipGetLastError();
}

int bq_dev_init_lists(const int32_t* S, int N, int G, int Gp, const int32_t* zrow, const int32_t* brow,
                      uint16_t* zcol, int32_t* big, void* stream)
{
    if (N <= 0) return (int)hipSuccess;
    hipLaunchKernelGGL(k_init_lists, dim3((N + 3) / 4), dim3(256), 0, (hipStream_t)stream, S, N, G, Gp, zrow, brow,
                       zcol, big);
    return (int)hipGetLastError();
}

int bq_dev_init_raw(const int32_t* S, int G, int Gp, int K, const int32_t* W, const double* ie,
                    const int32_t* chunks, int n_chunks, uint64_t* rs, void* stream)
{
    if (n_chunks <= 0) return (int)hipSuccess;
    if (K < 1 || K > BQ_KMAX) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_init_raw, dim3(n_chunks, (G + 255) / 256), dim3(256), 0, (hipStream_t)stream, S, G, Gp, K,
                       W, ie, chunks, (unsigned long long*)rs);
    return (int)hipGetLastError();
}

int bq_dev_init_hist_bg(const int32_t* S, int N, int G, int Gp, const int32_t* batch, uint32_t* hist_bg,
                        void* stream)
{
    const int runs = (N + BQ_HB_CELLS - 1) / BQ_HB_CELLS;
    const int gy = runs < 65535 ? runs : 65535;
    hipLaunchKernelGGL(k_init_hist_bg, dim3((G + 255) / 256, gy), dim3(256), 0, (hipStream_t)stream, S, N, G, Gp,
                       batch, hist_bg);
    return (int)hipGetLastError();
}
