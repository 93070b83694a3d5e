// host.cpp -- host side of libbusseq: initial state (BUSseq_gpu.cu:1163-1244),
// input readers (cu:1094-1098, 1145-1159), output writers (cu:1610-2018),
// the synthetic-data generator and error reporting.  Compiled with g++
// (-ffp-contract=off) so that the initial state uses the same numerics as the
// device code.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../../include/busseq.h"
#include "../../include/busseq/model.h"
#include "engine.h"

static thread_local std::string g_err;

void bq_set_error(const std::string& m) { g_err = m; }

extern "C" const char* busseq_last_error(void) { return g_err.c_str(); }

extern "C" void busseq_config_default(busseq_config* c)
{
    std::memset(c, 0, sizeof(*c));
    c->B = 4; c->n_b = nullptr; c->G = 3000; c->K = 5;
    c->seed = 0; c->n_iter = 4000; c->n_burnin = 2000; c->n_unchanged = 500;
    c->keep_traces = 0; c->precision = 32; c->chains = 1; c->chain0 = 0;
    c->tile_rows = 0; c->device = 0;   /* tile_rows 0: busseq_default_tile_rows */
}

// ---------------------------------------------------------------------------
// initial state, BUSseq_gpu.cu:1163-1244.  The passes over the N x G counts
// (state matrix, zero-entry lists, library sizes, raw-mean sums) run on the
// device (init.hip); the closed forms over cells and over (batch, subtype,
// gene) run here, in two stages around the raw-mean pass.
// ---------------------------------------------------------------------------

// stage 1: pi0, W0 and delta0 of one chain from the library sizes
void bq_init_cells(int B, const int* n_b, int K, const int64_t* lib, uint32_t seed, uint32_t chain, HostInit& o)
{
    std::vector<int> boff(B + 1, 0);
    for (int b = 0; b < B; ++b) boff[b + 1] = boff[b] + n_b[b];
    const int N = boff[B];
    // pi0_bk = (k+1)/(K(K+1)/2)  (cu:392-397)
    o.pi.assign((size_t)B * K, 0.0);
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < K; ++k) o.pi[(size_t)b * K + k] = (k + 1) / (K * (K + 1) / 2.0);
    // W0 ~ Cat(pi0_b); delta0_i = log s_i - log s_first(b)  (cu:419-443)
    o.W.assign(N, 0);
    o.delta.assign(N, 0.0);
    for (int b = 0; b < B; ++b)
        for (int i = boff[b]; i < boff[b + 1]; ++i) {
            bq_rng r;
            bq_rng_init(&r, seed, chain, (uint32_t)i, 0u, BQ_SWEEP_INIT, BQ_PH_INIT_W);
            double u = bq_uniform(&r);
            int k = 0;
            double tmp = o.pi[(size_t)b * K];
            while (u > tmp && k < K - 1) { k++; tmp += o.pi[(size_t)b * K + k]; }
            o.W[i] = k;
            o.delta[i] = bq_log((double)lib[i]) - bq_log((double)lib[boff[b]]);
        }
    o.mu_delta = o.delta;
}

// stage 2: the remaining parameters from the exact sums
// rs[(b*K + k)*G + g] = sum over cells i of batch b with W0_i = k of fix(log1p(Y_ig e^-delta0_i))
void bq_init_params(int B, const int* n_b, int G, int K, const uint64_t* rs, uint32_t seed, uint32_t chain,
                    HostInit& o)
{
    const int N = (int)o.W.size();
    std::vector<int> boff(B + 1, 0), nbk((size_t)B * K, 0);
    for (int b = 0; b < B; ++b) boff[b + 1] = boff[b] + n_b[b];
    for (int b = 0; b < B; ++b)
        for (int i = boff[b]; i < boff[b + 1] && i < N; ++i) nbk[(size_t)b * K + o.W[i]] += 1;
    // raw means r_bkg (cu:369-382, 1210-1221)
    std::vector<double> rm((size_t)B * K * G, 0.0);
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < K; ++k)
            for (int g = 0; g < G; ++g) {
                size_t j = ((size_t)b * K + k) * G + g;
                int n = nbk[(size_t)b * K + k];
                rm[j] = n == 0 ? 0.0 : bq_unfix(rs[j]) / n;
            }
    // mu_nu_b (cu:470-482, 1232-1241)
    std::vector<double> bm(B, 0.0);
    for (int b = 0; b < B; ++b) {
        unsigned __int128 tot = 0;
        for (int k = 0; k < K; ++k)
            for (int g = 0; g < G; ++g)
                tot += (unsigned __int128)(__int128)(int64_t)rs[((size_t)b * K + k) * G + g];
        bm[b] = (double)(__int128)tot * BQ_FIX_INV / ((double)n_b[b] * G);
    }
    o.mu_nu.assign(B, 0.0);
    for (int b = 1; b < B; ++b) o.mu_nu[b] = bm[b] - bm[0];
    // gamma0 = (0, -0.3) (cu:384-390)
    o.gamma.assign(2 * (size_t)B, 0.0);
    for (int b = 0; b < B; ++b) { o.gamma[b] = BQ_GAMMA0_INIT; o.gamma[B + b] = BQ_GAMMA1_INIT; }
    // phi0 = 5, nu0 (cu:399-417)
    o.phi.assign((size_t)B * G, BQ_PHI_INIT);
    o.nu.assign((size_t)B * G, 0.0);
    for (int b = 1; b < B; ++b)
        for (int g = 0; g < G; ++g) {
            double t1 = rm[((size_t)b * K) * G + g], t2 = rm[g];
            for (int k = 1; k < K; ++k) {
                t1 += rm[((size_t)b * K + k) * G + g];
                t2 += rm[(size_t)k * G + g];
            }
            o.nu[(size_t)b * G + g] = (t1 - t2) / K;
        }
    // p0 ~ U(0, 0.5) from the Philox stream, tau0 = 0.005 (cu:1172-1175)
    o.ptau.assign(2, 0.0);
    {
        bq_rng r;
        bq_rng_init(&r, seed, chain, 0u, 0u, BQ_SWEEP_INIT, BQ_PH_INIT_P);
        o.ptau[0] = bq_uniform(&r) * BQ_P_INIT_BOUND;
        o.ptau[1] = BQ_TAU0_INIT;
    }
    // alpha0, beta0, L0 (cu:445-462, 1186, 1227-1230)
    double p = o.ptau[0];
    double lrb = bq_log(p / (1.0 - p)) + (bq_log(BQ_TAU0_INIT) - bq_log(BQ_TAU1)) / 2.0;
    o.alpha.assign(G, 0.0);
    o.mu_alpha.assign(G, 0.0);
    o.beta.assign((size_t)K * G, 0.0);
    o.L.assign((size_t)K * G, 0);
    for (int g = 0; g < G; ++g) { o.alpha[g] = rm[g]; o.mu_alpha[g] = rm[g]; }
    for (int k = 1; k < K; ++k)
        for (int g = 0; g < G; ++g) {
            size_t j = (size_t)k * G + g;
            o.beta[j] = rm[j] - rm[g];
            double lr = lrb + (o.beta[j] * o.beta[j]) * (1.0 / BQ_TAU0_INIT - 1.0 / BQ_TAU1) / 2.0;
            bq_rng r;
            bq_rng_init(&r, seed, chain, (uint32_t)g, (uint32_t)k, BQ_SWEEP_INIT, BQ_PH_INIT_L);
            o.L[j] = bq_bern_logodds(bq_uniform(&r), lr);
        }
}

// ---------------------------------------------------------------------------
// input (cu:1094-1098, 1145-1159), strict: malformed or short input is an error
// ---------------------------------------------------------------------------
static bool read_file(const char* path, std::vector<char>& buf)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) { bq_set_error(std::string("cannot open ") + path + ": " + std::strerror(errno)); return false; }
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) { std::fclose(f); bq_set_error(std::string("cannot size ") + path); return false; }
    buf.resize((size_t)n + 1);
    size_t got = n ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    if (got != (size_t)n) { bq_set_error(std::string("short read of ") + path); return false; }
    buf[(size_t)n] = 0;
    return true;
}

// parse up to `want` whitespace-separated non-negative ints; returns count parsed, -1 on a bad token
static long parse_ints(const char* p, const char* end, long want, int32_t* out, std::string& bad)
{
    long n = 0;
    while (n < want) {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\f' || *p == '\v')) ++p;
        if (p >= end) break;
        const char* tok = p;
        bool neg = false;
        if (*p == '+' || *p == '-') { neg = *p == '-'; ++p; }
        int64_t v = 0;
        const char* d0 = p;
        while (p < end && *p >= '0' && *p <= '9') { v = v * 10 + (*p - '0'); if (v > 2147483647LL) break; ++p; }
        bool ws_end = p >= end || *p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\f' || *p == '\v';
        if (p == d0 || !ws_end || v > 2147483647LL) {
            const char* q = tok;
            while (q < end && q - tok < 24 && !(*q == ' ' || *q == '\t' || *q == '\n' || *q == '\r')) ++q;
            bad.assign(tok, q);
            return -1;
        }
        out[n++] = (int32_t)(neg ? -v : v);
    }
    return n;
}

extern "C" busseq_status busseq_read_dims(const char* path, int B, int* n_b)
{
    std::vector<char> buf;
    if (!read_file(path, buf)) return BUSSEQ_ERR_IO;
    std::string bad;
    std::vector<int32_t> v(B > 0 ? B : 1);
    long n = parse_ints(buf.data(), buf.data() + buf.size() - 1, B, v.data(), bad);
    if (n < 0) { bq_set_error(std::string(path) + ": not an integer: '" + bad + "'"); return BUSSEQ_ERR_FORMAT; }
    if (n < B) {
        bq_set_error(std::string(path) + ": expected " + std::to_string(B) + " batch sizes, found " + std::to_string(n));
        return BUSSEQ_ERR_FORMAT;
    }
    for (int b = 0; b < B; ++b) n_b[b] = v[b];
    return BUSSEQ_OK;
}

static inline bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }

// Binary count file: the 32-byte header below, then N*G little-endian int32,
// cells as rows.  Written by busseq_write_counts_binary; busseq_read_counts
// recognises it by the magic, so either form can be given to -c.
static const char kBinMagic[8] = {'B', 'Q', 'C', 'O', 'U', 'N', 'T', '1'};
struct BinHeader {
    char magic[8];
    int64_t N, G, reserved;
};

struct MappedFile {
    const char* p = nullptr;
    size_t n = 0;
    int fd = -1;
    ~MappedFile()
    {
        if (p && n) munmap((void*)p, n);
        if (fd >= 0) close(fd);
    }
    bool open_ro(const char* path)
    {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) { bq_set_error(std::string("cannot open ") + path + ": " + std::strerror(errno)); return false; }
        struct stat st;
        if (fstat(fd, &st) != 0) { bq_set_error(std::string("cannot size ") + path); return false; }
        n = (size_t)st.st_size;
        if (n == 0) return true;
        void* q = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (q == MAP_FAILED) { bq_set_error(std::string("cannot map ") + path + ": " + std::strerror(errno)); n = 0; return false; }
        p = (const char*)q;
        madvise(q, n, MADV_SEQUENTIAL);
        return true;
    }
};

// Text counts (cu:1145-1159), parsed by T threads: each thread takes the tokens
// that start in its byte range (ranges end at whitespace, so no token is cut),
// counts them, and after a prefix sum over the threads parses them into place.
// Errors are those of a serial reader: the first bad token among the first
// N*G, else a short file, else the first negative count.
static busseq_status read_counts_text(const char* path, const char* buf, size_t n, long want, int G, int32_t* Y)
{
    int T = 1;
#ifdef _OPENMP
    T = omp_get_max_threads();
#endif
    const size_t per = (size_t)4 << 20;   // at least 4 MB per thread
    if ((size_t)T > n / per + 1) T = (int)(n / per + 1);
    if (T < 1) T = 1;
    std::vector<size_t> lo(T + 1, n);
    lo[0] = 0;
    for (int t = 1; t < T; ++t) {
        size_t q = n / T * t;
        while (q < n && !is_ws(buf[q])) ++q;
        lo[t] = q;
    }
    for (int t = 1; t <= T; ++t) if (lo[t] < lo[t - 1]) lo[t] = lo[t - 1];
    std::vector<long> cnt(T + 1, 0);
    #pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; ++t) {
        long c = 0;
        char prev = lo[t] == 0 ? ' ' : buf[lo[t] - 1];
        for (size_t q = lo[t]; q < lo[t + 1]; ++q) {
            const char ch = buf[q];
            c += is_ws(prev) && !is_ws(ch);
            prev = ch;
        }
        cnt[t + 1] = c;
    }
    for (int t = 0; t < T; ++t) cnt[t + 1] += cnt[t];
    std::vector<long> bad_at(T, -1), neg_at(T, -1);
    std::vector<std::string> bad_tok(T);
    #pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; ++t) {
        long j = cnt[t];
        const char* p = buf + lo[t];
        const char* end = buf + lo[t + 1];
        while (j < want) {
            while (p < end && is_ws(*p)) ++p;
            if (p >= end) break;
            const char* tok = p;
            bool neg = false;
            if (*p == '+' || *p == '-') { neg = *p == '-'; ++p; }
            int64_t v = 0;
            const char* d0 = p;
            while (p < end && *p >= '0' && *p <= '9') { v = v * 10 + (*p - '0'); if (v > 2147483647LL) break; ++p; }
            if (p == d0 || !(p >= end || is_ws(*p)) || v > 2147483647LL) {
                const char* q = tok;
                while (q < end && q - tok < 24 && !is_ws(*q)) ++q;
                bad_at[t] = j;
                bad_tok[t].assign(tok, q);
                break;
            }
            const int32_t y = (int32_t)(neg ? -v : v);
            if (y < 0 && neg_at[t] < 0) neg_at[t] = j;
            Y[j++] = y;
        }
    }
    for (int t = 0; t < T; ++t)
        if (bad_at[t] >= 0) {
            bq_set_error(std::string(path) + ": not an integer count: '" + bad_tok[t] + "'");
            return BUSSEQ_ERR_FORMAT;
        }
    if (cnt[T] < want) {
        bq_set_error(std::string(path) + ": expected " + std::to_string(want) + " counts (N x G), found " + std::to_string(cnt[T]));
        return BUSSEQ_ERR_FORMAT;
    }
    for (int t = 0; t < T; ++t)
        if (neg_at[t] >= 0) {
            const long j = neg_at[t];
            bq_set_error(std::string(path) + ": negative count at cell " + std::to_string(j / G) + ", gene " + std::to_string(j % G));
            return BUSSEQ_ERR_FORMAT;
        }
    return BUSSEQ_OK;
}

extern "C" busseq_status busseq_read_counts(const char* path, int N, int G, int32_t* Y)
{
    if (!path || !Y || N < 0 || G < 0) { bq_set_error("bad arguments to busseq_read_counts"); return BUSSEQ_ERR_ARG; }
    MappedFile f;
    if (!f.open_ro(path)) return BUSSEQ_ERR_IO;
    const long want = (long)N * G;
    if (f.n >= sizeof(BinHeader) && std::memcmp(f.p, kBinMagic, sizeof(kBinMagic)) == 0) {
        BinHeader h;
        std::memcpy(&h, f.p, sizeof(h));
        if (h.N != N || h.G != G) {
            bq_set_error(std::string(path) + ": binary count file holds " + std::to_string(h.N) + " x " +
                         std::to_string(h.G) + " counts, expected " + std::to_string(N) + " x " + std::to_string(G));
            return BUSSEQ_ERR_FORMAT;
        }
        if (f.n != sizeof(BinHeader) + (size_t)want * sizeof(int32_t)) {
            bq_set_error(std::string(path) + ": binary count file has the wrong size");
            return BUSSEQ_ERR_FORMAT;
        }
        const int32_t* src = (const int32_t*)(f.p + sizeof(BinHeader));
        long first_neg = LONG_MAX;
        #pragma omp parallel for schedule(static) reduction(min : first_neg)
        for (long r = 0; r < N; ++r) {
            std::memcpy(Y + r * G, src + r * G, (size_t)G * sizeof(int32_t));
            for (long g = 0; g < G; ++g)
                if (src[r * G + g] < 0) { if (r * G + g < first_neg) first_neg = r * G + g; break; }
        }
        if (first_neg != LONG_MAX) {
            bq_set_error(std::string(path) + ": negative count at cell " + std::to_string(first_neg / G) + ", gene " + std::to_string(first_neg % G));
            return BUSSEQ_ERR_FORMAT;
        }
        return BUSSEQ_OK;
    }
    return read_counts_text(path, f.p ? f.p : "", f.n, want, G, Y);
}

extern "C" busseq_status busseq_write_counts_binary(const char* path, int N, int G, const int32_t* Y)
{
    if (!path || !Y || N < 0 || G < 0) { bq_set_error("bad arguments to busseq_write_counts_binary"); return BUSSEQ_ERR_ARG; }
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) { bq_set_error(std::string("cannot create ") + tmp + ": " + std::strerror(errno)); return BUSSEQ_ERR_IO; }
    BinHeader h;
    std::memcpy(h.magic, kBinMagic, sizeof(kBinMagic));
    h.N = N; h.G = G; h.reserved = 0;
    const size_t n = (size_t)N * G;
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && (n == 0 || std::fwrite(Y, sizeof(int32_t), n, f) == n);
    ok = std::fclose(f) == 0 && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        bq_set_error(std::string("cannot write ") + path + ": " + std::strerror(errno));
        return BUSSEQ_ERR_IO;
    }
    return BUSSEQ_OK;
}

// ---------------------------------------------------------------------------
// synthetic data from the BUSseq generative model
// ---------------------------------------------------------------------------
extern "C" busseq_status busseq_simulate(int B, const int* n_b, int G, int K, uint32_t ds,
                                         double g0_lo, double g0_hi, int32_t* Y, int32_t* W_true,
                                         int32_t* X_true)
{
    if (B < 1 || G < 1 || K < 1 || !n_b) { bq_set_error("simulate: bad dimensions"); return BUSSEQ_ERR_ARG; }
    const uint32_t key = 0x5EEDu;
    std::vector<int> boff(B + 1, 0);
    for (int b = 0; b < B; ++b) boff[b + 1] = boff[b] + n_b[b];
    const int N = boff[B];
    std::vector<double> alpha(G), beta((size_t)K * G, 0.0), nu((size_t)B * G, 0.0), phi((size_t)B * G);
    std::vector<double> g0(B), g1(B), pi((size_t)B * K);
    for (int g = 0; g < G; ++g) {
        bq_rng r; bq_rng_init(&r, ds, key, (uint32_t)g, 0u, 0u, BQ_PH_SIM_PARAM);
        alpha[g] = 1.5 + 0.5 * bq_normal(&r);
    }
    int m = (int)(0.1 * G + 0.5);
    std::vector<int> idx(G);
    for (int k = 1; k < K; ++k) {
        bq_rng r; bq_rng_init(&r, ds, key, (uint32_t)k, 1u, 0u, BQ_PH_SIM_PARAM);
        for (int g = 0; g < G; ++g) idx[g] = g;
        for (int j = 0; j < m; ++j) {
            int s = j + (int)(bq_uniform(&r) * (G - j));
            if (s >= G) s = G - 1;
            int t = idx[j]; idx[j] = idx[s]; idx[s] = t;
            double sign = bq_uniform(&r) < 0.5 ? -1.0 : 1.0;
            beta[(size_t)k * G + idx[j]] = 2.0 * sign + 0.5 * bq_normal(&r);
        }
    }
    for (int b = 0; b < B; ++b) {
        for (int g = 0; g < G; ++g) {
            bq_rng r; bq_rng_init(&r, ds, key, (uint32_t)g, 16u + (uint32_t)b, 0u, BQ_PH_SIM_PARAM);
            nu[(size_t)b * G + g] = b == 0 ? 0.0 : 0.5 * bq_normal(&r);
            phi[(size_t)b * G + g] = bq_rgamma(&r, 5.0, 2.0);
        }
        bq_rng r; bq_rng_init(&r, ds, key, (uint32_t)b, 400u, 0u, BQ_PH_SIM_PARAM);
        g0[b] = g0_lo + (g0_hi - g0_lo) * bq_uniform(&r);
        g1[b] = -0.3 + 0.2 * bq_uniform(&r);
        double tot = 0.0;
        for (int k = 0; k < K; ++k) { pi[(size_t)b * K + k] = bq_rgamma(&r, 1.0, 1.0); tot += pi[(size_t)b * K + k]; }
        for (int k = 0; k < K; ++k) pi[(size_t)b * K + k] /= tot;
    }
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < N; ++i) {
        int b = 0;
        while (i >= boff[b + 1]) ++b;
        bq_rng rc; bq_rng_init(&rc, ds, key, (uint32_t)i, 0u, 0u, BQ_PH_SIM_CELL);
        double delta = 0.2 * bq_normal(&rc);
        if (i == boff[b]) delta = 0.0;
        double u = bq_uniform(&rc), cum = 0.0;
        int w = K - 1;
        for (int k = 0; k < K; ++k) { cum += pi[(size_t)b * K + k]; if (u <= cum) { w = k; break; } }
        if (W_true) W_true[i] = w;
        for (int g = 0; g < G; ++g) {
            bq_rng r; bq_rng_init(&r, ds, key, (uint32_t)i, (uint32_t)g, 0u, BQ_PH_SIM_COUNT);
            double mu = bq_exp(alpha[g] + beta[(size_t)w * G + g] + nu[(size_t)b * G + g] + delta);
            int x = bq_rnbinom(&r, mu, phi[(size_t)b * G + g]);
            double pz = 1.0 / (1.0 + bq_exp(-(g0[b] + g1[b] * x)));
            int z = bq_uniform(&r) < pz;
            if (Y) Y[(size_t)i * G + g] = z ? 0 : x;
            if (X_true) X_true[(size_t)i * G + g] = x;
        }
    }
    return BUSSEQ_OK;
}

// ---------------------------------------------------------------------------
// output writers, byte-compatible with BUSseq_gpu.cu:1610-2018
// ---------------------------------------------------------------------------
namespace {

// "%lf" of v: six decimals of the exact binary value rounded half to even,
// as glibc's printf does, from 128-bit integer arithmetic (|v| < 9e12; other
// values go through snprintf).  Returns the length written to t (>= 400 bytes).
int fmt_f6(double v, char* t)
{
    if (!(std::fabs(v) < 9.0e12)) return std::snprintf(t, 400, "%lf", v);
    uint64_t bits;
    std::memcpy(&bits, &v, sizeof(bits));
    const bool neg = bits >> 63;
    const int ex = (int)((bits >> 52) & 0x7ff);
    uint64_t m = bits & ((1ull << 52) - 1);
    int sh;   // |v| = m 2^-sh, sh >= 1 since |v| < 2^52
    if (ex == 0) sh = 1074; else { m |= 1ull << 52; sh = 1075 - ex; }
    uint64_t q = 0;   // round(|v| 1e6)
    if (sh < 100) {
        const unsigned __int128 num = (unsigned __int128)m * 1000000u;   // < 2^73
        const unsigned __int128 one = (unsigned __int128)1;
        const unsigned __int128 rem = num & ((one << sh) - 1), half = one << (sh - 1);
        q = (uint64_t)(num >> sh);
        if (rem > half || (rem == half && (q & 1))) ++q;
    }
    char d[32];
    int n = 0, k = 0;
    uint64_t ip = q / 1000000u;
    uint32_t fp = (uint32_t)(q % 1000000u);
    if (neg) t[n++] = '-';
    do { d[k++] = (char)('0' + ip % 10); ip /= 10; } while (ip);
    while (k) t[n++] = d[--k];
    t[n++] = '.';
    for (int j = 5; j >= 0; --j) { t[n + j] = (char)('0' + fp % 10); fp /= 10; }
    return n + 6;
}

int fmt_i32(int32_t v, char* t)
{
    int n = 0, m = 0;
    uint32_t u = v < 0 ? (uint32_t)(-(int64_t)v) : (uint32_t)v;
    char d[12];
    do { d[m++] = (char)('0' + u % 10); u /= 10; } while (u);
    if (v < 0) t[n++] = '-';
    while (m) t[n++] = d[--m];
    return n;
}

// rows x cols matrix, tab-separated, newline per row (after an all-zero first
// row, cu:1705-1707, 1740-1742, when zero_row).  Blocks of rows are formatted
// by all threads and written in order.
template <typename V>
bool write_matrix(const std::string& path, const V* a, long rows, long cols, bool zero_row = false)
{
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) { bq_set_error("cannot write " + path + ": " + std::strerror(errno)); return false; }
    bool ok = true;
    if (zero_row)
        for (long c = 0; c < cols && ok; ++c) ok = std::fwrite(c == cols - 1 ? "0.000000\n" : "0.000000\t", 1, 9, f) == 9;
    int T = 1;
#ifdef _OPENMP
    T = omp_get_max_threads();
#endif
    const long per_row = cols * (sizeof(V) == 4 ? 12 : 24) + 1;
    const long rb = std::max(1L, (1L << 20) / per_row);   // rows per block, ~1 MB of text
    std::vector<std::vector<char>> buf(T);
    for (long r0 = 0; r0 < rows && ok; r0 += rb * T) {
        #pragma omp parallel for num_threads(T) schedule(static, 1)
        for (int t = 0; t < T; ++t) {
            std::vector<char>& b = buf[t];
            b.clear();
            const long lo = r0 + t * rb, hi = std::min(rows, lo + rb);
            char tmp[400];
            for (long r = lo; r < hi; ++r)
                for (long c = 0; c < cols; ++c) {
                    int n;
                    if constexpr (sizeof(V) == 4) n = fmt_i32((int32_t)a[r * cols + c], tmp);
                    else n = fmt_f6((double)a[r * cols + c], tmp);
                    tmp[n++] = c == cols - 1 ? '\n' : '\t';
                    b.insert(b.end(), tmp, tmp + n);
                }
        }
        for (int t = 0; t < T && ok; ++t)
            if (!buf[t].empty()) ok = std::fwrite(buf[t].data(), 1, buf[t].size(), f) == buf[t].size();
    }
    if (std::fclose(f) != 0) ok = false;
    if (!ok) bq_set_error("cannot write " + path + ": " + std::strerror(errno));
    return ok;
}

// values on one line, each followed by sep (cu:1725-1727, 1838-1840)
bool write_line(const std::string& path, const double* v, int n, char sep)
{
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) { bq_set_error("cannot write " + path + ": " + std::strerror(errno)); return false; }
    std::string line;
    char t[400];
    for (int j = 0; j < n; ++j) {
        int m = fmt_f6(v[j], t);
        t[m++] = sep;
        line.append(t, (size_t)m);
    }
    bool ok = std::fwrite(line.data(), 1, line.size(), f) == line.size();
    if (std::fclose(f) != 0) ok = false;
    if (!ok) bq_set_error("cannot write " + path + ": " + std::strerror(errno));
    return ok;
}
} // namespace

extern "C" busseq_status busseq_write_outputs(const busseq_result* r, const char* prefix, int print_preserved)
{
    if (!r || !prefix) { bq_set_error("write_outputs: null argument"); return BUSSEQ_ERR_ARG; }
    const std::string p(prefix);
    const long B = r->B, N = r->N, G = r->G, K = r->K, P = r->n_pres;
    bool ok = true;
    ok = ok && write_matrix(p + "_post_Z.txt", r->Z, N, G);
    ok = ok && write_matrix(p + "_post_X.txt", r->X, N, G);
    ok = ok && write_matrix(p + "_post_gamma.txt", r->gamma, 2, B);
    ok = ok && write_matrix(p + "_post_alpha.txt", r->alpha, G, 1);
    ok = ok && write_matrix(p + "_post_L.txt", r->L + G, K - 1, G);
    ok = ok && write_matrix(p + "_post_beta.txt", r->beta + G, K - 1, G, true);
    ok = ok && write_line(p + "_post_p.txt", r->ptau, 2, '\t');
    ok = ok && write_matrix(p + "_post_nu.txt", r->nu + G, B - 1, G, true);
    ok = ok && write_matrix(p + "_post_delta.txt", r->delta, N, 1);
    ok = ok && write_matrix(p + "_post_phi.txt", r->phi, B, G);
    ok = ok && write_matrix(p + "_post_W.txt", r->W, N, 1);
    ok = ok && write_matrix(p + "_post_pi.txt", r->pi, B, K);
    ok = ok && write_line(p + "_BIC.txt", &r->bic, 1, '\n');
    if (ok && print_preserved) {
        if (!r->tr_alpha) { bq_set_error("write_outputs: traces were not kept (keep_traces = 0)"); return BUSSEQ_ERR_ARG; }
        ok = ok && write_matrix(p + "_preserved_gamma.txt", r->tr_gamma, 2 * B, P);
        ok = ok && write_matrix(p + "_preserved_alpha.txt", r->tr_alpha, G, P);
        ok = ok && write_matrix(p + "_preserved_L.txt", r->tr_L, (K - 1) * G, P);
        ok = ok && write_matrix(p + "_preserved_beta.txt", r->tr_beta, G * (K - 1), P);
        ok = ok && write_matrix(p + "_preserved_p.txt", r->tr_ptau, 2, P);
        ok = ok && write_matrix(p + "_preserved_nu.txt", r->tr_nu, (B - 1) * G, P);
        ok = ok && write_matrix(p + "_preserved_delta.txt", r->tr_delta, N, P);
        ok = ok && write_matrix(p + "_preserved_phi.txt", r->tr_phi, B * G, P);
        ok = ok && write_matrix(p + "_preserved_W.txt", r->tr_W, N, P);
        ok = ok && write_matrix(p + "_preserved_pi.txt", r->tr_pi, B * K, P);
    }
    return ok ? BUSSEQ_OK : BUSSEQ_ERR_IO;
}

extern "C" void busseq_result_free(busseq_result* r)
{
    if (!r) return;
    void* ps[] = { r->n_b, r->X, r->Z, r->W, r->gamma, r->alpha, r->L, r->beta, r->ptau, r->nu, r->delta,
                   r->phi, r->pi, r->tr_gamma, r->tr_alpha, r->tr_L, r->tr_beta, r->tr_ptau, r->tr_nu,
                   r->tr_delta, r->tr_phi, r->tr_W, r->tr_pi };
    for (void* q : ps) std::free(q);
    std::memset(r, 0, sizeof(*r));
}

// ---------------------------------------------------------------------------
// streamed traces: one block of rows (iteration-major, as the device writes
// them) into the parameter-major host traces ([param * P + iter], cu:892-936);
// 256 parameters per task so that each task writes its rows' columns together
// ---------------------------------------------------------------------------
void bq_trace_transpose(const void* stage, int esz, size_t np, int rows, const int* col, size_t P, void* dst)
{
    const long ntile = (long)((np + 255) / 256);
    #pragma omp parallel for schedule(static) if (np * (size_t)rows > 65536)
    for (long tj = 0; tj < ntile; ++tj) {
        const size_t j0 = (size_t)tj * 256, j1 = j0 + 256 < np ? j0 + 256 : np;
        for (int q = 0; q < rows; ++q) {
            if (col[q] < 0) continue;
            if (esz == 8) {
                const double* sr = (const double*)stage + (size_t)q * np;
                double* d = (double*)dst + col[q];
                for (size_t j = j0; j < j1; ++j) d[j * P] = sr[j];
            } else {
                const int32_t* sr = (const int32_t*)stage + (size_t)q * np;
                int32_t* d = (int32_t*)dst + col[q];
                for (size_t j = j0; j < j1; ++j) d[j * P] = sr[j];
            }
        }
    }
}
