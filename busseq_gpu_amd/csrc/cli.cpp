// cli.cpp -- the BUSseq_gpu executable: the reference's command line
// (BUSseq_gpu.cu:1026-1134), progress lines and outputs, over libbusseq.
//
//   BUSseq_gpu -B 4 -N dim.txt -G 3000 -K 5 -s 13579 -c count.txt
//              -i 4000 -b 2000 -u 500 [-p] -o prefix
//
// Additive long options (defaults reproduce the reference's behaviour):
//   --precision=32|64   per-element arithmetic (default 32)
//   --tile-rows=R|auto  cells per reduction tile (default 16, 32 or 64 by the number
//                       of cells, busseq_default_tile_rows; auto: the fastest of
//                       16, 32, 64, 128 timed on this data, printed)
//   --device=D          HIP device ordinal (default 0)
//   --count-cache=PATH  binary copy of the -c counts: read instead of the text when
//                       it is at least as new as the -c file and holds N x G counts,
//                       otherwise written after the text is read
// The -c file itself may also be a binary count file (busseq_write_counts_binary).
#include <getopt.h>
#include <sys/stat.h>
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/busseq.h"

static int fail(const char* what)
{
    std::printf("Error: %s: %s\n", what, busseq_last_error());
    return 1;
}

// modification time of a file in nanoseconds, or -1 if it cannot be read
static long long mtime_ns(const char* path)
{
    struct stat st;
    if (stat(path, &st) != 0) return -1;
    return (long long)st.st_mtim.tv_sec * 1000000000LL + st.st_mtim.tv_nsec;
}

// counts from the -c file, or from its binary cache when that is current
static busseq_status read_counts(const char* count_data, const char* cache, int N, int G, int32_t* Y)
{
    if (cache[0]) {
        const long long tc = mtime_ns(cache), tt = mtime_ns(count_data);
        if (tc >= 0 && tt >= 0 && tc >= tt && busseq_read_counts(cache, N, G, Y) == BUSSEQ_OK) {
            std::printf("Counts read from the cache %s.\n", cache);
            return BUSSEQ_OK;
        }
    }
    busseq_status rc = busseq_read_counts(count_data, N, G, Y);
    if (rc != BUSSEQ_OK || !cache[0]) return rc;
    if (busseq_write_counts_binary(cache, N, G, Y) == BUSSEQ_OK)
        std::printf("Counts cached in %s.\n", cache);
    else
        std::printf("Warning: the count cache was not written: %s\n", busseq_last_error());
    return BUSSEQ_OK;
}

int main(int argc, char** argv)
{
    // hard-coded defaults, BUSseq_gpu.cu:1028-1038
    int B = 4;
    std::vector<int> n_b(200, 0);
    n_b[0] = 300; n_b[1] = 300; n_b[2] = 200; n_b[3] = 200;
    int G = 3000, K = 5, seed = 0;
    char count_data[4096] = "count_data/demo_count.txt";
    int n_iter = 4000, n_burnin = 2000, n_unchanged = 500, print_preserved = 0;
    char output_file[4096] = "demo_output";
    char n_file[4096] = "";
    char cache[4096] = "";
    int n_file_flag = 0, burnin_flag = 1, unchanged_flag = 1, seed_flag = 1;
    int precision = 32, tile_rows = 0, device = 0, tune = 0;

    static struct option longopts[] = {
        {"precision", required_argument, nullptr, 1000},
        {"tile-rows", required_argument, nullptr, 1001},
        {"device", required_argument, nullptr, 1002},
        {"count-cache", required_argument, nullptr, 1003},
        {nullptr, 0, nullptr, 0},
    };
    int opt;
    while ((opt = getopt_long(argc, argv, "B:N:G:K:s:c:i:b:u:po:", longopts, nullptr)) != -1) {
        switch (opt) {
        case 'B': B = std::atoi(optarg); break;
        case 'N': n_file_flag = 1; std::snprintf(n_file, sizeof(n_file), "%s", optarg); break;
        case 'G': G = std::atoi(optarg); break;
        case 'K': K = std::atoi(optarg); break;
        case 's': seed_flag = 0; seed = std::atoi(optarg); break;
        case 'c': std::snprintf(count_data, sizeof(count_data), "%s", optarg); break;
        case 'i': n_iter = std::atoi(optarg); break;
        case 'b': burnin_flag = 0; n_burnin = std::atoi(optarg); break;
        case 'u': unchanged_flag = 0; n_unchanged = std::atoi(optarg); break;
        case 'p': print_preserved = 1; break;
        case 'o': std::snprintf(output_file, sizeof(output_file), "%s", optarg); break;
        case 1000: precision = std::atoi(optarg); break;
        case 1001: tune = std::strcmp(optarg, "auto") == 0; tile_rows = tune ? 0 : std::atoi(optarg); break;
        case 1002: device = std::atoi(optarg); break;
        case 1003: std::snprintf(cache, sizeof(cache), "%s", optarg); break;
        default:
            std::printf("Error Usage: %s [-b batch] [-n sample_size_file] [-g gene] [-k celltype] [-c count_data_file)] "
                        "[-i iterations] [-b burnin] [-u unchanged_iterations] [-p print_preserved] [-o output_prefix] \n",
                        argv[0]);
            std::exit(1);
        }
    }
    // the reference's limit (its __constant__ batch arrays, cu:9-10); B < 2 meets its own check below
    if (B > 128) { std::printf("Error: -B must be at most 128.\n"); return 1; }
    // derived defaults apply only when arguments are given (cu:1092-1111)
    if (argc > 1) {
        if (n_file_flag && B > 0) {
            if (busseq_read_dims(n_file, B, n_b.data()) != BUSSEQ_OK) return fail("reading the -N file");
        }
        if (burnin_flag) n_burnin = n_iter / 2;
        if (unchanged_flag) n_unchanged = 0.3 * n_iter > 500 ? (int)(0.3 * n_iter) : 500;
    }
    if (seed_flag) {
        struct timeval tv;
        gettimeofday(&tv, nullptr);
        seed = (int)tv.tv_usec;
    }
    std::printf("Seed is %d\n", seed);
    // input checks, cu:1120-1134; a burn-in as long as the run is rejected too (the
    // reference would divide the posterior sums by zero recorded iterations)
    if (n_burnin >= n_iter) {
        std::printf("Burn-in iterations must be less than total number of iterations.\n");
        return -1;
    }
    if (B < 2) {
        std::printf("The batch number must be greater than one.\n");
        return -1;
    }
    for (int b = 0; b < B; ++b)
        if (n_b[b] < K) {
            std::printf("The sample size in any batch must be greater than the assumed cell type number.\n");
            return -1;
        }
    int N = 0;
    for (int b = 0; b < B; ++b) N += n_b[b];

    std::printf("Start reading file.\n");
    std::fflush(stdout);
    std::vector<int32_t> Y((size_t)N * G);
    if (read_counts(count_data, cache, N, G, Y.data()) != BUSSEQ_OK) return fail("reading the count file");

    std::printf("Start initialization.\n");
    std::fflush(stdout);
    busseq_config cfg;
    busseq_config_default(&cfg);
    cfg.B = B; cfg.n_b = n_b.data(); cfg.G = G; cfg.K = K;
    cfg.seed = (uint32_t)seed; cfg.n_iter = n_iter; cfg.n_burnin = n_burnin; cfg.n_unchanged = n_unchanged;
    cfg.keep_traces = print_preserved; cfg.precision = precision; cfg.tile_rows = tile_rows; cfg.device = device;
    if (tune) {
        double sps = 0.0;
        if (busseq_autotune(&cfg, Y.data(), 20, &cfg.tile_rows, &sps) != BUSSEQ_OK) return fail("tile autotuning");
        std::printf("Tile rows %d (autotuned, %.1f sweeps/s).\n", cfg.tile_rows, sps);
        std::fflush(stdout);
    }
    busseq_sampler* s = nullptr;
    if (busseq_create(&cfg, Y.data(), &s) != BUSSEQ_OK) return fail("initialization");

    std::printf("Start MCMC.\n");
    std::fflush(stdout);
    if (busseq_sweeps(s, n_burnin) != BUSSEQ_OK) { busseq_destroy(s); return fail("MCMC"); }
    std::printf("Burnin completed. Start recording iterations.\n");
    std::fflush(stdout);
    if (busseq_sweeps(s, n_iter - n_burnin) != BUSSEQ_OK) { busseq_destroy(s); return fail("MCMC"); }

    std::printf("Start posterior inference.\n");
    std::fflush(stdout);
    busseq_result r;
    std::memset(&r, 0, sizeof(r));
    if (busseq_finish(s, 0, &r) != BUSSEQ_OK) { busseq_destroy(s); return fail("posterior inference"); }
    busseq_destroy(s);

    std::printf("Start writing output files.\n");
    std::fflush(stdout);
    if (print_preserved) std::printf("Start writing preserved iterations.\n");
    busseq_status rc = busseq_write_outputs(&r, output_file, print_preserved);
    busseq_result_free(&r);
    if (rc != BUSSEQ_OK) return fail("writing outputs");
    return 0;
}
