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
#include <string>
#include <vector>

#include "../../include/busseq.h"

#define BATCH_SLOTS 128   // the largest B (the reference's __constant__ batch arrays, cu:9-10)

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

// What one run does: the reference's defaults (BUSseq_gpu.cu:1028-1038), what
// the command line changed, and the additive long options.
struct Settings {
    int batches = 4;
    std::vector<int> cells = std::vector<int>(BATCH_SLOTS, 0);   // per batch; the first four default below
    int genes = 3000, subtypes = 5, seed = 0;
    int iterations = 4000, burn_in = 2000, unchanged = 500;
    bool preserved = false;
    std::string counts = "count_data/demo_count.txt", prefix = "demo_output", dims, cache;
    // which of the options with derived defaults the command line set
    bool dims_set = false, burn_in_set = false, unchanged_set = false, seed_set = false;
    int precision = 32, tile_rows = 0, device = 0;
    bool tune = false;
    Settings() { cells[0] = 300; cells[1] = 300; cells[2] = 200; cells[3] = 200; }
};

static void usage(const char* prog)
{
    std::printf("Error Usage: %s [-b batch] [-n sample_size_file] [-g gene] [-k celltype] [-c count_data_file)] "
                "[-i iterations] [-b burnin] [-u unchanged_iterations] [-p print_preserved] [-o output_prefix] \n",
                prog);
    std::exit(1);
}

// the reference's short options (cu:1048) and the long ones
static void parse(int argc, char** argv, Settings& o)
{
    enum { OPT_PRECISION = 1000, OPT_TILE_ROWS, OPT_DEVICE, OPT_CACHE };
    static const struct option longopts[] = {
        {"precision", required_argument, nullptr, OPT_PRECISION},
        {"tile-rows", required_argument, nullptr, OPT_TILE_ROWS},
        {"device", required_argument, nullptr, OPT_DEVICE},
        {"count-cache", required_argument, nullptr, OPT_CACHE},
        {nullptr, 0, nullptr, 0},
    };
    for (int opt; (opt = getopt_long(argc, argv, "B:N:G:K:s:c:i:b:u:po:", longopts, nullptr)) != -1;) {
        const char* v = optarg;
        switch (opt) {
        case 'B': o.batches = std::atoi(v); break;
        case 'N': o.dims = v; o.dims_set = true; break;
        case 'G': o.genes = std::atoi(v); break;
        case 'K': o.subtypes = std::atoi(v); break;
        case 's': o.seed = std::atoi(v); o.seed_set = true; break;
        case 'c': o.counts = v; break;
        case 'i': o.iterations = std::atoi(v); break;
        case 'b': o.burn_in = std::atoi(v); o.burn_in_set = true; break;
        case 'u': o.unchanged = std::atoi(v); o.unchanged_set = true; break;
        case 'p': o.preserved = true; break;
        case 'o': o.prefix = v; break;
        case OPT_PRECISION: o.precision = std::atoi(v); break;
        case OPT_TILE_ROWS:
            o.tune = std::strcmp(v, "auto") == 0;
            o.tile_rows = o.tune ? 0 : std::atoi(v);
            break;
        case OPT_DEVICE: o.device = std::atoi(v); break;
        case OPT_CACHE: o.cache = v; break;
        default: usage(argv[0]);
        }
    }
}

// The values derived from others when the command line holds any argument
// (cu:1092-1111): the batch sizes from the -N file, burn-in n_iter/2 and the
// unchanged iterations max(0.3 n_iter, 500) unless given; then the seed from
// the clock unless given.  Returns 0, or an exit code.
static int complete(Settings& o, bool any_args)
{
    if (any_args) {
        if (o.dims_set && o.batches > 0 && busseq_read_dims(o.dims.c_str(), o.batches, o.cells.data()) != BUSSEQ_OK)
            return fail("reading the -N file");
        if (!o.burn_in_set) o.burn_in = o.iterations / 2;
        if (!o.unchanged_set) o.unchanged = 0.3 * o.iterations > 500 ? (int)(0.3 * o.iterations) : 500;
    }
    if (!o.seed_set) {
        struct timeval tv;
        gettimeofday(&tv, nullptr);
        o.seed = (int)tv.tv_usec;
    }
    return 0;
}

// The reference's input checks and messages (cu:1120-1134), with exit status
// -1; a burn-in as long as the run is rejected too (the reference would divide
// the posterior sums by zero recorded iterations).
static bool valid(const Settings& o)
{
    const char* why = nullptr;
    if (o.burn_in >= o.iterations) why = "Burn-in iterations must be less than total number of iterations.";
    else if (o.batches < 2) why = "The batch number must be greater than one.";
    else
        for (int b = 0; b < o.batches && !why; ++b)
            if (o.cells[b] < o.subtypes) why = "The sample size in any batch must be greater than the assumed cell type number.";
    if (why) std::printf("%s\n", why);
    return why == nullptr;
}

static void stage(const char* line)
{
    std::printf("%s\n", line);
    std::fflush(stdout);
}

int main(int argc, char** argv)
{
    Settings o;
    parse(argc, argv, o);
    // the reference's limit (its __constant__ batch arrays, cu:9-10); fewer than two
    // batches meet its own check below
    if (o.batches > BATCH_SLOTS) { std::printf("Error: -B must be at most 128.\n"); return 1; }
    if (int rc = complete(o, argc > 1)) return rc;
    std::printf("Seed is %d\n", o.seed);
    if (!valid(o)) return -1;
    const int B = o.batches, G = o.genes;
    int N = 0;
    for (int b = 0; b < B; ++b) N += o.cells[b];

    stage("Start reading file.");
    std::vector<int32_t> Y((size_t)N * G);
    if (read_counts(o.counts.c_str(), o.cache.c_str(), N, G, Y.data()) != BUSSEQ_OK)
        return fail("reading the count file");

    stage("Start initialization.");
    busseq_config cfg;
    busseq_config_default(&cfg);
    cfg.B = B; cfg.n_b = o.cells.data(); cfg.G = G; cfg.K = o.subtypes; cfg.seed = (uint32_t)o.seed;
    cfg.n_iter = o.iterations; cfg.n_burnin = o.burn_in; cfg.n_unchanged = o.unchanged;
    cfg.keep_traces = o.preserved; cfg.precision = o.precision; cfg.tile_rows = o.tile_rows; cfg.device = o.device;
    if (o.tune) {
        double sps = 0.0;
        if (busseq_autotune(&cfg, Y.data(), 20, &cfg.tile_rows, &sps) != BUSSEQ_OK) return fail("tile autotuning");
        std::printf("Tile rows %d (autotuned, %.1f sweeps/s).\n", cfg.tile_rows, sps);
        std::fflush(stdout);
    }
    busseq_sampler* s = nullptr;
    if (busseq_create(&cfg, Y.data(), &s) != BUSSEQ_OK) return fail("initialization");

    // burn-in, then the recorded sweeps, then the posterior summaries
    stage("Start MCMC.");
    busseq_result r;
    std::memset(&r, 0, sizeof(r));
    const char* failed = nullptr;
    if (busseq_sweeps(s, o.burn_in) != BUSSEQ_OK) failed = "MCMC";
    if (!failed) {
        stage("Burnin completed. Start recording iterations.");
        if (busseq_sweeps(s, o.iterations - o.burn_in) != BUSSEQ_OK) failed = "MCMC";
    }
    if (!failed) {
        stage("Start posterior inference.");
        if (busseq_finish(s, 0, &r) != BUSSEQ_OK) failed = "posterior inference";
    }
    busseq_destroy(s);
    if (failed) return fail(failed);

    stage("Start writing output files.");
    if (o.preserved) std::printf("Start writing preserved iterations.\n");
    const busseq_status rc = busseq_write_outputs(&r, o.prefix.c_str(), o.preserved ? 1 : 0);
    busseq_result_free(&r);
    if (rc != BUSSEQ_OK) return fail("writing outputs");
    return 0;
}
