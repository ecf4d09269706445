// vc_main.cpp -- command-line driver with the reference's surface
// (./output, src/main.cu:43-408 and src/config.h:14-243): same flags, same
// key stdout lines, same Results/Results.csv row (src/auxFunctions.cpp:580-614).
// The solve itself goes through the C ABI of libvcgpu.so.
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <vector>

#include "vcgpu.h"

namespace {

enum Version { STACK_ONLY, HYBRID, SEQUENTIAL };
enum Instance { MVC, PVC };

struct Config {
    Version version = HYBRID;
    Instance instance = MVC;
    unsigned k = 0;
    unsigned numBlocks = 0;
    bool userDefMemory = false;
    bool useGlobalMemory = false;
    unsigned blockDim = 0;
    unsigned globalListSize = 1u << 19;
    float globalListThreshold = 1.0f;
    unsigned startingDepth = 4;
    const char *outputFilePrefix = "1";
    const char *graphFileName = "GraphInput.txt";
    bool userListSize = false;
    bool userThreshold = false;
    unsigned indexBase = 1;
};

void usage()
{
    std::fprintf(stderr,
                 "\nArguments :\n\n"
                 "   -v <Version>                  hybrid, stackonly, sequential (default hybrid)\n"
                 "   -f <Graph File Name>          graph file, 'n m' header then m 1-indexed edges (default GraphInput.txt)\n"
                 "   -o <Output File Prefix>       identifier written to Results/Results.csv (default 1)\n"
                 "   -i <Instance>                 mvc or pvc (default mvc)\n"
                 "   -k <parameter>                k for pvc (default 0)\n"
                 "   -q <Global Work List Size>    worklist slots, power of 2 (default: sized to the grid)\n"
                 "   -t <Global Work List Threshold> donate while queued < t * size, 0.0-1.0 (default: only while workers wait)\n"
                 "   -d <Starting Depth>           static split depth for stackonly (default 4)\n"
                 "   -b <Block Dimension>          threads per workgroup, multiple of 64, <= 256 (default 256)\n"
                 "   -g <Use Global Memory>        0 = node storage in LDS, 1 = in HBM (default: LDS when it fits)\n"
                 "   -n <Number of Blocks>         workgroups (default occupancy x CUs)\n"
                 "   -z                            the graph file's vertices start at 0 (formated_*, lastfm-asia, *-cell)\n\n"
                 "   VC_SEED=<u32> fixes the RemoveEdge seed (default time(0), as the reference).\n\n");
}

const char *as_string(Version v) { return v == STACK_ONLY ? "stackonly" : v == HYBRID ? "hybrid" : "sequential"; }
const char *as_string(Instance i) { return i == MVC ? "mvc" : "pvc"; }

// getopt loop of src/config.h:178-243; invalid values end the program with
// status 0 there, status 2 here.
bool parse_args(int argc, char **argv, Config &c)
{
    int opt;
    while ((opt = getopt(argc, argv, "v:f:o:i:k:q:t:d:b:g:n:zh")) >= 0) {
        switch (opt) {
        case 'v':
            if (!std::strcmp(optarg, "stackonly")) c.version = STACK_ONLY;
            else if (!std::strcmp(optarg, "hybrid")) c.version = HYBRID;
            else if (!std::strcmp(optarg, "sequential")) c.version = SEQUENTIAL;
            else {
                std::fprintf(stderr, "Unrecognized -v option: %s\n", optarg);
                return false;
            }
            break;
        case 'f': c.graphFileName = optarg; break;
        case 'o': c.outputFilePrefix = optarg; break;
        case 'i':
            if (!std::strcmp(optarg, "mvc")) c.instance = MVC;
            else if (!std::strcmp(optarg, "pvc")) c.instance = PVC;
            else {
                std::fprintf(stderr, "Unrecognized -i option: %s\n", optarg);
                return false;
            }
            break;
        case 'k': c.k = (unsigned)std::atoi(optarg); break;
        case 'q': c.globalListSize = (unsigned)std::atoi(optarg); c.userListSize = true; break;
        case 't': c.globalListThreshold = (float)std::atof(optarg); c.userThreshold = true; break;
        case 'd': c.startingDepth = (unsigned)std::atoi(optarg); break;
        case 'b': c.blockDim = (unsigned)std::atoi(optarg); break;
        case 'g': c.userDefMemory = true; c.useGlobalMemory = std::atoi(optarg) != 0; break;
        case 'n': c.numBlocks = (unsigned)std::atoi(optarg); break;
        case 'z': c.indexBase = 0; break;
        case 'h': usage(); std::exit(0);
        default: std::fprintf(stderr, "\nUnrecognized option!\n"); return false;
        }
    }
    return true;
}

void append_results_row(const Config &c, const vc_csr &g, unsigned maxApprox, double timeMax, unsigned edgeApprox,
                        double timeEdge, unsigned minimum, unsigned kFound, const vc_launch_info *info, double time_ms)
{
    mkdir("Results", 0755);
    FILE *f = std::fopen("Results/Results.csv", "a");
    if (!f) {
        std::fprintf(stderr, "cannot open Results/Results.csv\n");
        return;
    }
    if (info) {
        // 22 columns of src/auxFunctions.cpp:589-593; numBlocks is the grid
        // actually launched.
        std::fprintf(f, "%s,%s,%u,%u,%s,%s,%d,%f,%d,%d,%d,%u,%d,%d,%u,%f,%u,%f,%u,%u,%u,%f\n", c.outputFilePrefix,
                     c.graphFileName, g.vertexNum, g.edgeNum, as_string(c.instance), as_string(c.version),
                     (int)info->list_size, info->list_size ? (double)info->list_threshold / info->list_size : 0.0,
                     (int)c.startingDepth, (int)info->use_global_memory, (int)info->block_dim, info->blocks_per_cu,
                     (int)(info->blocks_per_cu * info->block_dim), (int)info->num_blocks, maxApprox, timeMax,
                     edgeApprox, timeEdge, minimum, c.k, kFound, time_ms);
    } else {
        std::fprintf(f, "%s,%s,%u,%u,%s,%s,  ,  ,  ,  ,  ,  ,  ,  ,%u,%f,%u,%f,%u,%u,%u,%f\n", c.outputFilePrefix,
                     c.graphFileName, g.vertexNum, g.edgeNum, as_string(c.instance), as_string(c.version), maxApprox,
                     timeMax, edgeApprox, timeEdge, minimum, c.k, kFound, time_ms);
    }
    std::fclose(f);
}

double secs_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char **argv)
{
    Config c;
    if (!parse_args(argc, argv, c)) return 2;
    std::printf("\nGraph file: %s", c.graphFileName);
    std::printf("\nUUID: %s\n", c.outputFilePrefix);

    vc_csr_owned owned;
    vc_status st = vc_load_graph_file_indexed(c.graphFileName, c.indexBase, &owned);
    if (st != VC_OK) {
        std::fprintf(stderr, "cannot load %s: %s\n", c.graphFileName, vc_status_string(st));
        return 2;
    }
    const vc_csr &g = owned.csr;
    if (!vc_check_graph(&g)) {
        std::fprintf(stderr, "graph check failed (duplicate edge or self-loop): %s\n", c.graphFileName);
        vc_csr_free(&owned);
        return 2;
    }

    uint32_t seed = (uint32_t)std::time(nullptr);
    if (const char *s = std::getenv("VC_SEED")) seed = (uint32_t)std::strtoul(s, nullptr, 10);

    auto t0 = std::chrono::steady_clock::now();
    unsigned maxApprox = vc_approx_remove_max(&g, nullptr);
    double timeMax = secs_since(t0);
    std::printf("\nElapsed Time for Approximate Remove Max: %f\n", timeMax);
    std::printf("Approximate Remove Max Minimum is: %u\n", maxApprox);
    t0 = std::chrono::steady_clock::now();
    unsigned edgeApprox = vc_approx_remove_edge(&g, seed, nullptr);
    double timeEdge = secs_since(t0);
    std::printf("Elapsed Time for Approximate Remove Edge: %f\n", timeEdge);
    std::printf("Approximate Remove Edge Minimum is: %u\n", edgeApprox);
    std::fflush(stdout);
    unsigned minimum = maxApprox < edgeApprox ? maxApprox : edgeApprox;
    unsigned kFound = 0;

    if (c.version == SEQUENTIAL) {
        std::fprintf(stderr, "-v sequential is the reference's CPU solver and is not part of this GPU build\n");
        vc_csr_free(&owned);
        return 2;
    }
    if (c.userListSize && (c.globalListSize == 0 || (c.globalListSize & (c.globalListSize - 1)) != 0)) {
        std::fprintf(stderr, "-q must be a power of 2\n");
        vc_csr_free(&owned);
        return 2;
    }
    if (c.blockDim && (c.blockDim % 64u != 0u || c.blockDim > 1024u)) {
        // one search worker per 64-lane wave, at most 16 per workgroup
        std::fprintf(stderr, "-b must be a multiple of 64 from 64 to 1024\n");
        vc_csr_free(&owned);
        return 2;
    }

    vc_ctx *ctx = nullptr;
    if ((st = vc_ctx_create(0, &ctx)) != VC_OK) {
        std::printf("GPU Error: %s\n", vc_status_string(st));
        vc_csr_free(&owned);
        return 1;
    }
    vc_graphs *set = nullptr;
    if ((st = vc_graphs_upload(ctx, 1, &g, &set)) != VC_OK) {
        std::printf("GPU Error: %s\n", vc_status_string(st));
        vc_ctx_destroy(ctx);
        vc_csr_free(&owned);
        return 1;
    }
    vc_job job = {c.instance == PVC ? (uint32_t)VC_PVC : (uint32_t)VC_MVC, c.k, minimum, 0u, 0ull};
    vc_params p;
    std::memset(&p, 0, sizeof(p));
    p.version = c.version == STACK_ONLY ? VC_STACK_ONLY : VC_HYBRID;
    p.block_dim = c.blockDim;
    p.num_blocks = c.numBlocks;
    p.use_global_memory = c.userDefMemory ? (c.useGlobalMemory ? 1 : 0) : -1;
    p.list_size = c.userListSize ? c.globalListSize : 0u;
    p.list_threshold = c.userThreshold ? c.globalListThreshold : 0.f;
    p.start_depth = c.startingDepth;
    std::vector<uint8_t> cover(g.vertexNum ? g.vertexNum : 1);
    uint8_t *covers[1] = {cover.data()};
    vc_result r;
    vc_launch_info info;
    std::memset(&info, 0, sizeof(info));
    st = vc_solve(ctx, set, &job, &p, &r, covers, &info);
    if (st != VC_OK) {
        std::printf("GPU Error: %s\n", vc_status_string(st));
        vc_graphs_free(ctx, set);
        vc_ctx_destroy(ctx);
        vc_csr_free(&owned);
        return 1;
    }
    std::printf("\nOur Config :\n");
    std::printf("NumOfThreadPerBlocks : %u\n", info.block_dim);
    std::printf("NumOfBlocks : %u\n", info.num_blocks);
    std::printf("NumOfBlockPerSM : %u\n", info.blocks_per_cu);
    std::printf("NumOfThreadsPerSM : %u\n", info.blocks_per_cu * info.block_dim);
    std::printf("Tree nodes : %llu\n\n", (unsigned long long)info.tree_nodes);
    std::printf("Elapsed time: %fms \n", info.kernel_ms);
    if (c.instance == PVC) {
        kFound = r.k_found;
    } else {
        minimum = r.minimum;
    }
    append_results_row(c, g, maxApprox, timeMax, edgeApprox, timeEdge, minimum, kFound, &info, info.kernel_ms);

    if (c.instance == PVC) {
        if (kFound)
            std::printf("\nMinimum is less than or equal to K: %u\n\n", c.k);
        else
            std::printf("\nMinimum is greater than K: %u\n\n", c.k);
    } else {
        std::printf("\nSize of minimum vertex cover: %u\n\n", minimum);
    }
    vc_graphs_free(ctx, set);
    vc_ctx_destroy(ctx);
    vc_csr_free(&owned);
    return 0;
}
