// ref_driver.cpp -- driver for the reference's own CPU solver files, built by
// oracle/Makefile straight from the reference's src/ directory (never copied)
// into oracle/_ref/ref_driver.  TEST INFRASTRUCTURE ONLY: it exists to pin
// the C restatement in oracle/vc_oracle.c against the real reference code.
//
// The link wraps (ld --wrap) the calls that src/Sequential.cpp and
// src/SequentialParameterized.cpp make into src/helperFunctions.cpp, so the
// driver can report per-rule call counts and the exact number of node-loop
// iterations without touching the reference sources:
//   * a node starts at every leafReductionRule call that does not continue
//     the rule do/while of the previous node (src/Sequential.cpp:38-44);
//   * time() is wrapped so RemoveEdgeApproximateMVC's srand(time(0))
//     (src/auxFunctions.cpp:260) can be given a fixed seed.
//
// Usage: ref_driver GRAPH [--pvc K] [--min U] [--seed S] [--approx-only]
// Prints one JSON object.
#include "CSRGraphRep.h"
#include "Sequential.h"
#include "SequentialParameterized.h"
#include "auxFunctions.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>

static long long g_seed = -1;
static unsigned long long c_nodes, c_leaf, c_tri, c_high, c_square, c_delete;
static bool g_pending_continue = false, g_last_tri = false;

extern "C" {
time_t __real_time(time_t *t);
time_t __wrap_time(time_t *t)
{
    if (g_seed >= 0) {
        if (t) *t = (time_t)g_seed;
        return (time_t)g_seed;
    }
    return __real_time(t);
}

bool __real__Z17leafReductionRule8CSRGraphPiPj(CSRGraph g, int *d, unsigned *nd);
bool __wrap__Z17leafReductionRule8CSRGraphPiPj(CSRGraph g, int *d, unsigned *nd)
{
    if (!g_pending_continue) ++c_nodes;
    g_pending_continue = false;
    ++c_leaf;
    return __real__Z17leafReductionRule8CSRGraphPiPj(g, d, nd);
}

bool __real__Z21triangleReductionRule8CSRGraphPiPj(CSRGraph g, int *d, unsigned *nd);
bool __wrap__Z21triangleReductionRule8CSRGraphPiPj(CSRGraph g, int *d, unsigned *nd)
{
    ++c_tri;
    g_last_tri = __real__Z21triangleReductionRule8CSRGraphPiPj(g, d, nd);
    return g_last_tri;
}

bool __real__Z23highDegreeReductionRule8CSRGraphPiPji(CSRGraph g, int *d, unsigned *nd, int minimum);
bool __wrap__Z23highDegreeReductionRule8CSRGraphPiPji(CSRGraph g, int *d, unsigned *nd, int minimum)
{
    ++c_high;
    bool r = __real__Z23highDegreeReductionRule8CSRGraphPiPji(g, d, nd, minimum);
    g_pending_continue = g_last_tri || r;
    return r;
}

long long __real__Z16squareSequentiali(int num);
long long __wrap__Z16squareSequentiali(int num)
{
    ++c_square;
    return __real__Z16squareSequentiali(num);
}

int *__real__Z12deleteVertexR8CSRGraphjPiPj(CSRGraph &g, unsigned v, int *d, unsigned *nd);
int *__wrap__Z12deleteVertexR8CSRGraphjPiPj(CSRGraph &g, unsigned v, int *d, unsigned *nd)
{
    ++c_delete;
    return __real__Z12deleteVertexR8CSRGraphjPiPj(g, v, d, nd);
}
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s GRAPH [--pvc K] [--min U] [--seed S] [--approx-only]\n", argv[0]);
        return 2;
    }
    const char *path = argv[1];
    bool pvc = false, approx_only = false;
    unsigned k = 0;
    long long umin = -1;
    for (int i = 2; i < argc; ++i) {
        if (!strcmp(argv[i], "--pvc") && i + 1 < argc) {
            pvc = true;
            k = (unsigned)atoll(argv[++i]);
        } else if (!strcmp(argv[i], "--min") && i + 1 < argc) {
            umin = atoll(argv[++i]);
        } else if (!strcmp(argv[i], "--seed") && i + 1 < argc) {
            g_seed = atoll(argv[++i]);
        } else if (!strcmp(argv[i], "--approx-only")) {
            approx_only = true;
        }
    }
    FILE *probe = fopen(path, "r");
    if (!probe) {
        fprintf(stderr, "cannot open %s\n", path);
        return 2;
    }
    fclose(probe);

    CSRGraph graph = createCSRGraphFromFile(path);
    bool ok = check_graph(graph);
    unsigned umax = RemoveMaxApproximateMVC(graph);
    long long uedge = -1;
    if (g_seed >= 0) uedge = RemoveEdgeApproximateMVC(graph);
    unsigned minimum = umin >= 0 ? (unsigned)umin : umax;

    printf("{\"n\": %u, \"m\": %u, \"check\": %d, \"U_max\": %u, \"U_edge\": %lld", graph.vertexNum,
           graph.edgeNum, ok ? 1 : 0, umax, uedge);
    if (!approx_only) {
        CSRGraph copy;
        copy.copy(graph); // Sequential* free the graph they are given
        unsigned kFound = 0, result;
        c_nodes = c_leaf = c_tri = c_high = c_square = c_delete = 0;
        g_pending_continue = g_last_tri = false;
        auto t0 = std::chrono::steady_clock::now();
        if (pvc)
            result = SequentialParameterized(copy, minimum, k, &kFound);
        else
            result = Sequential(copy, minimum);
        auto t1 = std::chrono::steady_clock::now();
        printf(", \"minimum_in\": %u, \"pvc\": %d, \"k\": %u, \"result\": %u, \"kFound\": %u, \"nodes\": %llu, "
               "\"leaf_calls\": %llu, \"tri_calls\": %llu, \"high_calls\": %llu, \"square_calls\": %llu, "
               "\"delete_calls\": %llu, \"seconds\": %.6f",
               minimum, pvc ? 1 : 0, k, result, kFound, c_nodes, c_leaf, c_tri, c_high, c_square, c_delete,
               std::chrono::duration<double>(t1 - t0).count());
    }
    printf("}\n");
    return 0;
}
