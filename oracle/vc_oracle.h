/*
 * vc_oracle.h -- CPU oracle for the vertex-cover branch-and-reduce path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * CPU solver (src/Sequential.cpp, src/SequentialParameterized.cpp,
 * src/helperFunctions.cpp), its graph loader (src/auxFunctions.cpp:129-209),
 * its graph check (src/auxFunctions.cpp:308-354) and its two upper-bound
 * approximations (src/auxFunctions.cpp:44-127,211-306, src/CSRGraphRep.cpp:46-64).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker.  The product (the HIP library) never
 * links or calls anything here.
 *
 * Pinning (what checks each function against something other than itself):
 *   or_sequential / or_sequential_parameterized : the compiled reference
 *     (oracle/_ref, same MVC, same kFound, same node count and same per-rule
 *     call counts), the reference's MVCS table (test/runGraphs.py:8-30),
 *     brute force on tiny graphs, closed forms, an independent MILP solver.
 *   or_remove_max_approx / or_remove_edge_approx : oracle/_ref (bit-exact U;
 *     RemoveEdge with the reference's time(0) seed replaced by a fixed seed).
 *   or_graph_load : oracle/_ref (same CSR arrays) and the reference's data
 *     files.
 *   or_check_graph : the reference's data files, and the definition (no list
 *     repeats a vertex or names its own vertex) checked by brute force on
 *     random unsorted lists, which the reference's loader never produces.
 */
#ifndef VC_ORACLE_H
#define VC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CSR graph with the layout of the reference's CSRGraph
 * (src/CSRGraphRep.h:4-18): srcPtr[n+1], dst[2m] sorted per vertex,
 * degree[n] = adjacency length. */
typedef struct {
    uint32_t n;
    uint32_t m;
    uint32_t *srcPtr;
    uint32_t *dst;
    int32_t *degree;
} or_graph;

/* Counters of one oracle solve.  "nodes" is one iteration of the node loop
 * (src/Sequential.cpp:22-97); the *_calls fields count the calls the
 * reference's driver loop makes into src/helperFunctions.cpp, so they can be
 * compared with the instrumented reference build in oracle/_ref. */
typedef struct {
    uint64_t max_nodes;      /* in: 0 = unlimited; else stop after this many nodes */
    uint64_t nodes;          /* out */
    uint64_t leaf_calls;     /* out: leafReductionRule calls */
    uint64_t tri_calls;      /* out: triangleReductionRule calls */
    uint64_t high_calls;     /* out: highDegreeReductionRule calls */
    uint64_t square_calls;   /* out: squareSequential calls */
    uint64_t delete_calls;   /* out: deleteVertex calls made by the node loop */
    uint64_t branches;       /* out */
    int32_t max_top;         /* out: deepest stack index used */
    int32_t completed;       /* out: 1 = search finished, 0 = stopped at max_nodes */
} or_stats;

/* Loader (src/auxFunctions.cpp:129-209): header "n m", m lines of 1-indexed
 * "u v".  Returns 0 on success, negative on error (-1 open, -2 parse,
 * -3 vertex index out of range). */
int or_graph_load(const char *path, or_graph *g);

/* Build the same CSR from a 0-indexed edge list (edge i = (eu[i], ev[i])).
 * Returns 0 on success, -3 on an index out of range. */
int or_graph_from_edges(uint32_t n, uint32_t m, const uint32_t *eu, const uint32_t *ev, or_graph *g);

void or_graph_free(or_graph *g);

/* check_graph (src/auxFunctions.cpp:308-354): 1 if no duplicate edge and no
 * self-loop, else 0. */
int or_check_graph(const or_graph *g);

/* RemoveMaxApproximateMVC (src/auxFunctions.cpp:211-256). */
uint32_t or_remove_max_approx(const or_graph *g);

/* RemoveEdgeApproximateMVC (src/auxFunctions.cpp:265-306).  The reference
 * seeds rand() with time(0) inside every getRandom call (:258-263); here the
 * seed is an argument and glibc's srand/rand are used the same way. */
uint32_t or_remove_edge_approx(const or_graph *g, uint32_t seed);

/* Sequential(graph, minimum) (src/Sequential.cpp:3-104).  Returns the final
 * incumbent.  If cover is non-NULL (n bytes) it receives the cover found at
 * the last incumbent update ({v : deg[v] == -1}, src/Sequential.cpp:73-77);
 * *cover_found is set to 1 when such an update happened.  Unlike the
 * reference, the caller's graph is not freed (src/Sequential.cpp:99). */
uint32_t or_sequential(const or_graph *g, uint32_t minimum, or_stats *st,
                       uint8_t *cover, int32_t *cover_found);

/* SequentialParameterized(graph, minimum, k, &kFound)
 * (src/SequentialParameterized.cpp:4-111).  Returns minimum (unchanged, as
 * the reference does).  On kFound, cover receives {v : deg[v] == -1}. */
uint32_t or_sequential_parameterized(const or_graph *g, uint32_t minimum, uint32_t k,
                                     uint32_t *kFound, or_stats *st, uint8_t *cover);

/* 1 if cover (n bytes, nonzero = in cover) covers every edge of g. */
int or_cover_is_valid(const or_graph *g, const uint8_t *cover);

/* CPU baseline over many graphs: `threads` POSIX threads pull graph indices
 * from a shared counter and run or_remove_max_approx + or_sequential on each
 * (MVC).  results[i] = MVC (or the incumbent at the cap), nodes[i] = tree
 * nodes, done[i] = 1 when the search completed within max_nodes (0 = no cap).
 * Returns wall seconds. */
double or_solve_batch_mt(int count, const or_graph *graphs, int threads, uint64_t max_nodes,
                         uint32_t *results, uint64_t *nodes, int32_t *done);

#ifdef __cplusplus
}
#endif
#endif
