/*
 * vcgpu.h -- C ABI of the MI355X (gfx950) branch-and-reduce vertex-cover solver.
 *
 * The reference has no FFI layer: its boundary is main()'s launch block
 * (src/main.cu:300-344) handing by-value structs of device pointers to eight
 * __global__ kernels, plus the host helpers it calls before and after.  Each
 * entry point below replaces one of those pieces; the comment on each cites
 * the reference code it stands in for.  Plain C types only, no exit(): every
 * call returns a vc_status.
 *
 * Threading: one vc_ctx per host thread; calls on one ctx are serialised by
 * the caller.  A vc_solve call blocks until its persistent kernel finishes.
 */
#ifndef VCGPU_H
#define VCGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    VC_OK = 0,
    VC_EINVAL = 1,     /* bad argument (reference: exit(0) in src/config.h:104-105, 143-144) */
    VC_ENOMEM = 2,     /* device memory does not fit (reference: src/auxFunctions.cpp:399-402) */
    VC_EDEVICE = 3,    /* HIP error (reference: "GPU Error" exit(1), src/main.cu:348-352) */
    VC_ETIMEOUT = 4,   /* node budget or spin bound exhausted before the search finished */
    VC_ECANCELLED = 5, /* cancelled by the caller */
    VC_ENOTSUP = 6,    /* combination not supported */
    VC_EBADGRAPH = 7,  /* duplicate edge / self-loop / index out of range
                          (reference: assert(check_graph), src/auxFunctions.cpp:358) */
    VC_EAGAIN = 8      /* serving: the request has not finished yet */
} vc_status;

/* src/config.h:125-129 */
typedef enum { VC_MVC = 0, VC_PVC = 1 } vc_instance;

/* src/config.h:81-86 (SEQUENTIAL is the CPU path and has no device entry). */
typedef enum { VC_STACK_ONLY = 0, VC_HYBRID = 1 } vc_version;

/* CSR graph as src/CSRGraphRep.h:4-18 lays it out: srcPtr[vertexNum+1],
 * dst[2*edgeNum] sorted ascending per vertex, degree[v] = srcPtr[v+1]-srcPtr[v].
 * (srcPtrUncompressed is used only by the out-of-scope matching kernel.) */
typedef struct {
    uint32_t vertexNum;
    uint32_t edgeNum;
    const uint32_t *srcPtr;
    const uint32_t *dst;
    const int32_t *degree;
} vc_csr;

/* Per-graph job: what one search is asked.  upper_bound plays the role of
 * `minimum` in src/main.cu:73 (a valid cover size U from the host
 * approximations); for MVC the search looks for covers smaller than U and
 * returns U when none exists, as src/Sequential.cpp:67-77 does. */
typedef struct {
    uint32_t instance;    /* vc_instance */
    uint32_t k;           /* PVC parameter (-k, src/config.h:211-213) */
    uint32_t upper_bound; /* U; must be the size of some vertex cover */
    uint32_t reserved;
    uint64_t max_nodes;   /* per-graph tree-node budget, 0 = unlimited; a graph that reaches
                             it stops within a few percent of it (workers flush their node
                             counts more often as the budget runs out) and reports
                             VC_ETIMEOUT with its incumbent */
} vc_job;

/* Launch-level parameters: the reference's Config (src/config.h:162-176)
 * fields that shape the device solve.  0 / -1 mean "pick automatically". */
typedef struct {
    uint32_t version;           /* vc_version (-v) */
    uint32_t block_dim;         /* threads per workgroup (-b): a multiple of 64 up to 1024,
                                   one search worker per 64-thread wave; 0 = auto.  When an
                                   explicit workgroup's workers do not fit in LDS their node
                                   state goes to HBM (as with -g 1) */
    uint32_t num_blocks;        /* workgroups (-n); 0 = auto (occupancy x CUs).  At most
                                   65535 workers (num_blocks x block_dim / 64) per launch.
                                   With 0 and fewer roots (graphs x static paths) than half
                                   the workers, 512 workers start and the rest join once
                                   they are nearly all busy and the launch has run 0.8 ms
                                   (a small search ends before they do); an explicit
                                   num_blocks starts every worker at once */
    int32_t use_global_memory;  /* -g: -1 auto, 0 degree arrays in LDS, 1 in HBM */
    uint32_t list_size;         /* global worklist slots (-q, power of 2); 0 = auto */
    float list_threshold;       /* donate while queued < threshold*size (-t); <=0 = auto */
    uint32_t start_depth;       /* static split depth for VC_STACK_ONLY (-d) */
    uint32_t time_limit_ms;     /* wall-clock limit of the launch; 0 = one hour.  When it
                                   passes every worker drains and exits, and unfinished
                                   graphs report VC_ETIMEOUT. */
    uint32_t flags;             /* VC_FLAG_* */
} vc_params;

/* vc_params.flags: run the solve as a HIP graph (init kernel, persistent
 * solve kernel and result copies captured once per graph set and launch
 * shape, replayed on later calls with the same shape). */
#define VC_FLAG_GRAPH 1u
/* vc_params.flags: store degrees as int32 (with 64-bit sketch sums) even
 * when every graph has fewer than 32768 vertices, as graphs beyond that size
 * always do.  Only useful to exercise that storage variant on small inputs. */
#define VC_FLAG_WIDE 2u

/* Per-graph outcome.  Replaces the 4-byte minimum_d / kFound_d readback of
 * src/main.cu:355-359 and adds the node count and status. */
typedef struct {
    uint32_t minimum;        /* MVC: cover size found (== MVC when solved);
                                PVC: size of the cover found when k_found */
    uint32_t k_found;        /* PVC: 1 = a cover of size <= k exists */
    uint32_t solved;         /* 1 = search finished (answer exact) */
    uint32_t cover_from_gpu; /* 1 = cover[] was written by the device search;
                                0 = no cover < U exists (MVC) and cover[] is
                                untouched (the caller holds the U-cover) */
    uint64_t tree_nodes;     /* node-loop iterations spent on this graph */
    int32_t status;          /* vc_status for this graph */
    uint32_t done_us;        /* vc_solve: when this graph's search ended, in
                                microseconds after the launch's first worker
                                started (0 from a server) */
} vc_result;

/* Launch facts of the last vc_solve (the reference prints these,
 * src/main.cu:200-204, and writes them to Results.csv, src/auxFunctions.cpp:589-593). */
typedef struct {
    uint32_t num_blocks;
    uint32_t block_dim;
    uint32_t workers;
    uint32_t blocks_per_cu;
    uint32_t use_global_memory;
    uint32_t list_size;
    uint32_t list_threshold;
    uint32_t stack_depth;
    float kernel_ms;
    uint32_t cu_count;
    uint64_t tree_nodes;
    uint64_t donations;   /* subtrees put on the worklist */
    uint64_t steals;      /* subtrees taken from the worklist */
    float idle_fraction;  /* share of worker time spent waiting for work */
    float active_ms;      /* device wall clock from the first worker's start to the
                             last searching worker's exit (workers of a staged start
                             that never joined do not count; 0 from a server) */
    uint64_t state_bytes; /* HBM bytes of node state the search moved: every stack
                             push and pop, worklist steal, and (twice) donation copies
                             the graph's degree/sketch records, 8 bytes per vertex (16
                             in the int32 layout) in 16-byte vectors; 0 from a server */
} vc_launch_info;

typedef struct vc_ctx vc_ctx;
typedef struct vc_graphs vc_graphs;

/* Device context on one GPU.  Replaces the device queries of
 * src/main.cu:101-121. */
vc_status vc_ctx_create(int device, vc_ctx **out);
vc_status vc_ctx_destroy(vc_ctx *ctx);

/* Upload `count` graphs into one device-resident graph set (one HBM buffer).
 * count == 1 replaces allocateGraph (src/CSRGraphRep.cuh:3-27); larger
 * counts are the batched form.  Graphs are checked (check_graph,
 * src/auxFunctions.cpp:308-354, plus what the kernels rely on: offsets
 * consistent with degree[] and edgeNum, each adjacency sorted ascending as
 * createCSRGraphFromFile leaves it, and every edge listed from both ends)
 * and rejected with VC_EBADGRAPH. */
vc_status vc_graphs_upload(vc_ctx *ctx, int count, const vc_csr *graphs, vc_graphs **out);
vc_status vc_graphs_free(vc_ctx *ctx, vc_graphs *set);

/* Solve every graph of the set in ONE persistent launch: jobs[count],
 * results[count]; covers == NULL or covers[i] == NULL skips that cover,
 * otherwise covers[i] receives vertexNum bytes (1 = in cover).
 * Replaces src/main.cu:219-366 (stacks, worklist, launch, readback) for the
 * four kernels GlobalWorkList{,Parameterized}_{shared,global}_kernel
 * (src/GlobalWorkList.cuh:12-18, src/GlobalWorkListParameterized.cuh:12-18) and
 * LocalStacks{,Parameterized}_{shared,global}_kernel
 * (src/LocalStacks.cuh:10-16, src/LocalStacksParameterized.cuh:10-16). */
vc_status vc_solve(vc_ctx *ctx, const vc_graphs *set, const vc_job *jobs, const vc_params *params,
                   vc_result *results, uint8_t *const *covers, vc_launch_info *info);

/* Allocation-free variant of vc_solve.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and covers as vc_solve; it calls no hipMalloc, hipFree or
 * hipHostMalloc.  Its launch scratch (stacks, worklist, counters, worker
 * statistics, read-back) is one fixed device array of 128 MiB per GPU, which
 * the library holds from the moment it is loaded on that GPU until the process
 * ends.  The grid shrinks until the stacks fit in it; a set whose smallest
 * launch (one workgroup with exact-depth stacks) does not fit returns
 * VC_ENOMEM, as does a set whose read-back (32 bytes per graph plus one byte
 * per vertex) alone would fill it.  Every call on one GPU shares that array
 * and leaves its last contents there (each call clears what must start at
 * zero); calls that overlap on one GPU race, so the library serialises its
 * own calls per GPU with a host lock (each process has its own array).
 * VC_FLAG_GRAPH is refused with VC_ENOTSUP.  Results return by
 * device-to-host copies into results[] and covers[i]. */
vc_status vc_solve_noalloc(vc_ctx *ctx, const vc_graphs *set, const vc_job *jobs, const vc_params *params,
                           vc_result *results, uint8_t *const *covers, vc_launch_info *info);

/* Workgroups per CU the solve would run with (replaces
 * cudaOccupancyMaxActiveBlocksPerMultiprocessor, src/main.cu:131-191). */
vc_status vc_occupancy(vc_ctx *ctx, const vc_graphs *set, const vc_params *params, int *blocks_per_cu,
                       int *block_dim, int *use_global_memory);

/* Launch-shape autotuner (replaces the reference's fixed search,
 * setBlockDimAndUseGlobalMemoryAutoWorkListSize, src/auxFunctions.cpp:469-578):
 * times candidate shapes -- workgroup size, LDS or HBM node storage,
 * worklist size -- by solving `set` with `jobs`, and writes the fastest into
 * *best (other fields copied from *base).  Results are cached in the
 * context by (mean vertex count, mean density) bucket, so later sets of the
 * same kind are not re-timed.  *best_ms (optional) gets the winning kernel
 * time (0 when served from the cache). */
vc_status vc_autotune(vc_ctx *ctx, const vc_graphs *set, const vc_job *jobs, const vc_params *base,
                      vc_params *best, float *best_ms);

/* ---- serving: continuous batching over one persistent launch ----
 *
 * No counterpart in the reference (which solves one graph per process); this
 * is the request loop of the served configuration.  A server owns one
 * persistent solve kernel on its own stream.  vc_server_submit copies a
 * graph and returns a ticket at once; a host thread admits queued requests
 * as request slots and arena memory (graph + cover + the worker-stack
 * footprint check) free up, uploads them while the kernel runs, and
 * publishes them to the kernel, whose idle workers pick them up.  Search
 * state stays on the device; the kernel reports each completion to pinned
 * host memory. */
typedef struct vc_server vc_server;

typedef struct {
    uint32_t max_vertices;  /* largest graph the server admits (sizes LDS and stacks) */
    uint32_t slots;         /* requests on the device at once; 0 = 4096 */
    uint64_t arena_bytes;   /* HBM for the graphs + covers in flight; 0 = 1 GiB */
    uint32_t num_blocks;    /* workgroups of the persistent kernel; 0 = 2 workers per CU
                               (at most the occupancy): idle waves poll, so a
                               small grid serves a request stream faster and
                               leaves CUs to other kernels */
    uint32_t block_dim;     /* threads per workgroup (multiple of 64, <= 1024); 0 = auto */
    uint32_t stack_depth;   /* stack entries per worker; 0 = max_vertices + 1 */
    uint32_t time_limit_ms; /* lifetime limit of the server kernel; 0 = one hour */
} vc_server_config;

vc_status vc_server_start(vc_ctx *ctx, const vc_server_config *cfg, vc_server **out);
/* Queue one request (graph copied; U = job->upper_bound must be a valid
 * cover size).  *ticket identifies it in poll/wait.  Graphs larger than
 * max_vertices are rejected with VC_EINVAL; a request whose stack need
 * min(n, max(U, k+1)) + 1 exceeds the server's stack_depth with VC_ENOTSUP. */
vc_status vc_server_submit(vc_server *s, const vc_csr *graph, const vc_job *job, uint64_t *ticket);
/* VC_EAGAIN while pending; otherwise fills *result (and cover, vertexNum
 * bytes, if non-NULL) and forgets the ticket.  *done_ns (optional) is the
 * host monotonic-clock time at which the host thread saw the completion. */
vc_status vc_server_poll(vc_server *s, uint64_t ticket, vc_result *result, uint8_t *cover, uint64_t *done_ns);
/* Blocks until the request finishes or timeout_ms passes (VC_EAGAIN). */
vc_status vc_server_wait(vc_server *s, uint64_t ticket, uint32_t timeout_ms, vc_result *result, uint8_t *cover,
                         uint64_t *done_ns);
/* Host monotonic clock in ns (the clock of done_ns). */
uint64_t vc_now_ns(void);
/* Finishes every admitted and queued request, stops the kernel, frees all. */
vc_status vc_server_stop(vc_server *s, vc_launch_info *info);

/* ---- host helpers the reference's driver calls around the solve ---- */

/* Owned CSR filled by vc_load_graph_file; free with vc_csr_free. */
typedef struct {
    vc_csr csr;
    void *storage;
} vc_csr_owned;

/* createCSRGraphFromFile (src/auxFunctions.cpp:129-209): header "n m" then
 * m 1-indexed edges; adjacency sorted.  Errors are returned, not asserted. */
vc_status vc_load_graph_file(const char *path, vc_csr_owned *out);
/* The same loader with the first vertex numbered index_base (0 or 1; other
 * values are VC_EINVAL).  index_base 1 is vc_load_graph_file.  index_base 0
 * reads the reference's 0-indexed files (test/test_graphs/formated_* and
 * lastfm-asia.txt, whose MVC test/runGraphs.py:21-29 lists, and
 * data/GraphDatabase/Extra/{60,120}-cell.txt), which its own loader
 * (src/auxFunctions.cpp:151-153) shifts to vertex -1. */
vc_status vc_load_graph_file_indexed(const char *path, uint32_t index_base, vc_csr_owned *out);
/* Same CSR from a 0-indexed edge list. */
vc_status vc_csr_from_edges(uint32_t n, uint32_t m, const uint32_t *eu, const uint32_t *ev, vc_csr_owned *out);
void vc_csr_free(vc_csr_owned *g);

/* check_graph (src/auxFunctions.cpp:308-354): 1 = valid (no duplicate
 * neighbour in any order, no self-loop, indices and offsets in range),
 * computed in O(n + m) instead of the reference's n x n matrix. */
int vc_check_graph(const vc_csr *g);

/* RemoveMaxApproximateMVC (src/auxFunctions.cpp:211-256).  Returns U; if
 * cover != NULL it receives the greedy cover (vertexNum bytes).  Both
 * approximations give the reference's U bit for bit; their leaf and
 * triangle passes visit only vertices whose test can have changed, in the
 * reference's order, instead of rescanning every vertex. */
uint32_t vc_approx_remove_max(const vc_csr *g, uint8_t *cover);

/* RemoveEdgeApproximateMVC (src/auxFunctions.cpp:265-306); the reference
 * seeds with time(0) inside each random draw, here the seed is explicit. */
uint32_t vc_approx_remove_edge(const vc_csr *g, uint32_t seed, uint8_t *cover);

const char *vc_status_string(vc_status s);

#ifdef __cplusplus
}
#endif
#endif /* VCGPU_H */
