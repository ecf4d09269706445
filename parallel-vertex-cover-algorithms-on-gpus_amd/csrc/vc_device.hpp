// vc_device.hpp -- data layout shared by the gfx950 kernels and their host
// launcher.  All structures live in HBM and are written by the host launcher
// (init kernel) and the persistent solve kernel only.
#pragma once

#include <stdint.h>

namespace vcg {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kWave = 64;

// One graph of a device graph set.  Offsets are in 32-bit words into the set
// buffer (srcPtr[n+1], dst[2m], degree[n]).  Same arrays as the reference's
// CSRGraph (src/CSRGraphRep.h:4-18).
struct GDesc {
    uint32_t n, m;
    uint32_t off_src, off_dst, off_deg;
    uint32_t off_cover;  // first 32-bit word of this graph's cover bitmap
    uint32_t instance;   // 0 MVC, 1 PVC
    uint32_t k;
    uint32_t U;          // upper bound (valid cover size)
    uint32_t seq;        // serving: request sequence number echoed in the completion
    uint64_t max_nodes;  // 0 = unlimited
    // first word of the graph's adjacency bitmap (bit n1 * n + n2 set when
    // n1 and n2 are adjacent), 0 = none (the triangle rule then searches the
    // sorted adjacency list)
    uint32_t off_bits;
    uint32_t pad1[3];
};
static_assert(sizeof(GDesc) == 64, "GDesc is one 64-byte line");
// the serving loader reads these fields by 32-bit word index
static_assert(__builtin_offsetof(GDesc, n) == 0 && __builtin_offsetof(GDesc, off_src) == 8 &&
                  __builtin_offsetof(GDesc, off_deg) == 16 && __builtin_offsetof(GDesc, U) == 32,
              "GDesc word layout");

// Mutable search state of one graph (one 64-byte line, so graphs do not
// share lines).
struct GState {
    int32_t best;         // MVC incumbent; starts at U (src/main.cu:73 with VC 1)
    uint32_t found;       // PVC: a cover <= k was found (kFound_d, src/main.cu:249-251)
    uint32_t stop;        // 1 = drop remaining work (found or budget exhausted)
    uint32_t outstanding; // live subtrees: queued items + owned by a worker
    uint64_t nodes;       // tree nodes spent on this graph
    uint32_t cover_lock;
    int32_t cover_size;   // size of the cover held in the bitmap, -1 = none
    uint32_t done;        // outstanding reached 0
    uint32_t budget_hit;  // stop was set by the node budget
    uint32_t pad0;
    uint64_t done_tick;   // device wall clock when outstanding reached 0
    uint32_t pad[2];
};
static_assert(sizeof(GState) == 64, "GState is one 64-byte line");

// Launch-wide counters (each on its own line).
struct Global {
    // Packed termination word [enqueued:32 | registered:16 | waiting:16];
    // updated with single 64-bit atomics so a dequeue moves a worker out of
    // "waiting" and an item out of "enqueued" in one step (the reference packs
    // {numWaiting, numEnqueued} the same way, src/BWDWorkList.cuh:8-11,200-205).
    unsigned long long counter;
    uint32_t pad0[14];
    unsigned long long head;  // 64-bit ring positions (the reference uses two
    uint32_t pad1[14];        // 32-bit halves of one word, src/BWDWorkList.cuh:63-69)
    unsigned long long tail;
    uint32_t pad2[14];
    int32_t count;            // occupancy guard (src/BWDWorkList.cuh:78-110)
    uint32_t pad3[15];
    uint32_t next_root;       // root / static-path hand-out counter
    uint32_t pad4[15];
    uint32_t done_flag;       // set once termination is observed
    uint32_t error;           // nonzero: an internal bound was hit (see kErr*)
    uint32_t cancel;          // cancel request (host, or the deadline)
    uint32_t timed_out;       // the deadline passed before the search finished
    uint64_t t0;              // wall clock when the first solve worker started (0 = none yet)
    uint32_t pad5[10];
    // on a line of its own: every finished unit adds to it, and the line
    // above is read by every idle poll
    unsigned long long state_bytes;  // HBM bytes of node-state slot copies (stacks, worklist)
    uint32_t pad7[14];
    // serving: device copy of the host's request tail / shutdown flag, and the
    // time of the last host poll (polls are rate-limited to one per interval)
    uint32_t host_tail;
    uint32_t host_shutdown;
    uint64_t last_poll;
    uint32_t pad6[12];
};

// Serving mode: control words and completion records in pinned host memory.
struct HostCtrl {
    uint32_t tail;      // requests published so far (ring entries [0, tail))
    uint32_t pad0[15];
    uint32_t shutdown;  // 1 = exit once every admitted request is done
    uint32_t pad1[15];
    uint32_t error;     // the kernel's glob->error, published as its workers exit
    uint32_t pad2[15];
};

struct Completion {
    uint32_t seq;       // GDesc.seq of the request, written last
    uint32_t minimum;   // incumbent (MVC) / cover size found (PVC)
    uint32_t k_found;
    uint32_t flags;     // bit0 budget exhausted, bit1 cancelled
    uint64_t nodes;
    int32_t cover_size;
    uint32_t pad;
};

constexpr unsigned long long kWaitOne = 1ull;
constexpr unsigned long long kRegOne = 1ull << 16;
constexpr unsigned long long kEnqOne = 1ull << 32;

// glob->error codes
constexpr uint32_t kErrStackOverflow = 5;

// Per-worker counters written once when the worker exits.
struct WorkerStats {
    uint64_t nodes;       // tree nodes processed
    uint64_t idle_ticks;  // device wall-clock ticks spent waiting for work
    uint64_t donations;   // subtrees put on the worklist
    uint64_t steals;      // subtrees taken from the worklist
    uint64_t exit_tick;   // device wall clock when the worker exited
};

// Degree arrays are stored as DegT (int16 when every graph of the launch
// has fewer than 32768 vertices, else int32); a node slot on a stack or in
// the worklist is npad DegT values, a multiple of 64 bytes.
struct SolveArgs {
    const uint32_t *gbuf;   // graph set words
    const GDesc *gdesc;
    GState *gstate;
    uint32_t *covers;       // cover bitmaps
    Global *glob;
    void *stk;              // [workers][depth][npad] DegT
    uint2 *stk_meta;        // [workers][depth] = (d, branch vertex)
    void *q_deg;            // [qsize][npad] DegT
    uint4 *q_meta;          // [qsize] = (graph, d, branch vertex, 0)
    uint32_t *tickets;      // [qsize]
    WorkerStats *wstats;    // [workers] (pinned host memory for a server)
    uint32_t num_graphs;
    uint32_t npad;          // slot length in DegT values
    uint32_t lcap;          // mark-list capacity per worker (multiple of 4)
    uint32_t ncap;          // capacity of each rule-candidate list (multiple of 4)
    uint32_t mwords;        // mark-bitset words per worker (multiple of 4)
    uint32_t worker_bytes;  // bytes of per-worker node storage (LDS or HBM)
    uint32_t depth;         // stack slots per worker
    uint32_t qsize;         // power of two
    uint32_t qthreshold;    // donate while enqueued < qthreshold
    uint32_t workers;
    uint32_t version;       // 0 static split (stack only), 1 hybrid
    uint32_t start_depth;   // static split depth
    uint32_t roots;         // number of root items handed out by next_root
    uint64_t limit_ticks;   // deadline = glob->t0 + limit_ticks (device wall clock)
    char *hbm_nodes;        // [workers][worker_bytes] node storage in HBM mode
    // serving mode: roots come from a request ring published by a host
    // thread in pinned host memory; the worker that claims a request copies
    // its descriptor and graph into HBM (gdesc / gbuf, same word offsets)
    uint32_t serving;
    uint32_t ring_mask;        // ring entries - 1 (power of two)
    const uint32_t *req_ring;  // pinned: [ring] graph slot of each request
    const GDesc *host_desc;    // pinned: [slots] request descriptors
    uint32_t *host_arena;      // pinned: request graphs in, cover bitmaps out
    HostCtrl *host;            // pinned
    Completion *comp;          // pinned: one record per slot
    uint64_t poll_ticks;       // minimum interval between host polls
    // longest sleep of an idle worker between polls, as log2 of 512-cycle
    // sleeps: every poll reads launch-wide words with device scope, so
    // thousands of idle waves polling often load the memory channels that
    // the busy waves (and other kernels) use
    uint32_t idle_cap;
    // 1 = every graph of the launch has 2m < 2^16 and node storage is in
    // LDS: a worker keeps its current graph's srcPtr there as uint16
    uint32_t off16;
    // staged start of launches with few roots: workers wid >= active0 first
    // sleep about dormant_ticks and join only while the running ones are
    // nearly all busy (0 = every worker starts at once)
    uint32_t active0;
    uint64_t dormant_ticks;
    // node-budget flushes: workers per graph, the least number of workers
    // assumed to share a graph (vcg::flush_interval)
    uint32_t flush_share;
};

}  // namespace vcg
