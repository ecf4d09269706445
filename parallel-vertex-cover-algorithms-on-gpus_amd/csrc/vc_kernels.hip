// vc_kernels.hip -- gfx950 persistent branch-and-reduce kernel for minimum
// (MVC) and parameterized (PVC) vertex cover.
//
// One wave64 is one search worker.  A workgroup holds several independent
// workers; no workgroup barrier is ever used.  Each worker keeps the degree
// array of its current tree node (-1 = vertex in the cover) in LDS -- or, for
// graphs too large for LDS, in a per-worker HBM scratch area (the reference's
// `_global_` kernels) -- plus a mark bitset and a compact mark list used by
// the reduction rules.  Degrees are int16 when every graph of the launch has
// fewer than 32768 vertices, halving LDS use and stack/worklist traffic.
//
// Reference semantics kept (file:line in the reference):
//   leaf rule            src/helperFunctions.cuh:229-296
//   high-degree rule     src/helperFunctions.cuh:299-358
//   triangle rule        src/helperFunctions.cuh:361-447
//   rule fixpoint        src/GlobalWorkList.cuh:107-130
//   edge bound           src/GlobalWorkList.cuh:132-139,
//                        src/GlobalWorkListParameterized.cuh:135-137
//   branching            src/helperFunctions.cuh:156-215
//   local stacks         src/stack.cuh:10-32
//   broker worklist      src/BWDWorkList.cuh:13-265
//   static path split    src/LocalStacks.cuh:28-162
// Every rule is applied in mark/apply passes like the reference's GPU rules,
// but over one wave with ballot compaction instead of block-wide barriers.
// The branch vertex is the lowest-index vertex of maximum degree (the CPU
// tie-break of src/Sequential.cpp:50-58; the reference GPU tie-break depends
// on blockDim, src/helperFunctions.cuh:453-478).
#include <hip/hip_runtime.h>

#include "vc_device.hpp"

namespace vcg {

// ---------------------------------------------------------------------------
// wave primitives
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Orders this wave's accesses to its node storage across lanes.  In LDS a
// wave's DS instructions execute in issue order, so only the compiler has to
// be held back; in HBM mode a workgroup-scope fence also waits for the
// wave's stores before later loads.
template <bool HBM>
__device__ __forceinline__ void nsync()
{
    if constexpr (HBM) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// DPP steps inside a row of 16 lanes: quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_ror:4, row_ror:8.  After the four steps every lane of a row
// holds the row's reduction; the four rows are combined with readlane.
template <int CTRL>
__device__ __forceinline__ int dpp(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}

__device__ __forceinline__ int wave_sum(int x)
{
    x += dpp<0xB1>(x);
    x += dpp<0x4E>(x);
    x += dpp<0x124>(x);
    x += dpp<0x128>(x);
    return __builtin_amdgcn_readlane(x, 0) + __builtin_amdgcn_readlane(x, 16) + __builtin_amdgcn_readlane(x, 32) +
           __builtin_amdgcn_readlane(x, 48);
}

__device__ __forceinline__ int wave_max(int x)
{
    x = max(x, dpp<0xB1>(x));
    x = max(x, dpp<0x4E>(x));
    x = max(x, dpp<0x124>(x));
    x = max(x, dpp<0x128>(x));
    int a = max(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16));
    int b = max(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48));
    return max(a, b);
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t xu)
{
    int x = (int)xu;
    // unsigned min through the order-preserving flip of the sign bit
    x ^= (int)0x80000000;
    x = min(x, dpp<0xB1>(x));
    x = min(x, dpp<0x4E>(x));
    x = min(x, dpp<0x124>(x));
    x = min(x, dpp<0x128>(x));
    int a = min(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16));
    int b = min(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48));
    return (uint32_t)(min(a, b) ^ (int)0x80000000);
}

// ---------------------------------------------------------------------------
// HBM atomics (agent scope)
// ---------------------------------------------------------------------------

template <class T>
__device__ __forceinline__ T ld_rlx(T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_rlx(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T ld_acq(T *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_rel(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T fetch_add(T *p, T v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void fence_rel() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acq() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

__device__ __forceinline__ void backoff(uint32_t step)
{
    uint32_t reps = 1u << (step < 6 ? step : 6);
    for (uint32_t i = 0; i < reps; ++i) __builtin_amdgcn_s_sleep(8);
}

// True once the launch's time limit has passed; the first lane to see it
// raises cancel so every worker drains and exits.  Every wait loop of the
// kernel calls this, so no wave can spin forever.
__device__ __forceinline__ bool past_deadline(const SolveArgs &a)
{
    if (a.limit_ticks == 0) return false;
    uint64_t now = wall_clock64();
    if (now - ld_rlx(&a.glob->t0) <= a.limit_ticks) return false;
    st_rlx(&a.glob->timed_out, 1u);
    st_rlx(&a.glob->cancel, 1u);
    return true;
}

// ---------------------------------------------------------------------------
// one worker's view of its current tree node
// ---------------------------------------------------------------------------

template <typename DegT, bool HBM>
struct Node {
    DegT *deg;           // [npad]: degree, -1 = in cover
    uint32_t *mark;      // [mwords] bitset, all zero between passes
    uint32_t *list;      // [lcap]: vertices marked in the current pass
    const uint32_t *src; // graph srcPtr (HBM)
    const uint32_t *dst; // graph dst (HBM)
    uint32_t n;
    uint32_t lcap;
    using deg_t = DegT;

    __device__ __forceinline__ void sync() const { nsync<HBM>(); }
};

// Appends cand (kNone = nothing) of every lane to the mark list; with dedupe
// a vertex already marked in this pass is dropped.  At most lcap entries are
// kept: a candidate past the end is unmarked again and found by the next
// pass (applying any subset of a pass's marks is sound, since the whole set
// is).  Returns the new length (wave-uniform).
template <class N>
__device__ __forceinline__ uint32_t append(const N &w, uint32_t L, uint32_t cand, bool dedupe)
{
    bool keep = cand != kNone;
    uint32_t bit = 1u << (cand & 31u);
    if (dedupe && keep) {
        uint32_t old = __hip_atomic_fetch_or(&w.mark[cand >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        keep = (old & bit) == 0;
    }
    uint64_t m = __ballot(keep);
    if (keep) {
        uint32_t pos = L + lanes_below(m);
        if (pos < w.lcap)
            w.list[pos] = cand;
        else if (dedupe)
            __hip_atomic_fetch_and(&w.mark[cand >> 5], ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    return min(L + (uint32_t)__popcll(m), w.lcap);
}

// Apply phase of a rule: every listed vertex goes into the cover.  First all
// of them get -1, then each one decrements its neighbours that are still
// live (src/helperFunctions.cuh:270-289).  Listed vertices are distinct and
// live.  Neighbours of one vertex are distinct, so one vertex's decrements
// need no atomics; vertices are taken one after the other.
template <class N>
__device__ void apply_list(const N &w, uint32_t L)
{
    w.sync();
    for (uint32_t i = lane_id(); i < L; i += kWave) w.deg[w.list[i]] = -1;
    w.sync();
    for (uint32_t c0 = 0; c0 < L; c0 += kWave) {
        uint32_t cnt = min(L - c0, (uint32_t)kWave);
        uint32_t my_b = 0, my_e = 0;
        if (lane_id() < cnt) {
            uint32_t u = w.list[c0 + lane_id()];
            my_b = w.src[u];
            my_e = w.src[u + 1];
        }
        for (uint32_t j = 0; j < cnt; ++j) {
            uint32_t b = __builtin_amdgcn_readlane(my_b, j);
            uint32_t e = __builtin_amdgcn_readlane(my_e, j);
            for (uint32_t x = b + lane_id(); x < e; x += kWave) {
                uint32_t t = w.dst[x];
                int dt = w.deg[t];
                if (dt != -1) w.deg[t] = (typename N::deg_t)(dt - 1);
            }
            w.sync();
        }
    }
    for (uint32_t i = lane_id(); i < L; i += kWave) w.mark[w.list[i] >> 5] = 0u;
    w.sync();
}

// Leaf rule (src/helperFunctions.cuh:229-296): for every degree-1 vertex v
// mark its live neighbour u when deg(u) > 1, or deg(u) == 1 and u < v (so an
// isolated edge loses exactly one end).  Repeat until nothing is marked.
template <class N>
__device__ uint32_t rule_leaf(const N &w)
{
    uint32_t total = 0;
    for (;;) {
        uint32_t L = 0;
        for (uint32_t base = 0; base < w.n; base += kWave) {
            uint32_t v = base + lane_id();
            uint32_t cand = kNone;
            if (v < w.n && w.deg[v] == 1) {
                uint32_t e = w.src[v + 1];
                for (uint32_t x = w.src[v]; x < e; ++x) {
                    uint32_t u = w.dst[x];
                    int du = w.deg[u];
                    if (du > 0) {
                        if (du > 1 || u < v) cand = u;
                        break;
                    }
                }
            }
            L = append(w, L, cand, true);
        }
        if (L == 0) break;
        apply_list(w, L);
        total += L;
    }
    return total;
}

// Sorted-adjacency membership test (binarySearch, src/helperFunctions.cuh:21-36),
// half-open [lo, hi).
__device__ __forceinline__ bool adj_has(const uint32_t *dst, uint32_t lo, uint32_t hi, uint32_t x)
{
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        uint32_t y = dst[mid];
        if (y == x) return true;
        if (y < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return false;
}

// Degree-two triangle rule (src/helperFunctions.cuh:361-447): a degree-2
// vertex whose two live neighbours are adjacent puts both neighbours in the
// cover.  Skipped when a neighbour has degree 1, or degree 2 with a lower
// index, as in the reference.  All triangles of a pass are applied together.
template <class N>
__device__ uint32_t rule_triangle(const N &w)
{
    uint32_t total = 0;
    for (;;) {
        uint32_t L = 0;
        for (uint32_t base = 0; base < w.n; base += kWave) {
            uint32_t v = base + lane_id();
            uint32_t c1 = kNone, c2 = kNone;
            if (v < w.n && w.deg[v] == 2) {
                uint32_t n1 = kNone, n2 = kNone;
                bool keep = false;
                uint32_t e = w.src[v + 1];
                for (uint32_t x = w.src[v]; x < e; ++x) {
                    uint32_t u = w.dst[x];
                    int du = w.deg[u];
                    if (du > 0) {
                        if (du == 1 || (du == 2 && u < v)) {
                            keep = true;
                            break;
                        }
                        if (n1 == kNone) {
                            n1 = u;
                        } else {
                            n2 = u;
                            break;
                        }
                    }
                }
                if (!keep && n2 != kNone && adj_has(w.dst, w.src[n2], w.src[n2 + 1], n1)) {
                    c1 = n1;
                    c2 = n2;
                }
            }
            L = append(w, L, c1, true);
            L = append(w, L, c2, true);
        }
        if (L == 0) break;
        apply_list(w, L);
        total += L;
    }
    return total;
}

// High-degree rule (src/helperFunctions.cuh:299-358): v goes into the cover
// when deg(v) + d + (deleted so far in this call) >= T.
template <class N>
__device__ uint32_t rule_high(const N &w, uint32_t d, uint32_t T)
{
    uint32_t total = 0;
    for (;;) {
        uint32_t L = 0;
        for (uint32_t base = 0; base < w.n; base += kWave) {
            uint32_t v = base + lane_id();
            uint32_t cand = kNone;
            if (v < w.n) {
                int dv = w.deg[v];
                if (dv > 0 && (uint32_t)dv + d + total >= T) cand = v;
            }
            L = append(w, L, cand, false);
        }
        if (L == 0) break;
        apply_list(w, L);
        total += L;
    }
    return total;
}

// Rule fixpoint in the order of src/GlobalWorkList.cuh:107-130.  Returns the
// new number of vertices in the cover.
template <class N>
__device__ uint32_t reduce_node(const N &w, uint32_t d, uint32_t T)
{
    uint32_t it = 0, nl, nt = 0, nh = 0;
    do {
        nl = rule_leaf(w);
        d += nl;
        if (it == 0 || nl > 0 || nh > 0) {
            nt = rule_triangle(w);
            d += nt;
        } else {
            nt = 0;
        }
        if (it == 0 || nl > 0 || nt > 0) {
            nh = rule_high(w, d, T);
            d += nh;
        } else {
            nh = 0;
        }
        ++it;
    } while (nt > 0 || nh > 0);
    return d;
}

struct Summary {
    uint32_t edges;  // live edges (findNumOfEdges, src/helperFunctions.cuh:483-504)
    int32_t maxdeg;  // findMaxDegree (src/helperFunctions.cuh:450-481)
    uint32_t maxv;   // lowest index of maximum degree
};

template <class N>
__device__ Summary summarize(const N &w)
{
    int s = 0, md = 0;
    uint32_t mv = kNone;
    for (uint32_t base = 0; base < w.n; base += kWave) {
        uint32_t v = base + lane_id();
        if (v < w.n) {
            int dv = w.deg[v];
            if (dv > 0) s += dv;
            if (dv > md) {
                md = dv;
                mv = v;
            }
        }
    }
    Summary r;
    r.edges = (uint32_t)wave_sum(s) / 2u;
    r.maxdeg = wave_max(md);
    uint32_t cand = (md == r.maxdeg && r.maxdeg > 0) ? mv : kNone;
    r.maxv = wave_min_u32(cand);
    return r;
}

// Left child in place: v into the cover (deleteMaxDegreeVertex,
// src/helperFunctions.cuh:200-215).
template <class N>
__device__ void take_vertex(const N &w, uint32_t v)
{
    w.sync();
    if (lane_id() == 0) w.deg[v] = -1;
    w.sync();
    uint32_t e = w.src[v + 1];
    for (uint32_t x = w.src[v] + lane_id(); x < e; x += kWave) {
        uint32_t t = w.dst[x];
        int dt = w.deg[t];
        if (dt != -1) w.deg[t] = (typename N::deg_t)(dt - 1);
    }
    w.sync();
}

// Right child: every live neighbour of v into the cover
// (deleteNeighborsOfMaxDegreeVertex, src/helperFunctions.cuh:156-184), in
// rounds of at most lcap vertices.  Returns how many vertices were added
// (deg(v) before the call).
template <class N>
__device__ uint32_t take_neighbours(const N &w, uint32_t v)
{
    uint32_t total = 0;
    const uint32_t b = w.src[v], e = w.src[v + 1];
    for (;;) {
        uint32_t L = 0;
        for (uint32_t x0 = b; x0 < e && L < w.lcap; x0 += kWave) {
            uint32_t x = x0 + lane_id();
            uint32_t cand = kNone;
            if (x < e) {
                uint32_t t = w.dst[x];
                if (w.deg[t] != -1) cand = t;
            }
            L = append(w, L, cand, false);
        }
        if (L == 0) break;
        apply_list(w, L);
        total += L;
    }
    return total;
}

// Node storage <-> HBM slot copies (bytes % 16 == 0, both sides 16-aligned).
template <bool HBM>
__device__ __forceinline__ void copy_out(void *__restrict__ g, const void *s, uint32_t bytes)
{
    const int4 *s4 = reinterpret_cast<const int4 *>(s);
    int4 *g4 = reinterpret_cast<int4 *>(g);
    for (uint32_t i = lane_id(); i < bytes / 16; i += kWave) g4[i] = s4[i];
}
template <bool HBM>
__device__ __forceinline__ void copy_in(void *s, const void *__restrict__ g, uint32_t bytes)
{
    const int4 *g4 = reinterpret_cast<const int4 *>(g);
    int4 *s4 = reinterpret_cast<int4 *>(s);
    for (uint32_t i = lane_id(); i < bytes / 16; i += kWave) s4[i] = g4[i];
    nsync<HBM>();
}

// ---------------------------------------------------------------------------
// broker worklist (src/BWDWorkList.cuh) with 64-bit ring positions
// ---------------------------------------------------------------------------

// Writes one subtree (degrees at src, cover size d, pending branch vertex v
// of graph g) into the ring if the occupancy guard admits it.  The caller
// has already decided that a donation is wanted.
__device__ bool try_enqueue(const SolveArgs &a, const void *src, uint32_t g, uint32_t d, uint32_t v,
                            uint32_t slot_bytes)
{
    uint32_t ok = 0, slot = 0, ticket = 0;
    if (lane_id() == 0) {
        int c = fetch_add(&a.glob->count, 1);
        if (c < (int)a.qsize) {
            ok = 1;
            unsigned long long pos = fetch_add(&a.glob->tail, 1ull);
            slot = (uint32_t)(pos & (unsigned long long)(a.qsize - 1));
            ticket = (uint32_t)(2ull * (pos / a.qsize));
            uint32_t spins = 0;
            while (ld_acq(&a.tickets[slot]) != ticket) {
                backoff(spins < 3 ? spins : 3);
                ++spins;
                if (past_deadline(a)) break;
            }
            // the subtree being donated is one more outstanding unit of g
            fetch_add(&a.gstate[g].outstanding, 1u);
        } else {
            fetch_add(&a.glob->count, -1);
        }
    }
    ok = uni(ok);
    if (!ok) return false;
    slot = uni(slot);
    ticket = uni(ticket);
    fence_acq();
    copy_out<false>(static_cast<char *>(a.q_deg) + (size_t)slot * slot_bytes, src, slot_bytes);
    if (lane_id() == 0) a.q_meta[slot] = make_uint4(g, d, v, 0u);
    fence_rel();
    if (lane_id() == 0) {
        st_rel(&a.tickets[slot], ticket + 1u);
        fetch_add(&a.glob->counter, kEnqOne);
    }
    return true;
}

// Takes one item if the occupancy guard says one is there.  On success the
// degree array is in the worker's node storage and (g, d, v) are returned.
template <class N>
__device__ bool try_dequeue(const SolveArgs &a, const N &w, uint32_t &g, uint32_t &d, uint32_t &v, uint32_t slot_bytes)
{
    uint32_t ok = 0, slot = 0, ticket = 0;
    if (lane_id() == 0 && ld_rlx(&a.glob->count) > 0) {
        int c = fetch_add(&a.glob->count, -1);
        if (c > 0) {
            ok = 1;
            unsigned long long pos = fetch_add(&a.glob->head, 1ull);
            slot = (uint32_t)(pos & (unsigned long long)(a.qsize - 1));
            ticket = (uint32_t)(2ull * (pos / a.qsize) + 1ull);
            uint32_t spins = 0;
            while (ld_acq(&a.tickets[slot]) != ticket) {
                backoff(spins < 3 ? spins : 3);
                ++spins;
                if (past_deadline(a)) break;
            }
        } else {
            fetch_add(&a.glob->count, 1);
        }
    }
    ok = uni(ok);
    if (!ok) return false;
    slot = uni(slot);
    ticket = uni(ticket);
    fence_acq();
    copy_in<false>(w.deg, static_cast<const char *>(a.q_deg) + (size_t)slot * slot_bytes, slot_bytes);
    w.sync();
    uint4 meta = a.q_meta[slot];
    g = uni(meta.x);
    d = uni(meta.y);
    v = uni(meta.z);
    fence_rel();
    if (lane_id() == 0) {
        st_rel(&a.tickets[slot], ticket + 1u);
        // leave "waiting" and take the item off "enqueued" in one step
        fetch_add(&a.glob->counter, (unsigned long long)(0ull - (kEnqOne + kWaitOne)));
    }
    return true;
}

// ---------------------------------------------------------------------------
// per-graph bookkeeping
// ---------------------------------------------------------------------------

template <class N>
__device__ void write_cover(const SolveArgs &a, const N &w, const GDesc &gd)
{
    uint32_t *cw = a.covers + gd.off_cover;
    for (uint32_t base = 0; base < w.n; base += kWave) {
        uint32_t v = base + lane_id();
        uint64_t m = __ballot(v < w.n && w.deg[v] == -1);
        if (lane_id() == 0) {
            cw[base / 32] = (uint32_t)m;
            cw[base / 32 + 1] = (uint32_t)(m >> 32);
        }
    }
}

// A cover of size d was reached (maxDegree == 0).
template <class N>
__device__ void record_cover(const SolveArgs &a, const N &w, const GDesc &gd, uint32_t g, uint32_t d)
{
    GState *gs = &a.gstate[g];
    if (gd.instance == 0) {
        // MVC: atomicMin(minimum, d) (src/GlobalWorkList.cuh:150-155); the
        // improving worker also publishes the cover under a per-graph lock.
        int old = 0;
        if (lane_id() == 0) old = atomicMin(&gs->best, (int)d);
        old = (int)uni((uint32_t)old);
        if ((int)d < old) {
            if (lane_id() == 0) {
                while (atomicCAS(&gs->cover_lock, 0u, 1u) != 0u) {
                    backoff(2);
                    if (past_deadline(a)) break;
                }
            }
            fence_acq();
            int cur = (int)uni((uint32_t)ld_rlx(&gs->best));
            if (cur == (int)d) {
                write_cover(a, w, gd);
                if (lane_id() == 0) st_rlx(&gs->cover_size, (int32_t)d);
            }
            fence_rel();
            if (lane_id() == 0) st_rel(&gs->cover_lock, 0u);
        }
    } else {
        // PVC: atomicOr(kFound, 1) (src/GlobalWorkListParameterized.cuh:148-156);
        // the first finder publishes its cover and stops the graph.
        if (d <= gd.k) {
            uint32_t first = 0;
            if (lane_id() == 0) first = atomicCAS(&gs->found, 0u, 1u) == 0u;
            if (uni(first)) {
                write_cover(a, w, gd);
                if (lane_id() == 0) st_rlx(&gs->cover_size, (int32_t)d);
                fence_rel();
            }
            if (lane_id() == 0) st_rlx(&gs->stop, 1u);
        }
    }
}

__device__ __forceinline__ void flush_nodes(const SolveArgs &a, const GDesc &gd, uint32_t g, uint64_t &pending)
{
    if (pending == 0) return;
    if (lane_id() == 0) {
        GState *gs = &a.gstate[g];
        uint64_t before = fetch_add(&gs->nodes, (uint64_t)pending);
        if (gd.max_nodes && before + pending >= gd.max_nodes) {
            st_rlx(&gs->budget_hit, 1u);
            st_rlx(&gs->stop, 1u);
        }
    }
    pending = 0;
}

__device__ __forceinline__ void finish_unit(const SolveArgs &a, uint32_t g)
{
    if (lane_id() == 0) {
        uint32_t old = fetch_add(&a.gstate[g].outstanding, 0xFFFFFFFFu);
        if (old == 1u) st_rlx(&a.gstate[g].done, 1u);
    }
}

// Idle worker: take a queued item, else a new root, else wait; returns
// 0 = terminate, 1 = queued item (degrees already in node storage), 2 = root.
struct Acquired {
    uint32_t kind, g, d, v, root;
};

template <class N>
__device__ Acquired acquire(const SolveArgs &a, const N &w, uint32_t slot_bytes)
{
    Acquired r = {0u, kNone, 0u, kNone, 0u};
    if (lane_id() == 0) fetch_add(&a.glob->counter, kWaitOne);
    uint32_t spins = 0;
    for (;;) {
        if (a.version == 1) {
            uint32_t g, d, v;
            if (try_dequeue(a, w, g, d, v, slot_bytes)) {
                r.kind = 1;
                r.g = g;
                r.d = d;
                r.v = v;
                return r;
            }
        }
        if (uni(ld_rlx(&a.glob->next_root)) < a.roots) {
            uint32_t idx = kNone;
            if (lane_id() == 0) {
                fetch_add(&a.glob->counter, (unsigned long long)(0ull - kWaitOne));
                idx = fetch_add(&a.glob->next_root, 1u);
                if (idx >= a.roots) fetch_add(&a.glob->counter, kWaitOne);
            }
            idx = uni(idx);
            if (idx < a.roots) {
                r.kind = 2;
                r.root = idx;
                return r;
            }
        }
        unsigned long long c = 0;
        if (lane_id() == 0) c = ld_rlx(&a.glob->counter);
        uint32_t lo = uni((uint32_t)c), hi = uni((uint32_t)(c >> 32));
        uint32_t waiting = lo & 0xFFFFu, reg = lo >> 16;
        if (waiting == reg && hi == 0u && uni(ld_rlx(&a.glob->next_root)) >= a.roots) {
            if (lane_id() == 0) st_rlx(&a.glob->done_flag, 1u);
            return r;
        }
        if (uni(ld_rlx(&a.glob->done_flag)) || uni(ld_rlx(&a.glob->cancel))) return r;
        uint32_t late = 0;
        if (lane_id() == 0) late = past_deadline(a);
        if (uni(late)) return r;
        backoff(++spins);
    }
}

__device__ __forceinline__ long long sq64(int x) { return (long long)x * (long long)x; }

// ---------------------------------------------------------------------------
// the persistent solve kernel
// ---------------------------------------------------------------------------

constexpr uint32_t kFlushEvery = 1024;

template <typename DegT, bool HBM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void solve_kernel(SolveArgs a)
{
    extern __shared__ int4 smem4[];
    const uint32_t wib = threadIdx.x / kWave;
    const uint32_t wid = blockIdx.x * (blockDim.x / kWave) + wib;
    if (wid >= a.workers) return;

    // node storage: deg [npad DegT] | list [lcap u32] | mark [mwords u32]
    char *store = HBM ? a.hbm_nodes + (size_t)wid * a.worker_bytes
                      : reinterpret_cast<char *>(smem4) + (size_t)wib * a.worker_bytes;
    const uint32_t slot_bytes = a.npad * (uint32_t)sizeof(DegT);
    Node<DegT, HBM> w;
    w.deg = reinterpret_cast<DegT *>(store);
    w.list = reinterpret_cast<uint32_t *>(store + slot_bytes);
    w.mark = w.list + a.lcap;
    w.src = nullptr;
    w.dst = nullptr;
    w.n = 0;
    w.lcap = a.lcap;
    for (uint32_t i = lane_id(); i < a.mwords; i += kWave) w.mark[i] = 0u;
    w.sync();

    char *stk = static_cast<char *>(a.stk) + (size_t)wid * a.depth * slot_bytes;
    uint2 *stk_meta = a.stk_meta + (size_t)wid * a.depth;

    if (lane_id() == 0) fetch_add(&a.glob->counter, kRegOne);

    uint32_t g = kNone;
    GDesc gd = {};
    int top = -1, bot = 0;  // local stack entries bot..top (bot moves up on donation)
    bool have = false;
    uint32_t d = 0;
    uint32_t path = 0, descend = 0;
    uint64_t pending = 0, total_nodes = 0, idle_ticks = 0, donations = 0, steals = 0;

    for (;;) {
        if (!have) {
            if (top < bot) {
                top = -1;
                bot = 0;
            }
            if (top >= 0) {
                // pop (popStack, src/stack.cuh:10-20), then build the stored
                // right child
                copy_in<HBM>(w.deg, stk + (size_t)top * slot_bytes, slot_bytes);
                uint2 m = stk_meta[top];
                d = uni(m.x);
                uint32_t v = uni(m.y);
                --top;
                d += take_neighbours(w, v);
                have = true;
            } else {
                if (g != kNone) {
                    total_nodes += pending;
                    flush_nodes(a, gd, g, pending);
                    finish_unit(a, g);
                    g = kNone;
                }
                if (lane_id() == 0 && ld_rlx(&a.glob->cancel)) st_rlx(&a.glob->done_flag, 1u);
                const uint64_t t_idle = wall_clock64();
                Acquired q = acquire(a, w, slot_bytes);
                idle_ticks += wall_clock64() - t_idle;
                if (q.kind == 0) break;
                steals += q.kind == 1;
                if (q.kind == 1) {
                    g = q.g;
                } else {
                    g = a.version == 1 ? q.root : (q.root >> a.start_depth);
                }
                gd = a.gdesc[g];
                w.n = gd.n;
                w.src = a.gbuf + gd.off_src;
                w.dst = a.gbuf + gd.off_dst;
                if (q.kind == 1) {
                    d = q.d;
                    if (q.v != kNone) d += take_neighbours(w, q.v);
                    descend = 0;
                } else {
                    const int32_t *deg0 = reinterpret_cast<const int32_t *>(a.gbuf + gd.off_deg);
                    for (uint32_t i = lane_id(); i < a.npad; i += kWave)
                        w.deg[i] = (DegT)(i < gd.n ? deg0[i] : -1);
                    w.sync();
                    d = 0;
                    if (a.version == 1) {
                        descend = 0;
                    } else {
                        path = q.root & ((1u << a.start_depth) - 1u);
                        descend = a.start_depth;
                    }
                }
                have = true;
            }
        }

        GState *gs = &a.gstate[g];
        if (uni(ld_rlx(&gs->stop))) {
            // graph finished by someone else (PVC found, budget): drop local work
            top = -1;
            bot = 0;
            have = false;
            continue;
        }
        if (++pending >= kFlushEvery) {
            total_nodes += pending;
            flush_nodes(a, gd, g, pending);
            uint32_t stop_all = 0;
            if (lane_id() == 0) stop_all = ld_rlx(&a.glob->cancel) || past_deadline(a);
            if (uni(stop_all)) {
                top = -1;
                bot = 0;
                have = false;
                continue;
            }
        }

        const bool mvc = gd.instance == 0;
        uint32_t T = mvc ? (uint32_t)uni((uint32_t)ld_rlx(&gs->best)) : gd.k + 1u;
        d = reduce_node(w, d, T);
        Summary s = summarize(w);

        bool prune;
        if (mvc) {
            int best = (int)uni((uint32_t)ld_rlx(&gs->best));
            prune = (int)d >= best || (long long)s.edges >= sq64(best - (int)d - 1) + 1;
        } else {
            prune = d > gd.k || (long long)s.edges >= sq64((int)gd.k - (int)d) + 1;
        }
        if (prune) {
            have = false;
            descend = 0;
            continue;
        }
        if (s.maxdeg == 0) {
            record_cover(a, w, gd, g, d);
            have = false;
            descend = 0;
            continue;
        }
        const uint32_t v = s.maxv;
        if (descend > 0) {
            // static split: follow this worker's path bit (src/LocalStacks.cuh:121-157)
            uint32_t bit = (path >> (a.start_depth - descend)) & 1u;
            --descend;
            if (bit == 0u) {
                take_vertex(w, v);
                d += 1;
            } else {
                d += take_neighbours(w, v);
            }
            continue;
        }
        // Branch: the right child (N(v) into the cover) is stored as its
        // parent state plus v and built when it is popped.
        if (top + 1 >= (int)a.depth) {
            if (lane_id() == 0) st_rlx(&a.glob->error, kErrStackOverflow);
            have = false;
            continue;
        }
        ++top;
        copy_out<HBM>(stk + (size_t)top * slot_bytes, w.deg, slot_bytes);
        if (lane_id() == 0) stk_meta[top] = make_uint2(d, v);
        if (a.version == 1) {
            // Donate while workers wait and the worklist holds fewer items
            // than there are waiting workers (and fewer than the -t
            // threshold).  The OLDEST local entry goes: it is the shallowest
            // and so the largest subtree this worker holds.
            unsigned long long c = 0;
            if (lane_id() == 0) c = ld_rlx(&a.glob->counter);
            int waiting = (int)(uni((uint32_t)c) & 0xFFFFu);
            int enq = (int)uni((uint32_t)(c >> 32));
            if (waiting > 0 && enq < waiting && enq < (int)a.qthreshold && top >= bot) {
                uint2 m = stk_meta[bot];
                if (try_enqueue(a, stk + (size_t)bot * slot_bytes, g, uni(m.x), uni(m.y), slot_bytes)) {
                    ++bot;
                    ++donations;
                }
            }
        }
        take_vertex(w, v);
        d += 1;
    }
    if (lane_id() == 0) {
        WorkerStats ws;
        ws.nodes = total_nodes;
        ws.idle_ticks = idle_ticks;
        ws.donations = donations;
        ws.steals = steals;
        a.wstats[wid] = ws;
    }
}

// Per-launch initialisation with plain stores (no memset nodes, so the whole
// solve can be captured in a HIP graph).  units = outstanding subtrees per
// graph at start: 1 root (hybrid) or 2^start_depth paths (static split).
__global__ void init_kernel(SolveArgs a, uint32_t units)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t g = tid; g < a.num_graphs; g += stride) {
        GState s = {};
        s.best = (int32_t)a.gdesc[g].U;
        s.outstanding = units;
        s.cover_size = -1;
        a.gstate[g] = s;
    }
    for (uint32_t i = tid; i < a.qsize; i += stride) a.tickets[i] = 0u;
    for (uint32_t i = tid; i < a.workers; i += stride) a.wstats[i] = WorkerStats{};
    if (tid == 0) {
        Global z = {};
        z.t0 = wall_clock64();
        *a.glob = z;
    }
}

template __global__ void solve_kernel<int16_t, false>(SolveArgs);
template __global__ void solve_kernel<int32_t, false>(SolveArgs);
template __global__ void solve_kernel<int32_t, true>(SolveArgs);

}  // namespace vcg

// ===========================================================================
// host launcher and C ABI (include/vcgpu.h)
// ===========================================================================

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "vcgpu.h"

using vcg::GDesc;
using vcg::GState;
using vcg::Global;
using vcg::SolveArgs;

struct vc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cu_count = 0;
    int wall_khz = 100000;  // device wall-clock rate
    size_t total_mem = 0;
    size_t lds_per_cu = 0;
    // launch scratch, grown on demand and reused across solves
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
};

struct vc_graphs {
    int count = 0;
    uint32_t *d_words = nullptr;  // CSR arrays of every graph
    GDesc *d_desc = nullptr;      // per-graph descriptors (job fields rewritten per solve)
    uint32_t *d_covers = nullptr; // cover bitmaps
    GState *d_state = nullptr;
    std::vector<GDesc> h_desc;
    uint32_t nmax = 0;
    uint32_t cover_words = 0;
};

namespace {

inline bool hip_ok(hipError_t e) { return e == hipSuccess; }

inline uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }

inline uint32_t next_pow2(uint32_t x)
{
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Validation of an uploaded CSR: what createCSRGraphFromFile produces and
// check_graph accepts (src/auxFunctions.cpp:129-209, 308-354), plus the
// sortedness the triangle rule's binary search relies on.
bool csr_is_valid(const vc_csr &g)
{
    const uint32_t n = g.vertexNum, m = g.edgeNum;
    if (n > 0 && (!g.srcPtr || !g.degree)) return false;
    if (m > 0 && !g.dst) return false;
    if (n == 0) return m == 0;
    if (g.srcPtr[0] != 0 || g.srcPtr[n] != 2ull * m) return false;
    for (uint32_t v = 0; v < n; ++v) {
        uint32_t b = g.srcPtr[v], e = g.srcPtr[v + 1];
        if (e < b || e > 2ull * m) return false;
        if ((int64_t)g.degree[v] != (int64_t)(e - b)) return false;
        for (uint32_t x = b; x < e; ++x) {
            uint32_t u = g.dst[x];
            if (u >= n || u == v) return false;
            if (x > b && g.dst[x - 1] >= u) return false;  // unsorted or duplicate
        }
    }
    return true;
}

// Kernel variants: degree type x node storage.
enum Variant { kLds16 = 0, kLds32 = 1, kHbm32 = 2 };

const void *kernel_of(int variant)
{
    switch (variant) {
    case kLds16: return reinterpret_cast<const void *>(vcg::solve_kernel<int16_t, false>);
    case kLds32: return reinterpret_cast<const void *>(vcg::solve_kernel<int32_t, false>);
    default: return reinterpret_cast<const void *>(vcg::solve_kernel<int32_t, true>);
    }
}

struct Plan {
    int variant;
    uint32_t npad, lcap, mwords, worker_bytes, deg_bytes, wpb, block_dim, blocks_per_cu, num_blocks, workers;
    uint32_t depth, qsize, qthreshold, roots, units;
    size_t smem;
    size_t off_glob, off_stk, off_meta, off_qdeg, off_qmeta, off_tickets, off_wnodes, off_hbm, total;
};

// Per-worker node storage for a degree type: deg [npad] | list [lcap] | mark [mwords].
void size_storage(Plan &pl, uint32_t nmax, uint32_t deg_bytes)
{
    pl.deg_bytes = deg_bytes;
    pl.npad = round_up(nmax, 64u / deg_bytes);
    pl.lcap = std::min<uint32_t>(round_up(nmax, 4u), 1024u);
    pl.mwords = round_up((nmax + 31u) / 32u, 4u);
    pl.worker_bytes = round_up(pl.npad * deg_bytes + 4u * pl.lcap + 4u * pl.mwords, 64u);
}

vc_status make_plan(vc_ctx *ctx, const vc_graphs *set, const vc_job *jobs, const vc_params *p, Plan &pl)
{
    const uint32_t version = p->version;
    if (version != VC_HYBRID && version != VC_STACK_ONLY) return VC_EINVAL;
    if (p->use_global_memory < -1 || p->use_global_memory > 1) return VC_EINVAL;
    uint32_t block_dim = p->block_dim ? p->block_dim : 256u;
    if (block_dim % vcg::kWave != 0 || block_dim > 256u) return VC_EINVAL;
    if (version == VC_STACK_ONLY && p->start_depth > 20u) return VC_EINVAL;
    const uint32_t nmax = std::max<uint32_t>(set->nmax, 1u);
    const size_t lds_limit = ctx->lds_per_cu ? ctx->lds_per_cu : 160u * 1024u;

    // Degree arrays in LDS when at least one worker fits per CU (int16 while
    // every degree fits), otherwise in HBM (-g 1 forces HBM, -g 0 forbids it).
    pl.variant = nmax < 32768u ? kLds16 : kLds32;
    size_storage(pl, nmax, pl.variant == kLds16 ? 2u : 4u);
    bool hbm = p->use_global_memory == 1 || pl.worker_bytes > lds_limit;
    if (hbm && p->use_global_memory == 0) return VC_ENOTSUP;
    if (hbm) {
        pl.variant = kHbm32;
        size_storage(pl, nmax, 4u);
    }
    // a worker is one wave: shrink the workgroup when LDS would not hold it
    uint32_t wpb = block_dim / vcg::kWave;
    if (!hbm && !p->block_dim)
        while (wpb > 1 && (size_t)wpb * pl.worker_bytes > lds_limit) wpb /= 2;
    pl.wpb = wpb;
    pl.block_dim = wpb * vcg::kWave;
    pl.smem = hbm ? 0 : (size_t)pl.wpb * pl.worker_bytes;
    if (pl.smem > lds_limit) return VC_ENOTSUP;

    int bpc = 0;
    if (!hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kernel_of(pl.variant), (int)pl.block_dim, pl.smem)))
        return VC_EDEVICE;
    if (bpc < 1) return VC_ENOTSUP;
    pl.blocks_per_cu = (uint32_t)bpc;
    pl.num_blocks = p->num_blocks ? p->num_blocks : (uint32_t)bpc * (uint32_t)ctx->cu_count;

    // Stack depth: the stack never holds more entries than the cover size of
    // the current node, which stays below max(U, k+1) (prune at d >= U for
    // MVC, d > k for PVC), and never more than n.
    uint32_t need = 1;
    for (int i = 0; i < set->count; ++i) {
        uint32_t lim = jobs[i].instance == VC_MVC ? jobs[i].upper_bound : std::max(jobs[i].upper_bound, jobs[i].k + 1u);
        lim = std::min(lim, set->h_desc[i].n);
        need = std::max(need, lim);
    }
    pl.depth = need + 1u;

    // Worklist: size and donation threshold (-q, -t).
    pl.workers = pl.num_blocks * pl.wpb;
    if (p->list_size) {
        if ((p->list_size & (p->list_size - 1u)) != 0u) return VC_EINVAL;
        pl.qsize = p->list_size;
    } else {
        pl.qsize = next_pow2(std::max<uint32_t>(1024u, 2u * pl.workers));
    }
    if (p->list_threshold > 0.f) {
        pl.qthreshold = (uint32_t)(p->list_threshold * (float)pl.qsize);
    } else {
        pl.qthreshold = pl.qsize;  // demand-driven: donate only while workers wait
    }

    // Fit stacks + worklist into device memory; shrink the grid if needed
    // (the reference halves the worklist instead, src/auxFunctions.cpp:507-512).
    const size_t slot = (size_t)pl.npad * pl.deg_bytes;
    size_t budget = ctx->total_mem / 10 * 7;
    for (;;) {
        pl.workers = pl.num_blocks * pl.wpb;
        size_t off = 0;
        auto take = [&](size_t bytes) {
            size_t o = off;
            off += (bytes + 255) / 256 * 256;
            return o;
        };
        pl.off_glob = take(sizeof(Global));
        pl.off_stk = take((size_t)pl.workers * pl.depth * slot);
        pl.off_meta = take((size_t)pl.workers * pl.depth * sizeof(uint2));
        pl.off_qdeg = take((size_t)pl.qsize * slot);
        pl.off_qmeta = take((size_t)pl.qsize * sizeof(uint4));
        pl.off_tickets = take((size_t)pl.qsize * sizeof(uint32_t));
        pl.off_wnodes = take((size_t)pl.workers * sizeof(vcg::WorkerStats));
        pl.off_hbm = take(pl.variant == kHbm32 ? (size_t)pl.workers * pl.worker_bytes : 0);
        pl.total = off;
        if (pl.total <= budget || pl.num_blocks <= 1) break;
        pl.num_blocks = std::max<uint32_t>(1u, pl.num_blocks / 2u);
    }
    if (pl.total > budget) return VC_ENOMEM;

    pl.units = version == VC_HYBRID ? 1u : (1u << p->start_depth);
    pl.roots = (uint32_t)set->count * pl.units;
    return VC_OK;
}

}  // namespace

extern "C" {

const char *vc_status_string(vc_status s)
{
    switch (s) {
    case VC_OK: return "ok";
    case VC_EINVAL: return "invalid argument";
    case VC_ENOMEM: return "out of device memory";
    case VC_EDEVICE: return "device error";
    case VC_ETIMEOUT: return "node budget or spin bound exhausted";
    case VC_ECANCELLED: return "cancelled";
    case VC_ENOTSUP: return "not supported";
    case VC_EBADGRAPH: return "invalid graph (duplicate edge, self-loop, unsorted or out-of-range adjacency)";
    }
    return "unknown status";
}

vc_status vc_ctx_create(int device, vc_ctx **out)
{
    if (!out) return VC_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (!hip_ok(hipGetDeviceCount(&ndev)) || device < 0 || device >= ndev) return VC_EDEVICE;
    if (!hip_ok(hipSetDevice(device))) return VC_EDEVICE;
    vc_ctx *c = new vc_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (!hip_ok(hipGetDeviceProperties(&prop, device))) {
        delete c;
        return VC_EDEVICE;
    }
    c->cu_count = prop.multiProcessorCount;
    c->total_mem = prop.totalGlobalMem;
    c->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
    int khz = 0;
    if (hip_ok(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device)) && khz > 0) c->wall_khz = khz;
    if (!hip_ok(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) || !hip_ok(hipEventCreate(&c->ev0)) ||
        !hip_ok(hipEventCreate(&c->ev1))) {
        delete c;
        return VC_EDEVICE;
    }
    *out = c;
    return VC_OK;
}

vc_status vc_ctx_destroy(vc_ctx *c)
{
    if (!c) return VC_EINVAL;
    hipSetDevice(c->device);
    if (c->scratch) hipFree(c->scratch);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return VC_OK;
}

vc_status vc_graphs_upload(vc_ctx *ctx, int count, const vc_csr *graphs, vc_graphs **out)
{
    if (!ctx || !out || count < 1 || !graphs) return VC_EINVAL;
    *out = nullptr;
    hipSetDevice(ctx->device);
    for (int i = 0; i < count; ++i)
        if (!csr_is_valid(graphs[i])) return VC_EBADGRAPH;

    vc_graphs *s = new vc_graphs();
    s->count = count;
    s->h_desc.resize(count);
    size_t words = 0;
    uint32_t cover_words = 0;
    for (int i = 0; i < count; ++i) {
        const vc_csr &g = graphs[i];
        GDesc &d = s->h_desc[i];
        std::memset(&d, 0, sizeof(d));
        d.n = g.vertexNum;
        d.m = g.edgeNum;
        d.off_src = (uint32_t)words;
        words += g.vertexNum + 1u;
        d.off_dst = (uint32_t)words;
        words += 2ull * g.edgeNum;
        d.off_deg = (uint32_t)words;
        words += g.vertexNum;
        words = (words + 3) / 4 * 4;
        d.off_cover = cover_words;
        cover_words += round_up((g.vertexNum + 31u) / 32u, 2u);
        s->nmax = std::max(s->nmax, g.vertexNum);
        if (words >= 0xFFFFFFF0ull) {
            delete s;
            return VC_ENOMEM;
        }
    }
    s->cover_words = std::max<uint32_t>(cover_words, 2u);
    std::vector<uint32_t> host(std::max<size_t>(words, 4), 0u);
    for (int i = 0; i < count; ++i) {
        const vc_csr &g = graphs[i];
        const GDesc &d = s->h_desc[i];
        if (g.vertexNum) {
            std::memcpy(&host[d.off_src], g.srcPtr, sizeof(uint32_t) * (g.vertexNum + 1u));
            std::memcpy(&host[d.off_deg], g.degree, sizeof(int32_t) * g.vertexNum);
        } else {
            host[d.off_src] = 0u;
        }
        if (g.edgeNum) std::memcpy(&host[d.off_dst], g.dst, sizeof(uint32_t) * 2ull * g.edgeNum);
    }
    bool ok = hip_ok(hipMalloc(&s->d_words, host.size() * sizeof(uint32_t))) &&
              hip_ok(hipMalloc(&s->d_desc, sizeof(GDesc) * count)) &&
              hip_ok(hipMalloc(&s->d_covers, sizeof(uint32_t) * s->cover_words)) &&
              hip_ok(hipMalloc(&s->d_state, sizeof(GState) * count));
    if (ok)
        ok = hip_ok(hipMemcpy(s->d_words, host.data(), host.size() * sizeof(uint32_t), hipMemcpyHostToDevice)) &&
             hip_ok(hipMemcpy(s->d_desc, s->h_desc.data(), sizeof(GDesc) * count, hipMemcpyHostToDevice));
    if (!ok) {
        vc_graphs_free(ctx, s);
        return VC_ENOMEM;
    }
    *out = s;
    return VC_OK;
}

vc_status vc_graphs_free(vc_ctx *ctx, vc_graphs *s)
{
    if (!s) return VC_EINVAL;
    if (ctx) hipSetDevice(ctx->device);
    if (s->d_words) hipFree(s->d_words);
    if (s->d_desc) hipFree(s->d_desc);
    if (s->d_covers) hipFree(s->d_covers);
    if (s->d_state) hipFree(s->d_state);
    delete s;
    return VC_OK;
}

vc_status vc_occupancy(vc_ctx *ctx, const vc_graphs *set, const vc_params *params, int *blocks_per_cu, int *block_dim,
                       int *use_global_memory)
{
    if (!ctx || !set || !params) return VC_EINVAL;
    hipSetDevice(ctx->device);
    std::vector<vc_job> jobs(set->count);
    for (int i = 0; i < set->count; ++i) {
        jobs[i] = vc_job{VC_MVC, 0u, set->h_desc[i].n, 0u, 0ull};
    }
    Plan pl;
    vc_status st = make_plan(ctx, set, jobs.data(), params, pl);
    if (st != VC_OK) return st;
    if (blocks_per_cu) *blocks_per_cu = (int)pl.blocks_per_cu;
    if (block_dim) *block_dim = (int)pl.block_dim;
    if (use_global_memory) *use_global_memory = pl.variant == kHbm32;
    return VC_OK;
}

vc_status vc_solve(vc_ctx *ctx, const vc_graphs *set_c, const vc_job *jobs, const vc_params *params,
                   vc_result *results, uint8_t *const *covers, vc_launch_info *info)
{
    if (!ctx || !set_c || !jobs || !params || !results) return VC_EINVAL;
    vc_graphs *set = const_cast<vc_graphs *>(set_c);
    hipSetDevice(ctx->device);
    for (int i = 0; i < set->count; ++i) {
        const vc_job &j = jobs[i];
        if (j.instance != VC_MVC && j.instance != VC_PVC) return VC_EINVAL;
        if (j.upper_bound > set->h_desc[i].n) return VC_EINVAL;
    }
    Plan pl;
    vc_status st = make_plan(ctx, set, jobs, params, pl);
    if (st != VC_OK) return st;

    // job fields of the descriptors
    bool changed = false;
    for (int i = 0; i < set->count; ++i) {
        GDesc &d = set->h_desc[i];
        const vc_job &j = jobs[i];
        if (d.instance != j.instance || d.k != j.k || d.U != j.upper_bound || d.max_nodes != j.max_nodes) {
            d.instance = j.instance;
            d.k = j.k;
            d.U = j.upper_bound;
            d.max_nodes = j.max_nodes;
            changed = true;
        }
    }
    if (changed && !hip_ok(hipMemcpyAsync(set->d_desc, set->h_desc.data(), sizeof(GDesc) * set->count,
                                          hipMemcpyHostToDevice, ctx->stream)))
        return VC_EDEVICE;

    if (ctx->scratch_bytes < pl.total) {
        if (ctx->scratch) hipFree(ctx->scratch);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        if (!hip_ok(hipMalloc(&ctx->scratch, pl.total))) return VC_ENOMEM;
        ctx->scratch_bytes = pl.total;
    }
    char *base = static_cast<char *>(ctx->scratch);
    SolveArgs a;
    std::memset(&a, 0, sizeof(a));
    a.gbuf = set->d_words;
    a.gdesc = set->d_desc;
    a.gstate = set->d_state;
    a.covers = set->d_covers;
    a.glob = reinterpret_cast<Global *>(base + pl.off_glob);
    a.stk = base + pl.off_stk;
    a.stk_meta = reinterpret_cast<uint2 *>(base + pl.off_meta);
    a.q_deg = base + pl.off_qdeg;
    a.q_meta = reinterpret_cast<uint4 *>(base + pl.off_qmeta);
    a.tickets = reinterpret_cast<uint32_t *>(base + pl.off_tickets);
    a.wstats = reinterpret_cast<vcg::WorkerStats *>(base + pl.off_wnodes);
    a.num_graphs = (uint32_t)set->count;
    a.npad = pl.npad;
    a.lcap = pl.lcap;
    a.mwords = pl.mwords;
    a.worker_bytes = pl.worker_bytes;
    a.depth = pl.depth;
    a.qsize = pl.qsize;
    a.qthreshold = pl.qthreshold;
    a.workers = pl.workers;
    a.version = params->version;
    a.start_depth = params->version == VC_STACK_ONLY ? params->start_depth : 0u;
    a.roots = pl.roots;
    {
        // deadline on the device wall clock (every wait loop checks it)
        uint64_t ms = params->time_limit_ms ? params->time_limit_ms : 3600ull * 1000ull;
        a.limit_ticks = ms * (uint64_t)ctx->wall_khz;
    }
    a.hbm_nodes = pl.variant == kHbm32 ? base + pl.off_hbm : nullptr;

    hipLaunchKernelGGL(vcg::init_kernel, dim3(256), dim3(256), 0, ctx->stream, a, pl.units);
    if (!hip_ok(hipGetLastError())) return VC_EDEVICE;
    hipEventRecord(ctx->ev0, ctx->stream);
    switch (pl.variant) {
    case kLds16:
        hipLaunchKernelGGL((vcg::solve_kernel<int16_t, false>), dim3(pl.num_blocks), dim3(pl.block_dim), pl.smem,
                           ctx->stream, a);
        break;
    case kLds32:
        hipLaunchKernelGGL((vcg::solve_kernel<int32_t, false>), dim3(pl.num_blocks), dim3(pl.block_dim), pl.smem,
                           ctx->stream, a);
        break;
    default:
        hipLaunchKernelGGL((vcg::solve_kernel<int32_t, true>), dim3(pl.num_blocks), dim3(pl.block_dim), 0,
                           ctx->stream, a);
        break;
    }
    if (!hip_ok(hipGetLastError())) return VC_EDEVICE;
    hipEventRecord(ctx->ev1, ctx->stream);

    std::vector<GState> hs(set->count);
    Global hg;
    std::vector<vcg::WorkerStats> wn(pl.workers);
    bool ok = hip_ok(hipMemcpyAsync(hs.data(), set->d_state, sizeof(GState) * set->count, hipMemcpyDeviceToHost,
                                    ctx->stream)) &&
              hip_ok(hipMemcpyAsync(&hg, a.glob, sizeof(Global), hipMemcpyDeviceToHost, ctx->stream)) &&
              hip_ok(hipMemcpyAsync(wn.data(), a.wstats, sizeof(vcg::WorkerStats) * pl.workers, hipMemcpyDeviceToHost,
                                    ctx->stream));
    std::vector<uint32_t> bitmaps;
    bool want_covers = false;
    if (covers)
        for (int i = 0; i < set->count; ++i) want_covers |= covers[i] != nullptr;
    if (ok && want_covers) {
        bitmaps.resize(set->cover_words);
        ok = hip_ok(hipMemcpyAsync(bitmaps.data(), set->d_covers, sizeof(uint32_t) * set->cover_words,
                                   hipMemcpyDeviceToHost, ctx->stream));
    }
    if (!ok || !hip_ok(hipStreamSynchronize(ctx->stream))) return VC_EDEVICE;
    float ms = 0.f;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);

    uint64_t total_nodes = 0, idle = 0, donations = 0, steals = 0;
    for (const vcg::WorkerStats &x : wn) {
        total_nodes += x.nodes;
        idle += x.idle_ticks;
        donations += x.donations;
        steals += x.steals;
    }
    vc_status overall = hg.error ? VC_EDEVICE : (hg.timed_out ? VC_ETIMEOUT : VC_OK);
    for (int i = 0; i < set->count; ++i) {
        const GState &s = hs[i];
        const vc_job &j = jobs[i];
        vc_result &r = results[i];
        std::memset(&r, 0, sizeof(r));
        r.tree_nodes = s.nodes;
        bool finished = s.done != 0u && !hg.error;
        if (j.instance == VC_MVC) {
            r.minimum = (uint32_t)s.best;
            r.solved = finished && !s.budget_hit;
            r.cover_from_gpu = s.cover_size >= 0 && s.cover_size == s.best;
        } else {
            r.k_found = s.found;
            r.minimum = s.found ? (uint32_t)s.cover_size : 0u;
            r.solved = finished && (s.found || !s.budget_hit);
            r.cover_from_gpu = s.found != 0u;
        }
        r.status = hg.error ? VC_EDEVICE : (r.solved ? VC_OK : (hg.cancel && !hg.timed_out ? VC_ECANCELLED : VC_ETIMEOUT));
        if (covers && covers[i] && r.cover_from_gpu) {
            const uint32_t *bm = &bitmaps[set->h_desc[i].off_cover];
            for (uint32_t v = 0; v < set->h_desc[i].n; ++v) covers[i][v] = (bm[v >> 5] >> (v & 31u)) & 1u;
        }
    }
    if (info) {
        info->num_blocks = pl.num_blocks;
        info->block_dim = pl.block_dim;
        info->workers = pl.workers;
        info->blocks_per_cu = pl.blocks_per_cu;
        info->use_global_memory = pl.variant == kHbm32;
        info->list_size = pl.qsize;
        info->list_threshold = pl.qthreshold;
        info->stack_depth = pl.depth;
        info->kernel_ms = ms;
        info->cu_count = (uint32_t)ctx->cu_count;
        info->tree_nodes = total_nodes;
        info->donations = donations;
        info->steals = steals;
        // share of worker time spent waiting for work
        double span = (double)ms * 1e-3 * (double)ctx->wall_khz * 1e3 * (double)pl.workers;
        info->idle_fraction = span > 0 ? (float)std::min(1.0, (double)idle / span) : 0.f;
    }
    return overall;
}

}  // extern "C"
