// vc_kernels.hip -- gfx950 persistent branch-and-reduce kernel for minimum
// (MVC) and parameterized (PVC) vertex cover.
//
// One wave64 is one search worker.  A workgroup holds several independent
// workers; no workgroup barrier is ever used.  Each worker keeps the degree
// array of its current tree node (negative = vertex in the cover) in LDS -- or, for
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

#include "vcgpu.h"
#include "vc_bound.hpp"
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

// Inclusive prefix sum over the 64 lanes: DPP row_shr 1/2/3 on the input,
// row_shr 4/8 with bank masks inside each row of 16, then row_bcast 15/31
// to carry across rows (the GCN/CDNA scan sequence).
__device__ __forceinline__ int wave_incl_scan(int x)
{
    int v = x;
    v += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, x, 0x113, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xE, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xC, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, true);
    return v;
}

// Inclusive prefix maximum over the 64 lanes of non-negative values: DPP
// row_shr 1/2/4/8 (lanes shifted in from outside the row read 0), then
// row_bcast 15/31 across rows.
__device__ __forceinline__ int wave_incl_max(int x)
{
    int v = x;
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, true));
    return v;
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

// Worklist hand-offs without agent-scope release/acquire.  On gfx950 every
// XCD has its own L2, so an agent-scope release writes back the XCD's whole
// L2 (buffer_wbl2 sc1) and an agent-scope acquire invalidates it
// (buffer_inv sc1) -- and the next loads of every wave on that XCD, of
// adjacency lists and bitmaps too, go to HBM again.  Worklist slots are
// instead written and read only with agent-coherent accesses (sc1, what
// relaxed agent-scope atomics compile to), so a hand-off only has to order
// them: the producer waits until the slot's stores are complete
// (vmcnt(0)) before it stores the ticket, and the consumer issues the
// slot's loads after it has seen the ticket; the consumer waits for its
// loads the same way before it frees the slot.  The workgroup-scope fences
// only keep the compiler from moving memory accesses across these points.
__device__ __forceinline__ void handoff_publish()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0x3f70);  // vmcnt(0): this wave's loads and stores are complete
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
}
__device__ __forceinline__ void handoff_observe() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); }

// Sleeps 2^min(step, cap) x 512 cycles.
__device__ __forceinline__ void backoff(uint32_t step, uint32_t cap = 6u)
{
    uint32_t reps = 1u << (step < cap ? step : cap);
    for (uint32_t i = 0; i < reps; ++i) __builtin_amdgcn_s_sleep(8);
}

// Sleep of an idle worker between polls of the worklist.  Every poll reads
// launch-wide words that all workers share, so with thousands of idle
// workers the polls themselves slow down the few busy ones (a search of
// 5,000 tree nodes took 4.2 ms on 7,168 workers and 0.57 ms on 512).  The
// interval therefore grows with the number of waiting workers -- about one
// s_sleep(8) unit (~0.2 us) per 16 of them, i.e. a fixed total poll rate --
// with a jitter from the clock's low bits so that workers that went idle
// together do not poll together; kMaxIdleReps bounds how long a worker
// takes to see the end of the launch.
constexpr uint32_t kMaxIdleReps = 256u;
__device__ __forceinline__ void backoff_idle(uint32_t step, uint32_t cap, uint32_t waiting)
{
    uint32_t reps = 1u << (step < cap ? step : cap);
    uint32_t share = waiting >> 4;
    if (share > kMaxIdleReps) share = kMaxIdleReps;
    if (share > reps) {
        const uint32_t r = (uint32_t)wall_clock64() & 0xFFu;  // uniform in [share/2, 3 share/2)
        reps = share / 2u + ((share * r) >> 8);
    }
    for (uint32_t i = 0; i < reps; ++i) __builtin_amdgcn_s_sleep(8);
}

// True once the launch's time limit has passed; the first lane to see it
// raises cancel so every worker drains and exits.  Every wait loop of the
// kernel calls this, so no wave can spin forever.  The elapsed time is
// signed: the wall clocks of different XCDs differ slightly, and a worker
// can read its clock just before the start time another XCD recorded (an
// unsigned difference would then wrap and cancel the launch).
__device__ __forceinline__ bool past_deadline(const SolveArgs &a)
{
    if (a.limit_ticks == 0) return false;
    const int64_t elapsed = (int64_t)(wall_clock64() - ld_rlx(&a.glob->t0));
    if (elapsed <= (int64_t)a.limit_ticks) return false;
    st_rlx(&a.glob->timed_out, 1u);
    st_rlx(&a.glob->cancel, 1u);
    return true;
}

// ---------------------------------------------------------------------------
// one worker's view of its current tree node
// ---------------------------------------------------------------------------

// Neighbour sketch of a tree node.  Beside its degree (number of live
// neighbours; negative = in the cover) every vertex v keeps
//   s1(v) = sum of the ids of its live neighbours
//   s2(v) = sum of the squared ids of its live neighbours
// so a degree-1 vertex's neighbour is s1(v), and a degree-2 vertex's two
// neighbours n1 < n2 follow from n1 + n2 = s1 and n1^2 + n2^2 = s2:
// (n2 - n1)^2 = 2 s2 - s1^2.  That replaces the reference's walk over the
// whole original adjacency list to find the live neighbours
// (src/helperFunctions.cuh:240-252, 372-392) with local reads.
//
// Layouts (one stack / worklist slot = the arrays up to s2):
//  * int16 graphs (n < 2^15), 8 bytes per vertex: one 32-bit word
//    [s1 mod 2^16 : 16 | degree : 16] and a 32-bit s2.  The values decoded
//    (one id, the sum of two ids, n1^2 + n2^2 < 2^31) fit, and one 32-bit
//    atomic updates degree and s1 together: a degree being decremented is
//    >= 1, so no borrow crosses into s1, whose own wrap-around leaves the
//    word.
//  * int32 graphs, 16 bytes per vertex: int32 degrees, 32-bit s1, 64-bit s2.
template <typename DegT> struct SketchOf {
    using s2_t = uint32_t;
    using cand_t = uint16_t;  // vertex ids below 2^15
};
template <> struct SketchOf<int32_t> {
    using s2_t = unsigned long long;
    using cand_t = uint32_t;
};
// bytes of one node (degrees + sketch) per vertex
template <typename DegT> constexpr uint32_t node_bytes_per_vertex()
{
    return sizeof(DegT) == 2 ? 8u : 16u;
}

template <typename DegT, bool HBM, bool OFF16 = false, bool ADJ = false>
struct Node {
    static constexpr bool kPacked = sizeof(DegT) == 2;
    static constexpr bool kHbm = HBM;
    using s2_t = typename SketchOf<DegT>::s2_t;
    using cand_t = typename SketchOf<DegT>::cand_t;
    char *base;          // start of the node state (what a slot copy moves)
    uint32_t *dw;        // [npad] packed [s1:16 | degree:16] (int16 layout)
    int32_t *deg32;      // [npad] degrees (int32 layout)
    uint32_t *s1w;       // [npad] s1 (int32 layout)
    s2_t *s2;            // [npad] sum of squared live neighbour ids
    cand_t *cand;        // [ncap]: vertices whose degree fell 2 -> 1 since the last leaf pass
    cand_t *cand3;       // [ncap]: vertices whose degree fell 3 -> 2 since the last triangle pass
    uint32_t ncap;
    uint32_t *mark;      // [mwords] bitset, all zero between passes
    cand_t *list;        // [lcap]: vertices marked in the current pass
    int *xs;             // [64] scratch of apply_list: the owner of each edge of a window
    const uint32_t *src; // graph srcPtr (HBM)
    uint16_t *off16;     // [n+1] the same offsets in LDS, or nullptr (see off())
    const uint32_t *dst; // graph dst (HBM)
    uint16_t *adj16;     // [2m] the same adjacency in LDS (ADJ), or nullptr (see nb())
    const uint32_t *bits; // graph adjacency bitmap (HBM), or nullptr
    uint32_t n;
    uint32_t lcap;

    __device__ __forceinline__ void sync() const { nsync<HBM>(); }

    // dst[x] of the current graph: from the worker's LDS copy in the ADJ
    // kernel (launches of few graphs whose adjacency fits beside the node
    // state), else from HBM
    __device__ __forceinline__ uint32_t nb(uint32_t x) const
    {
        if constexpr (ADJ)
            return adj16[x];
        else
            return dst[x];
    }
    // srcPtr[i] of the current graph: from the worker's LDS copy in the
    // OFF16 kernel (batch launches whose graphs all have 2m < 2^16), else
    // from HBM
    __device__ __forceinline__ uint32_t off(uint32_t i) const
    {
        if constexpr (OFF16)
            return off16[i];
        else
            return src[i];
    }

    __device__ __forceinline__ int deg(uint32_t v) const
    {
        if constexpr (kPacked)
            return (int)(int16_t)(uint16_t)dw[v];
        else
            return deg32[v];
    }
    __device__ __forceinline__ uint32_t s1(uint32_t v) const
    {
        if constexpr (kPacked)
            return dw[v] >> 16;
        else
            return s1w[v];
    }
    // Degree of v (-1 for v >= n) and s1(v), read at a valid index without a
    // branch; one read in the packed layout.
    __device__ __forceinline__ int deg_s1(uint32_t v, uint32_t n, uint32_t &s) const
    {
        const uint32_t i = v < n ? v : 0u;
        int d;
        if constexpr (kPacked) {
            const uint32_t x = dw[i];
            d = (int)(int16_t)(uint16_t)x;
            s = x >> 16;
        } else {
            d = deg32[i];
            s = s1w[i];
        }
        return v < n ? d : -1;
    }
    // Degrees and s2 of the vertex pair (v, v + 1), v even, -1 past n: one
    // 8-byte read per array (16 bytes for 64-bit s2), issued together, so a
    // scan waits once per pair.  An even v < n <= npad has v + 1 < npad, and
    // v >= n reads at 0, so every read stays inside the arrays.
    __device__ __forceinline__ void pair(uint32_t v, uint32_t n, int &d0, int &d1) const
    {
        const uint32_t i = v < n ? v : 0u;
        int a, b;
        if constexpr (kPacked) {
            const uint2 x = *reinterpret_cast<const uint2 *>(dw + i);
            a = (int)(int16_t)(uint16_t)x.x;
            b = (int)(int16_t)(uint16_t)x.y;
        } else {
            const int2 x = *reinterpret_cast<const int2 *>(deg32 + i);
            a = x.x;
            b = x.y;
        }
        d0 = v < n ? a : -1;
        d1 = v + 1u < n ? b : -1;
    }
    __device__ __forceinline__ void pair_s2(uint32_t v, uint32_t n, s2_t &q0, s2_t &q1) const
    {
        const uint32_t i = v < n ? v : 0u;
        if constexpr (kPacked) {
            const uint2 x = *reinterpret_cast<const uint2 *>(s2 + i);
            q0 = x.x;
            q1 = x.y;
        } else {
            const ulonglong2 x = *reinterpret_cast<const ulonglong2 *>(s2 + i);
            q0 = x.x;
            q1 = x.y;
        }
    }
    // degree := -1 (v joins the cover).  Its sketch is not read again, and
    // later decrements through drop() only make the degree more negative
    // (at most its original degree of them: an int16 degree stays below 0,
    // so no borrow reaches s1).
    __device__ __forceinline__ void remove(uint32_t v) const
    {
        if constexpr (kPacked)
            reinterpret_cast<uint16_t *>(dw)[2u * v] = 0xFFFFu;
        else
            deg32[v] = -1;
    }
    // Vertex t loses its neighbour u; several lanes may hit the same t.
    // Returns t's degree before the update: negative when t is already in
    // the cover (the update then changes nothing that is read again).
    __device__ __forceinline__ int drop(uint32_t t, uint32_t u) const
    {
        int old;
        if constexpr (kPacked) {
            const uint32_t o = __hip_atomic_fetch_add(&dw[t], 0u - ((u << 16) + 1u), __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_WAVEFRONT);
            old = (int)(int16_t)(uint16_t)o;
        } else {
            old = __hip_atomic_fetch_add(&deg32[t], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            __hip_atomic_fetch_add(&s1w[t], 0u - u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
        __hip_atomic_fetch_add(&s2[t], (s2_t)0 - (s2_t)u * (s2_t)u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
        return old;
    }
    // The same with t updated by this lane only (dt = its current degree).
    __device__ __forceinline__ void drop_plain(uint32_t t, uint32_t u, int dt) const
    {
        if constexpr (kPacked) {
            dw[t] -= (u << 16) + 1u;
        } else {
            deg32[t] = dt - 1;
            s1w[t] -= u;
        }
        s2[t] -= (s2_t)u * (s2_t)u;
    }
    __device__ __forceinline__ void init_vertex(uint32_t v, int d, uint32_t sum, s2_t sq) const
    {
        if constexpr (kPacked) {
            dw[v] = (sum << 16) | (uint32_t)(uint16_t)(int16_t)d;
        } else {
            deg32[v] = d;
            s1w[v] = sum;
        }
        s2[v] = sq;
    }
};

// Appends cand (kNone = nothing) of every lane to the mark list; with dedupe
// a vertex already marked in this pass is dropped.  At most lcap entries are
// kept: a candidate past the end is unmarked again and found by the next
// pass (applying any subset of a pass's marks is sound, since the whole set
// is).  Returns the new length (wave-uniform).
template <class N>
__device__ __forceinline__ uint32_t append(const N &w, uint32_t L, uint32_t cand, bool dedupe)
{
    // most scan chunks mark nothing: leave at once (a wave-uniform branch)
    if (__ballot(cand != kNone) == 0ull) return L;
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
            w.list[pos] = (typename N::cand_t)cand;
        else if (dedupe)
            __hip_atomic_fetch_and(&w.mark[cand >> 5], ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    return uni(min(L + (uint32_t)__popcll(m), w.lcap));
}

// Degree of v, or -1 for v >= n (padding lanes of a scan chunk, empty
// candidate lanes).  The read goes to a clamped index, so a chunk needs no
// branch around it; n >= 1 wherever a chunk runs.
template <class N>
__device__ __forceinline__ int deg_or_none(const N &w, uint32_t v, uint32_t n)
{
    const int d = w.deg(v < n ? v : 0u);
    return v < n ? d : -1;
}

// Rule candidates of the current node.  A vertex can only become a leaf
// (degree 1) or an unexamined degree-2 vertex through a decrement: the
// decrements of every apply list each vertex whose degree falls 2 -> 1 (a
// leaf-rule candidate) or 3 -> 2 (a triangle-rule candidate).  Degrees never
// grow along a path, so a vertex enters each list at most once between two
// passes of its rule.  A count beyond the list capacity makes that rule's
// next pass scan every vertex, as does a root (full1 / full2).
struct Cands {
    uint32_t n1, n2;   // listed leaf / triangle candidates
    bool full1, full2; // the next leaf / triangle pass scans every vertex
};

__device__ __forceinline__ void cands_reset(Cands &c, bool full)
{
    c.n1 = 0;
    c.n2 = 0;
    c.full1 = full;
    c.full2 = full;
}

// Lists t by its degree `old` before a decrement (lanes with another old
// degree, or none, add nothing; every lane must call).
template <class N>
__device__ __forceinline__ void add_cands(const N &w, Cands &c, int old, uint32_t t)
{
    const uint64_t m1 = __ballot(old == 2);
    const uint64_t m2 = __ballot(old == 3);
    if ((m1 | m2) == 0ull) return;  // the common case: a wave-uniform exit
    if (old == 2) {
        const uint32_t pos = c.n1 + lanes_below(m1);
        if (pos < w.ncap) w.cand[pos] = (typename N::cand_t)t;
    } else if (old == 3) {
        const uint32_t pos = c.n2 + lanes_below(m2);
        if (pos < w.ncap) w.cand3[pos] = (typename N::cand_t)t;
    }
    c.n1 = uni(c.n1 + (uint32_t)__popcll(m1));
    c.n2 = uni(c.n2 + (uint32_t)__popcll(m2));
}

// Live vertex t loses its live neighbour u (several lanes may hit the same t).
// Exact integer square root of a perfect square below 2^30 (32-bit sums) or
// 2^62 (64-bit sums); the caller checks r * r == x.
// (the hardware square root is within 1 ulp: rounding gives the exact root
// of a perfect square below 2^30)
__device__ __forceinline__ uint32_t isqrt_round(uint32_t x)
{
    return (uint32_t)(__builtin_amdgcn_sqrtf((float)x) + 0.5f);
}
__device__ __forceinline__ unsigned long long isqrt_round(unsigned long long x)
{
    return (unsigned long long)(__builtin_sqrt((double)x) + 0.5);
}

// Top bit of s2 on a degree-2 vertex: the triangle rule examined it and it
// did not fire, i.e. a neighbour had degree 1 or degree 2 and a lower index
// (the skip), or its two neighbours are not adjacent.  Both outcomes are
// final while deg(v) == 2: the neighbour pair cannot change, degrees never
// grow along a path of the tree, a neighbour of degree 1 keeps v as its one
// neighbour, and a degree-2 neighbour can only fall to 1.  So the bit holds
// in every descendant node, and a settled vertex is not decoded again; once
// deg(v) < 2 the bit is never read.  Below the bit, n1^2 + n2^2 < 2^31
// (int16 degrees) or 2^63, so it never carries.
template <typename S2> constexpr S2 kSettled = (S2)1 << (8 * sizeof(S2) - 1);

// The two live neighbours n1 < n2 of a degree-2 vertex from its sketch
// (s2 without the kSettled bit).  False when the sketch does not decode
// to two distinct vertices below n (impossible for a consistent node).
template <class N>
__device__ __forceinline__ bool two_neighbours(const N &w, uint32_t v, typename N::s2_t q, uint32_t &n1,
                                               uint32_t &n2)
{
    using S2 = typename N::s2_t;
    const uint32_t s = w.s1(v);
    const S2 x = (S2)2 * q - (S2)s * (S2)s;
    const S2 r = isqrt_round(x);
    if (r * r != x || r == 0 || r > (S2)s) return false;
    n1 = (uint32_t)(((S2)s - r) >> 1);
    n2 = (uint32_t)(((S2)s + r) >> 1);
    return n2 < w.n;
}

// Apply phase of a rule: every listed vertex goes into the cover.  First all
// of them get -1, then each decrements its neighbours that are still live
// (src/helperFunctions.cuh:270-289).  Listed vertices are distinct and live.
// The adjacency lists of up to 64 listed vertices are laid end to end over
// the 64 lanes (wave prefix sum of their lengths), so one round of global
// loads serves 64 edges whatever the degrees.  Each edge is decremented
// without a prior degree read: drop() returns the old degree, and a
// negative one (a neighbour already in the cover) is ignored.
// One vertex u0 (wave-uniform, live) into the cover: its id and adjacency
// bounds are scalars, so no list walk, prefix scan or owner search.
template <class N>
__device__ __forceinline__ void apply_one(const N &w, uint32_t u0, Cands &c)
{
    // u0 is not among its own neighbours: its removal and the decrements
    // below touch different vertices and need no ordering between them
    if (lane_id() == 0) w.remove(u0);
    const uint32_t b0 = uni(w.off(u0)), e0 = uni(w.off(u0 + 1));
#pragma unroll 1
    for (uint32_t x0 = b0; x0 < e0; x0 += kWave) {
        const uint32_t x = x0 + lane_id();
        uint32_t t = kNone;
        int old = -1;
        if (x < e0) {
            t = w.nb(x);
            old = w.drop(t, u0);
        }
        add_cands(w, c, old, t);
    }
    w.sync();
}

template <class N>
__device__ void apply_list(const N &w, uint32_t L, Cands &c)
{
    L = uni(L);
    w.sync();
    if (L == 1u) {
        // one listed vertex (most passes of the high-degree rule)
        const uint32_t u0 = uni(w.list[0]);
        if (lane_id() == 0) w.mark[u0 >> 5] = 0u;
        apply_one(w, u0, c);
        return;
    }
    for (uint32_t i = lane_id(); i < L; i += kWave) w.remove(w.list[i]);
    w.sync();
    const int lane = (int)lane_id();
    int *const slot = w.xs;  // [64]: owner of each edge of one 64-edge window
    for (uint32_t c0 = 0; c0 < L; c0 += kWave) {
        const uint32_t cnt = min(L - c0, (uint32_t)kWave);
        int b = 0, len = 0, u = 0;
        if ((uint32_t)lane < cnt) {
            u = (int)w.list[c0 + lane];
            b = (int)w.off(u);
            len = (int)w.off(u + 1) - b;
        }
        const int incl = wave_incl_scan(len);
        const int excl = incl - len;
        const int total = __builtin_amdgcn_readlane(incl, kWave - 1);
        if (cnt == 1u) {
            // one listed vertex (most leaf passes): its own edges, no owner search
            const uint32_t b0 = (uint32_t)__builtin_amdgcn_readlane(b, 0);
            const uint32_t u0 = (uint32_t)__builtin_amdgcn_readlane(u, 0);
#pragma unroll 1
            for (int x0 = 0; x0 < total; x0 += kWave) {
                const int x = x0 + lane;
                uint32_t t = kNone;
                int old = -1;
                if (x < total) {
                    t = w.nb(b0 + (uint32_t)x);
                    old = w.drop(t, u0);
                }
                add_cands(w, c, old, t);
            }
            continue;
        }
#pragma unroll 1
        for (int x0 = 0; x0 < total; x0 += kWave) {
            // owner of edge x = the last entry j with excl_j <= x.  Entry j0
            // owns x0; each later entry that starts inside the window writes
            // its index at its start, and a prefix max over the lanes carries
            // every owner to the edges after its start.
            const int j0 = 63 - __clzll(__ballot((uint32_t)lane < cnt && excl <= x0));
            slot[lane] = j0;
            if ((uint32_t)lane < cnt && len > 0 && excl > x0 && excl < x0 + kWave) slot[excl - x0] = lane;
            w.sync();
            const int own = wave_incl_max(slot[lane]);
            const int ob = __shfl(b, own), ox = __shfl(excl, own), ou = __shfl(u, own);
            const int x = x0 + lane;
            uint32_t t = kNone;
            int old = -1;
            if (x < total) {
                t = w.nb((uint32_t)(ob + (x - ox)));
                old = w.drop(t, (uint32_t)ou);
            }
            add_cands(w, c, old, t);
            w.sync();
        }
    }
    for (uint32_t i = lane_id(); i < L; i += kWave) w.mark[w.list[i] >> 5] = 0u;
    w.sync();
}

// Per-lane statistics of the degree array, gathered by every rule scan in the
// same pass that reads the degrees for the rule.  After a scan that marked
// nothing they describe the current node exactly; that replaces the separate
// findNumOfEdges / findMaxDegree passes and tells which rule scans cannot
// mark anything.
struct Lanes {
    int s;        // sum of live degrees
    int md;       // max degree of the lane's vertices
    uint32_t mv;  // lowest index with degree md in the lane
};

__device__ __forceinline__ void lanes_reset(Lanes &st)
{
    st.s = 0;
    st.md = 0;
    st.mv = kNone;
}

// Vertices reach a lane in increasing order, so a strict > keeps the lowest
// index of the maximum (src/Sequential.cpp:50-58).
__device__ __forceinline__ void lanes_add(Lanes &st, int dv, uint32_t v)
{
    if (dv > 0) st.s += dv;
    if (dv > st.md) {
        st.md = dv;
        st.mv = v;
    }
}

// Leaf rule (src/helperFunctions.cuh:229-296): for every degree-1 vertex v
// mark its live neighbour u when deg(u) > 1, or deg(u) == 1 and u < v (so an
// isolated edge loses exactly one end).  Repeat until nothing is marked.
// The reference scans every vertex in every pass.  Here a pass looks at the
// leaf candidates only: a node enters the rules with no degree-1 vertex
// (its parent ended at the rules' fixpoint) except at a root, and every later
// degree-1 vertex got there by a decrement that listed it.  So a pass over
// the candidates marks exactly what a full scan would.  Full scans are kept
// for roots (c.full1) and after a pass whose marks hit the list capacity
// (the unmarked rest is found by the next full pass, as in the reference).
template <class N>
__device__ __forceinline__ uint32_t leaf_mark(const N &w, uint32_t v, uint32_t n)
{
    // branch-free: every lane reads at a valid index, the result is selected
    uint32_t s;
    const bool leaf = w.deg_s1(v, n, s) == 1;  // false past n and for empty candidates
    const uint32_t u = leaf ? s : 0u;
    const int du = deg_or_none(w, u, n);
    const bool mark = leaf & ((du > 1) | ((du == 1) & (u < v)));
    return mark ? u : kNone;
}

template <class N>
__device__ uint32_t rule_leaf(const N &w, Cands &c)
{
    uint32_t total = 0;
    for (;;) {
        // nothing listed (on entry, or by the last apply): nothing to mark
        // (a pass per node ended this way, counted on the batched workload)
        if (!c.full1 && c.n1 == 0u) break;
        uint32_t L = 0;
        if (c.full1 || c.n1 > w.ncap) {
            for (uint32_t base = 0, nv = uni(w.n); base < nv; base += kWave)
                L = append(w, L, leaf_mark(w, base + lane_id(), nv), true);
        } else {
            const uint32_t nv = uni(w.n), nc = c.n1;
            if (nc <= (uint32_t)kWave) {
                // one chunk of candidates (the common case): a pass that marks
                // one vertex, from one or more leaves, applies it directly
                const uint32_t i = lane_id();
                const uint32_t u = leaf_mark(w, i < nc ? (uint32_t)w.cand[i] : kNone, nv);
                const uint64_t m = __ballot(u != kNone);
                if (m == 0ull) {
                    c.n1 = 0;
                    c.full1 = false;
                    break;
                }
                const uint32_t u0 = __builtin_amdgcn_readlane(u, (int)__builtin_ctzll(m));
                if (__ballot(u != kNone && u != u0) == 0ull) {
                    c.n1 = 0;  // consumed; the apply lists the next candidates
                    apply_one(w, u0, c);
                    total += 1;
                    c.full1 = false;
                    continue;
                }
                L = append(w, L, u, true);
            } else {
                for (uint32_t i0 = 0; i0 < nc; i0 += kWave) {
                    const uint32_t i = i0 + lane_id();
                    L = append(w, L, leaf_mark(w, i < nc ? (uint32_t)w.cand[i] : kNone, nv), true);
                }
            }
        }
        c.n1 = 0;  // consumed; the apply below lists the next candidates
        if (L == 0) {
            c.full1 = false;
            break;
        }
        apply_list(w, L, c);
        total += L;
        c.full1 = L >= w.lcap;
    }
    return total;
}

// Sorted-adjacency membership test (binarySearch, src/helperFunctions.cuh:21-36),
// half-open [lo, hi).
template <class N>
__device__ __forceinline__ bool adj_has(const N &w, uint32_t lo, uint32_t hi, uint32_t x)
{
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        uint32_t y = w.nb(mid);
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
// One vertex of a triangle pass: v of degree dv with sketch sum q.  Sets
// c1, c2 to the two neighbours to cover, or settles v.
template <class N>
__device__ __forceinline__ void triangle_at(const N &w, uint32_t v, int dv, typename N::s2_t q, uint32_t &c1,
                                            uint32_t &c2)
{
    using S2 = typename N::s2_t;
    uint32_t n1, n2;
    if (dv == 2 && (q & kSettled<S2>) == 0 && two_neighbours(w, v, q, n1, n2)) {
        // skipped when a neighbour has degree 1, or degree 2 and a
        // lower index (src/helperFunctions.cuh:379-385)
        const int d1 = w.deg(n1), d2 = w.deg(n2);
        const bool skip = (d1 == 1) | ((d1 == 2) & (n1 < v)) | (d2 == 1) | ((d2 == 2) & (n2 < v));
        bool adj;
        if (w.bits) {
            const uint32_t bit = n1 * w.n + n2;
            adj = (w.bits[bit >> 5] >> (bit & 31u)) & 1u;
        } else {
            adj = adj_has(w, w.off(n2), w.off(n2 + 1), n1);
        }
        if (!skip && adj) {
            c1 = n1;
            c2 = n2;
        } else {
            w.s2[v] = q | kSettled<S2>;
        }
    }
}

// The triangle rule over its candidates: an unexamined degree-2 vertex is
// either a root's (c.full2: every vertex is looked at) or got there by a
// 3 -> 2 decrement that listed it, and examined ones carry the settled bit
// until their degree changes (kSettled).  So a pass over the candidates
// finds every triangle a full scan would.  A pass whose marks hit the list
// capacity (the unmarked rest stays unexamined) makes the next pass full.
template <class N>
__device__ uint32_t rule_triangle(const N &w, Cands &c)
{
    uint32_t total = 0;
    for (;;) {
        uint32_t L = 0;
        const uint32_t nv = uni(w.n);
        if (c.full2 || c.n2 > w.ncap) {
            // two vertices per lane: (v0, v0 + 1)
            for (uint32_t base = 0; base < nv; base += 2u * kWave) {
                const uint32_t v0 = base + 2u * lane_id();
                int d0, d1;
                typename N::s2_t q0, q1;
                w.pair(v0, nv, d0, d1);
                w.pair_s2(v0, nv, q0, q1);
                uint32_t a1 = kNone, a2 = kNone, b1 = kNone, b2 = kNone;
                triangle_at(w, v0, d0, q0, a1, a2);
                triangle_at(w, v0 + 1u, d1, q1, b1, b2);
                if (__ballot((a1 != kNone) | (b1 != kNone)) != 0ull) {
                    L = append(w, L, a1, true);
                    L = append(w, L, a2, true);
                    L = append(w, L, b1, true);
                    L = append(w, L, b2, true);
                }
            }
        } else {
            const uint32_t nc = c.n2;
            for (uint32_t i0 = 0; i0 < nc; i0 += kWave) {
                const uint32_t i = i0 + lane_id();
                const uint32_t v = i < nc ? (uint32_t)w.cand3[i] : kNone;
                const uint32_t vi = v < nv ? v : 0u;
                const int dv = v < nv ? w.deg(vi) : -1;
                const typename N::s2_t q = w.s2[vi];
                uint32_t a1 = kNone, a2 = kNone;
                triangle_at(w, vi, dv, q, a1, a2);
                if (__ballot(a1 != kNone) != 0ull) {
                    L = append(w, L, a1, true);
                    L = append(w, L, a2, true);
                }
            }
        }
        c.n2 = 0;  // consumed; the apply below lists the next candidates
        c.full2 = false;
        if (L == 0) break;
        c.full2 = L >= w.lcap;
        apply_list(w, L, c);
        total += L;
    }
    return total;
}

// One scan of every vertex, two per lane, for the high-degree rule
// (src/helperFunctions.cuh:299-358: v goes into the cover when
// deg(v) + d >= T, all marks of a pass applied together) and the node's
// statistics in the same reads: live edge count, maximum degree and its
// lowest index (findNumOfEdges / findMaxDegree,
// src/helperFunctions.cuh:450-504).  Returns the number of marks (listed in
// w.list, at most lcap; the rest are found by the next scan).  When nothing
// is marked, st describes the node exactly.
template <class N>
__device__ uint32_t scan_high(const N &w, uint32_t d, uint32_t T, Lanes &st)
{
    uint32_t L = 0;
    lanes_reset(st);
    for (uint32_t base = 0, nv = uni(w.n); base < nv; base += 2u * kWave) {
        const uint32_t v0 = base + 2u * lane_id();
        int d0, d1;
        w.pair(v0, nv, d0, d1);  // -1 past n: adds nothing below
        lanes_add(st, d0, v0);
        lanes_add(st, d1, v0 + 1u);
        const bool h0 = (d0 > 0) & ((uint32_t)d0 + d >= T);
        const bool h1 = (d1 > 0) & ((uint32_t)d1 + d >= T);
        if (__ballot(h0 | h1) != 0ull) {
            L = append(w, L, h0 ? v0 : kNone, false);
            L = append(w, L, h1 ? v0 + 1u : kNone, false);
        }
    }
    return L;
}

struct Summary {
    uint32_t edges;  // live edges (findNumOfEdges, src/helperFunctions.cuh:483-504)
    int32_t maxdeg;  // findMaxDegree (src/helperFunctions.cuh:450-481)
    uint32_t maxv;   // lowest index of maximum degree
};

// Rule fixpoint of one node.  The reference applies leaf, triangle and
// high-degree passes in a fixed round (src/GlobalWorkList.cuh:107-130) with
// every pass a scan of all vertices.  Here the leaf and triangle rules run
// from their candidate lists until neither marks anything, then one scan of
// every vertex applies the high-degree rule and gathers the node's
// statistics; after a high-degree deletion the candidate-driven rules run
// again.  The node leaves when that scan marks nothing: every rule is then at
// its fixpoint, as at the end of the reference's round, and each deletion is
// one of the three rules applied to the current graph, so the search stays
// exact (the tree can differ, as the reference's GPU and CPU trees already
// do).  Returns the new cover size and, in *sum, the node's edge count and
// max-degree vertex.
template <class N>
__device__ uint32_t reduce_node(const N &w, uint32_t d, uint32_t T, Summary *sum, Cands &c)
{
    Lanes st;
    for (;;) {
        for (;;) {
            d += rule_leaf(w, c);
            const uint32_t nt = rule_triangle(w, c);
            d += nt;
            if (nt == 0) break;  // the leaf rule was at its fixpoint before this triangle pass
        }
        const uint32_t L = scan_high(w, d, T, st);
        if (L == 0) break;
        apply_list(w, L, c);
        d += L;
    }
    const int md = wave_max(st.md);
    sum->edges = (uint32_t)wave_sum(st.s) / 2u;
    sum->maxdeg = md;
    sum->maxv = wave_min_u32((st.md == md && md > 0) ? st.mv : kNone);
    return d;
}

// Left child in place: v into the cover (deleteMaxDegreeVertex,
// src/helperFunctions.cuh:200-215).
template <class N>
__device__ void take_vertex(const N &w, uint32_t v, Cands &c)
{
    w.sync();
    if (lane_id() == 0) w.remove(v);
    w.sync();
    const uint32_t b = w.off(v), e = w.off(v + 1);
    for (uint32_t x0 = b; x0 < e; x0 += kWave) {
        const uint32_t x = x0 + lane_id();
        uint32_t t = kNone;
        int dt = -1;
        if (x < e) {
            t = w.nb(x);
            dt = w.deg(t);
            // the neighbours of one vertex are distinct: plain updates
            if (dt >= 0) w.drop_plain(t, v, dt);
        }
        add_cands(w, c, dt, t);
    }
    w.sync();
}

// Right child: every live neighbour of v into the cover
// (deleteNeighborsOfMaxDegreeVertex, src/helperFunctions.cuh:156-184), in
// rounds of at most lcap vertices.  Returns how many vertices were added
// (deg(v) before the call).
template <class N>
__device__ uint32_t take_neighbours(const N &w, uint32_t v, Cands &c)
{
    uint32_t total = 0;
    const uint32_t b = w.off(v), e = w.off(v + 1);
    for (;;) {
        uint32_t L = 0;
        for (uint32_t x0 = b; x0 < e && L < w.lcap; x0 += kWave) {
            uint32_t x = x0 + lane_id();
            uint32_t cand = kNone;
            if (x < e) {
                uint32_t t = w.nb(x);
                if (w.deg(t) >= 0) cand = t;
            }
            L = append(w, L, cand, false);
        }
        if (L == 0) break;
        apply_list(w, L, c);
        total += L;
        if (L < w.lcap) break;  // no entry was dropped: every live neighbour is done
    }
    return total;
}

// Root node of a graph: its degrees, and the sketch sums over every
// neighbour (all are live).  Padding vertices [n, npad) are in the cover.
template <class N>
__device__ void init_root(const N &w, const int32_t *deg0, uint32_t npad)
{
    using S2 = typename N::s2_t;
    for (uint32_t v = lane_id(); v < npad; v += kWave) {
        uint32_t s = 0;
        S2 q = 0;
        int dv = -1;
        if (v < w.n) {
            dv = deg0[v];
            const uint32_t e = w.off(v + 1);
            for (uint32_t x = w.off(v); x < e; ++x) {
                const uint32_t u = w.nb(x);
                s += u;
                q += (S2)u * (S2)u;
            }
        }
        w.init_vertex(v, dv, s, q);
    }
    w.sync();
}

// Copies the node state of the first cnt vertices between two node-state
// layouts with the same npad (node storage <-> a stack or worklist slot).
// A slot is sized for the largest graph of the launch; only the current
// graph's vertices travel (rounded up to whole 16-byte vectors).  The packed
// layout's two arrays are copied as one run of 16-byte vectors over both
// (the second array starts npad * 4 bytes in, npad a multiple of 32).  Each
// lane loads up to four vectors before it stores any, so a copy from HBM
// (a pop or a steal) waits for one round trip per 256 vectors (4 KB)
// instead of one per 64.
// 16-byte vectors in flight per lane in a register copy between node
// storage and a slot (measured: 2 beat 1 by 0.8% and 4, which spills, lost 3%)
constexpr uint32_t kCopyVecs = 2;

template <class N>
__device__ __forceinline__ void copy_state(char *__restrict__ dst, const char *__restrict__ src, uint32_t npad,
                                           uint32_t cnt)
{
    const uint32_t q = (cnt + 3u) >> 2;  // 16-byte vectors per array of 4-byte entries
    const int4 *s4 = reinterpret_cast<const int4 *>(src);
    int4 *d4 = reinterpret_cast<int4 *>(dst);
    const uint32_t o1 = npad >> 2;       // second array, at npad * 4 bytes
    if constexpr (N::kPacked) {
        // vector k < q is entry k of the first array, k >= q entry k - q of the second
        const uint32_t total = 2u * q;
        for (uint32_t k0 = lane_id(); k0 < total; k0 += kCopyVecs * kWave) {
            int4 x[kCopyVecs];
#pragma unroll
            for (uint32_t j = 0; j < kCopyVecs; ++j) {
                const uint32_t k = k0 + j * kWave;
                if (k < total) x[j] = s4[k < q ? k : o1 + (k - q)];
            }
#pragma unroll
            for (uint32_t j = 0; j < kCopyVecs; ++j) {
                const uint32_t k = k0 + j * kWave;
                if (k < total) d4[k < q ? k : o1 + (k - q)] = x[j];
            }
        }
    } else {
        const uint32_t o2 = npad >> 1;   // 8-byte s2 entries, at npad * 8 bytes
        for (uint32_t i = lane_id(); i < q; i += kWave) {
            const int4 x = s4[i], y = s4[o1 + i], z0 = s4[o2 + 2u * i], z1 = s4[o2 + 2u * i + 1u];
            d4[i] = x;
            d4[o1 + i] = y;
            d4[o2 + 2u * i] = z0;
            d4[o2 + 2u * i + 1u] = z1;
        }
    }
}

// Agent-coherent 16-byte accesses to a worklist slot: buffer loads and
// stores with the sc1 cache policy (the policy relaxed agent-scope atomics
// use), through a buffer resource that spans the slot.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kCoherent = 16;                 // cache policy bits: sc1
constexpr int kRawBufferFlags = 0x00020000;   // gfx9 buffer resource word 3: 32-bit data format
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slot_rsrc(const void *slot, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(slot), 0, (int)bytes, kRawBufferFlags);
}

// Node state into a worklist slot (shared) from the worker's own stack slot
// or node storage (src); the same vectors as copy_state.
template <class N>
__device__ __forceinline__ void copy_state_to_ring(__amdgpu_buffer_rsrc_t dst, const char *__restrict__ src,
                                                   uint32_t npad, uint32_t cnt)
{
    const uint32_t q = (cnt + 3u) >> 2;
    const u32x4 *s4 = reinterpret_cast<const u32x4 *>(src);
    const uint32_t o1 = npad >> 2;
    if constexpr (N::kPacked) {
        const uint32_t total = 2u * q;
        for (uint32_t k0 = lane_id(); k0 < total; k0 += kCopyVecs * kWave) {
            u32x4 x[kCopyVecs];
#pragma unroll
            for (uint32_t j = 0; j < kCopyVecs; ++j) {
                const uint32_t k = k0 + j * kWave;
                if (k < total) x[j] = s4[k < q ? k : o1 + (k - q)];
            }
#pragma unroll
            for (uint32_t j = 0; j < kCopyVecs; ++j) {
                const uint32_t k = k0 + j * kWave;
                if (k < total) __builtin_amdgcn_raw_buffer_store_b128(x[j], dst, 16u * (k < q ? k : o1 + (k - q)), 0, kCoherent);
            }
        }
    } else {
        const uint32_t o2 = npad >> 1;
        for (uint32_t i = lane_id(); i < q; i += kWave) {
            const u32x4 x = s4[i], y = s4[o1 + i], z0 = s4[o2 + 2u * i], z1 = s4[o2 + 2u * i + 1u];
            __builtin_amdgcn_raw_buffer_store_b128(x, dst, 16u * i, 0, kCoherent);
            __builtin_amdgcn_raw_buffer_store_b128(y, dst, 16u * (o1 + i), 0, kCoherent);
            __builtin_amdgcn_raw_buffer_store_b128(z0, dst, 16u * (o2 + 2u * i), 0, kCoherent);
            __builtin_amdgcn_raw_buffer_store_b128(z1, dst, 16u * (o2 + 2u * i + 1u), 0, kCoherent);
        }
    }
}

// Node state from a worklist slot (shared) into the worker's HBM node
// storage.
template <class N>
__device__ __forceinline__ void copy_state_from_ring(char *__restrict__ dst, __amdgpu_buffer_rsrc_t src, uint32_t npad,
                                                     uint32_t cnt)
{
    const uint32_t q = (cnt + 3u) >> 2;
    u32x4 *d4 = reinterpret_cast<u32x4 *>(dst);
    const uint32_t o1 = npad >> 2;
    if constexpr (N::kPacked) {
        const uint32_t total = 2u * q;
        for (uint32_t k0 = lane_id(); k0 < total; k0 += kCopyVecs * kWave) {
            u32x4 x[kCopyVecs];
#pragma unroll
            for (uint32_t j = 0; j < kCopyVecs; ++j) {
                const uint32_t k = k0 + j * kWave;
                if (k < total) x[j] = __builtin_amdgcn_raw_buffer_load_b128(src, 16u * (k < q ? k : o1 + (k - q)), 0, kCoherent);
            }
#pragma unroll
            for (uint32_t j = 0; j < kCopyVecs; ++j) {
                const uint32_t k = k0 + j * kWave;
                if (k < total) d4[k < q ? k : o1 + (k - q)] = x[j];
            }
        }
    } else {
        const uint32_t o2 = npad >> 1;
        for (uint32_t i = lane_id(); i < q; i += kWave) {
            const u32x4 x = __builtin_amdgcn_raw_buffer_load_b128(src, 16u * i, 0, kCoherent);
            const u32x4 y = __builtin_amdgcn_raw_buffer_load_b128(src, 16u * (o1 + i), 0, kCoherent);
            const u32x4 z0 = __builtin_amdgcn_raw_buffer_load_b128(src, 16u * (o2 + 2u * i), 0, kCoherent);
            const u32x4 z1 = __builtin_amdgcn_raw_buffer_load_b128(src, 16u * (o2 + 2u * i + 1u), 0, kCoherent);
            d4[i] = x;
            d4[o1 + i] = y;
            d4[o2 + 2u * i] = z0;
            d4[o2 + 2u * i + 1u] = z1;
        }
    }
}

// A stack or worklist slot (HBM) into LDS node storage, packed layout:
// every lane's 16-byte vector goes straight from global memory into LDS
// (global_load_lds_dwordx4: lane i of a load writes 16 bytes at the
// wave-uniform LDS address + 16 i), with no registers in between, so every
// load of the copy is in flight at once and the copy waits for one round
// trip.  Lanes past the copied vectors are masked off and write nothing.
// kShared: the slot is a worklist slot, loaded agent-coherent (sc1).
template <class N, bool kShared = false>
__device__ __forceinline__ void copy_state_to_lds(char *dst, const char *src, uint32_t npad, uint32_t cnt)
{
    static_assert(N::kPacked && !N::kHbm, "LDS node storage, packed layout");
    const uint32_t q = (cnt + 3u) >> 2;  // 16-byte vectors per array of 4-byte entries
    const uint32_t abytes = npad * 4u;   // second array at npad * 4 bytes
#pragma unroll
    for (uint32_t arr = 0; arr < 2u; ++arr) {
        const char *s = src + arr * abytes;
        char *d = dst + arr * abytes;
        for (uint32_t k0 = 0; k0 < q; k0 += kWave) {
            const uint32_t k = k0 + lane_id();
            if (k < q)
                __builtin_amdgcn_global_load_lds(reinterpret_cast<const int4 *>(s) + k,
                                                 reinterpret_cast<int4 *>(d) + k0, 16, 0, kShared ? 16 : 0);
        }
    }
    __builtin_amdgcn_s_waitcnt(0x3f70);  // vmcnt(0): every vector has landed in LDS
}

// HBM bytes one copy_state of cnt vertices reads or writes on the slot side
// (16-byte vectors: two 4-byte arrays packed, or 4 + 4 + 8 bytes per vertex).
template <class N>
__device__ __forceinline__ uint32_t copy_bytes(uint32_t cnt)
{
    const uint32_t q = (cnt + 3u) >> 2;
    return N::kPacked ? 32u * q : 64u * q;
}

// ---------------------------------------------------------------------------
// broker worklist (src/BWDWorkList.cuh) with 64-bit ring positions
// ---------------------------------------------------------------------------

// Writes one subtree (degrees at src, cover size d, pending branch vertex v
// of graph g) into the ring if the occupancy guard admits it.  The caller
// has already decided that a donation is wanted.
template <class N>
__device__ bool try_enqueue(const SolveArgs &a, const N &w, const char *src, uint32_t g, uint32_t d, uint32_t v,
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
            // the slot is free once its previous consumer has stored 2 lap
            while (ld_rlx(&a.tickets[slot]) != ticket) {
                backoff(spins < 3 ? spins : 3);
                ++spins;
                if (past_deadline(a)) {
                    ok = 0;  // the launch is being cancelled: leave the slot alone
                    break;
                }
            }
            // the subtree being donated is one more outstanding unit of g
            if (ok) fetch_add(&a.gstate[g].outstanding, 1u);
        } else {
            fetch_add(&a.glob->count, -1);
        }
    }
    ok = uni(ok);
    if (!ok) return false;
    slot = uni(slot);
    ticket = uni(ticket);
    handoff_observe();
    copy_state_to_ring<N>(slot_rsrc(static_cast<char *>(a.q_deg) + (size_t)slot * slot_bytes, slot_bytes), src, a.npad,
                          w.n);
    if (lane_id() == 0) {
        const u32x4 m = {g, d, v, 0u};
        __builtin_amdgcn_raw_buffer_store_b128(m, slot_rsrc(&a.q_meta[slot], 16u), 0u, 0, kCoherent);
    }
    handoff_publish();
    if (lane_id() == 0) {
        st_rlx(&a.tickets[slot], ticket + 1u);
        fetch_add(&a.glob->counter, kEnqOne);
    }
    return true;
}

// Takes one item if the occupancy guard says one is there.  On success the
// degree array is in the worker's node storage and (g, d, v) are returned.
template <bool SERVE, class N>
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
            while (ld_rlx(&a.tickets[slot]) != ticket) {
                backoff(spins < 3 ? spins : 3);
                ++spins;
                if (past_deadline(a)) {
                    ok = 0;  // the launch is being cancelled: leave the slot alone
                    break;
                }
            }
        } else {
            fetch_add(&a.glob->count, 1);
        }
    }
    ok = uni(ok);
    if (!ok) return false;
    slot = uni(slot);
    ticket = uni(ticket);
    // serving: the request's descriptor and graph were written by the worker
    // that claimed it (load_request, agent-scope release); they are read
    // after an agent-scope acquire.  Batch graphs are read-only.
    if constexpr (SERVE)
        fence_acq();
    else
        handoff_observe();
    const u32x4 m = __builtin_amdgcn_raw_buffer_load_b128(slot_rsrc(&a.q_meta[slot], 16u), 0u, 0, kCoherent);
    g = uni(m.x);
    d = uni(m.y);
    v = uni(m.z);
    if constexpr (N::kPacked && !N::kHbm)
        copy_state_to_lds<N, true>(w.base, static_cast<const char *>(a.q_deg) + (size_t)slot * slot_bytes, a.npad,
                                   uni(a.gdesc[g].n));
    else
        copy_state_from_ring<N>(w.base, slot_rsrc(static_cast<const char *>(a.q_deg) + (size_t)slot * slot_bytes, slot_bytes),
                                a.npad, uni(a.gdesc[g].n));
    w.sync();
    handoff_publish();  // the slot's loads are complete before it is freed
    if (lane_id() == 0) {
        st_rlx(&a.tickets[slot], ticket + 1u);
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
    for (uint32_t base = 0, nv = uni(w.n); base < nv; base += kWave) {
        uint32_t v = base + lane_id();
        uint64_t m = __ballot(v < w.n && w.deg(v) < 0);
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
            // A wait cut short by the deadline leaves the lock to its holder:
            // no cover is written (the launch ends unfinished anyway) and the
            // lock is not released by a wave that does not own it.
            uint32_t owned = 0;
            if (lane_id() == 0) {
                for (;;) {
                    if (atomicCAS(&gs->cover_lock, 0u, 1u) == 0u) {
                        owned = 1;
                        break;
                    }
                    backoff(2);
                    if (past_deadline(a)) break;
                }
            }
            if (uni(owned)) {
                fence_acq();
                int cur = (int)uni((uint32_t)ld_rlx(&gs->best));
                if (cur == (int)d) {
                    write_cover(a, w, gd);
                    if (lane_id() == 0) st_rlx(&gs->cover_size, (int32_t)d);
                }
                fence_rel();
                if (lane_id() == 0) st_rel(&gs->cover_lock, 0u);
            }
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

// Nodes a worker processes on graph g before it adds them to g's count.
// Without a budget: kFlushEvery.  With one, what the worker's last view of
// the count (`seen`) leaves of the budget, divided by twice the number of
// workers that may share the graph (rounded to a power of two), so their
// unflushed nodes stay below half of what is left and a budget-capped graph
// stops within a few percent of its budget.  (With fixed 1,024-node
// flushes, graphs that many workers shared at the end of a batch ran up to
// 3.5 times their budget.)  The sharers: the graph's outstanding units, but
// at least the launch's workers per graph (a.flush_share) -- a graph can
// be shared by more workers than it has units at the moment, e.g. when a
// single graph's search widens or a staged launch's sleeping workers join
// (counting units alone let one overrun its budget by a quarter).  Dividing
// by all workers in a batch made every flush 64 nodes and the batched step
// 4% slower.
constexpr uint32_t kFlushEvery = 1024;
template <bool kByWorkers>
__device__ __forceinline__ uint32_t flush_interval(const SolveArgs &a, const GDesc &gd, uint32_t g, uint64_t seen)
{
    if (!gd.max_nodes) return kFlushEvery;
    if (seen >= gd.max_nodes) return 1u;
    const uint32_t units = kByWorkers ? max(a.workers, 1u)
                                      : max(ld_rlx(&a.gstate[g].outstanding), max(a.flush_share, 1u));
    const uint32_t shift = 1u + (units > 1u ? 32u - (uint32_t)__builtin_clz(units - 1u) : 0u);
    const uint64_t k = (gd.max_nodes - seen) >> shift;
    return k >= kFlushEvery ? kFlushEvery : (k > 0u ? (uint32_t)k : 1u);
}

// Adds the nodes spent on graph g since the last flush to its count (and
// stops it at its budget), and the HBM bytes of the slot copies made for it
// (bytes_per_copy each) to the launch total.  Returns the next flush
// interval when there was something to add.
template <bool kByWorkers>
__device__ __forceinline__ void flush_nodes(const SolveArgs &a, const GDesc &gd, uint32_t g, uint64_t &pending,
                                            uint32_t &copies, uint32_t bytes_per_copy, uint32_t &flush_every,
                                            bool unit_end)
{
    // copy bytes go to one launch-wide word: added at the end of a unit or
    // every 256 copies, not at every (near a budget, frequent) node flush
    if (unit_end || copies >= 256u) {
        if (lane_id() == 0 && copies != 0u)
            fetch_add(&a.glob->state_bytes, (unsigned long long)copies * bytes_per_copy);
        copies = 0;
    }
    if (pending == 0) return;
    uint32_t next = 0;
    if (lane_id() == 0) {
        GState *gs = &a.gstate[g];
        const uint64_t after = fetch_add(&gs->nodes, (uint64_t)pending) + pending;
        if (gd.max_nodes && after >= gd.max_nodes) {
            st_rlx(&gs->budget_hit, 1u);
            st_rlx(&gs->stop, 1u);
        }
        next = flush_interval<kByWorkers>(a, gd, g, after);
    }
    flush_every = uni(next);
    pending = 0;
}

// One outstanding subtree of graph g is finished.  The last one marks the
// graph done; in serving mode the whole wave then copies the cover bitmap to
// the request's pinned host area and lane 0 writes the completion record
// (fields first, the sequence number last, with a system-scope release in
// between so the host only ever sees complete records).
template <bool SERVE>
__device__ void finish_unit(const SolveArgs &a, uint32_t g, const GDesc &gd)
{
    GState *gs = &a.gstate[g];
    uint32_t last = 0;
    if (lane_id() == 0) {
        // serving: the last finisher reads the cover other workers recorded
        // (acquire); a batch result is read by the host after the launch
        uint32_t old = SERVE ? __hip_atomic_fetch_add(&gs->outstanding, 0xFFFFFFFFu, __ATOMIC_ACQ_REL,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                             : fetch_add(&gs->outstanding, 0xFFFFFFFFu);
        last = old == 1u;
        if (last) {
            st_rlx(&gs->done_tick, (uint64_t)wall_clock64());
            st_rlx(&gs->done, 1u);
        }
    }
    if constexpr (!SERVE) {
        return;
    } else {
    if (!uni(last)) return;
    fence_acq();
    const uint32_t found = uni(ld_rlx(&gs->found));
    const int32_t cs = (int32_t)uni((uint32_t)ld_rlx(&gs->cover_size));
    const uint32_t best = uni((uint32_t)ld_rlx(&gs->best));
    const bool has_cover = gd.instance == 0 ? (cs >= 0 && (uint32_t)cs == best) : found != 0u;
    if (has_cover) {
        const uint32_t words = ((gd.n + 63u) / 64u) * 2u;
        for (uint32_t i = lane_id(); i < words; i += kWave)
            __hip_atomic_store(&a.host_arena[gd.off_cover + i], a.covers[gd.off_cover + i], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
    }
    if (lane_id() == 0) {
        Completion *c = &a.comp[g];
        const int sys = __HIP_MEMORY_SCOPE_SYSTEM;
        uint32_t minimum = gd.instance == 0 ? best : (found ? (uint32_t)cs : 0u);
        uint32_t flags = (ld_rlx(&gs->budget_hit) ? 1u : 0u) | (ld_rlx(&a.glob->cancel) ? 2u : 0u) |
                         (has_cover ? 4u : 0u);
        __hip_atomic_store(&c->minimum, minimum, __ATOMIC_RELAXED, sys);
        __hip_atomic_store(&c->k_found, found, __ATOMIC_RELAXED, sys);
        __hip_atomic_store(&c->flags, flags, __ATOMIC_RELAXED, sys);
        __hip_atomic_store(&c->nodes, ld_rlx(&gs->nodes), __ATOMIC_RELAXED, sys);
        __hip_atomic_store(&c->cover_size, cs, __ATOMIC_RELAXED, sys);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    if (lane_id() == 0) __hip_atomic_store(&a.comp[g].seq, gd.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Serving: the worker that claimed request slot g brings it into HBM -- the
// descriptor into gdesc[g], the graph words into the device arena at the
// same offsets, and a fresh GState -- before anyone else can reach it (other
// workers only see g through this worker's donations, which publish with
// release/acquire on the worklist tickets).
__device__ void load_request(const SolveArgs &a, uint32_t g)
{
    const uint32_t *hd = reinterpret_cast<const uint32_t *>(&a.host_desc[g]);
    uint32_t *dd = reinterpret_cast<uint32_t *>(const_cast<GDesc *>(&a.gdesc[g]));
    uint32_t word = 0;
    if (lane_id() < 16u) {
        word = __hip_atomic_load(&hd[lane_id()], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        dd[lane_id()] = word;
    }
    // GDesc words: 0 n, 2 off_src, 4 off_deg, 8 U
    const uint32_t n = __builtin_amdgcn_readlane(word, 0);
    const uint32_t lo = __builtin_amdgcn_readlane(word, 2);
    const uint32_t hi = __builtin_amdgcn_readlane(word, 4) + n;
    const uint32_t U = __builtin_amdgcn_readlane(word, 8);
    uint32_t *dev = const_cast<uint32_t *>(a.gbuf);
    for (uint32_t i = lo + lane_id(); i < hi; i += kWave) dev[i] = a.host_arena[i];
    if (lane_id() == 0) {
        GState s = {};
        s.best = (int32_t)U;
        s.outstanding = 1u;
        s.cover_size = -1;
        a.gstate[g] = s;
    }
    fence_rel();
    fence_acq();
}

// Number of roots that may be handed out.  Batch: all of them.  Serving: the
// host's request tail, polled at most once per poll interval by one worker
// (system-scope acquire on pinned memory) and shared through a device copy.
template <bool SERVE>
__device__ __forceinline__ uint32_t roots_available(const SolveArgs &a)
{
    if constexpr (!SERVE) return a.roots;
    uint32_t t = 0;
    if (lane_id() == 0) {
        t = ld_rlx(&a.glob->host_tail);
        uint64_t now = wall_clock64();
        uint64_t last = ld_rlx(&a.glob->last_poll);
        if (now - last > a.poll_ticks &&
            __hip_atomic_compare_exchange_strong(&a.glob->last_poll, &last, now, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
            uint32_t ht = __hip_atomic_load(&a.host->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
            uint32_t sd = __hip_atomic_load(&a.host->shutdown, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
            if (ht > t) {
                __hip_atomic_fetch_max(&a.glob->host_tail, ht, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                t = ht;
            }
            if (sd) st_rlx(&a.glob->host_shutdown, 1u);
        }
    }
    return uni(t);
}

// Idle worker: take a queued item, else a new root, else wait; returns
// 0 = terminate, 1 = queued item (degrees already in node storage), 2 = root.
struct Acquired {
    uint32_t kind, g, d, v, root;
};

// Work for an idle worker: a queued subtree (kind 1), a root (kind 2: a
// graph, a static path, or a served request), or nothing (kind 0: the
// launch is over).  A worker counts as "waiting" in the termination word
// while it is in here; `counted` says the caller already added it (a worker
// registers as registered-and-waiting in one atomic at its start).
//
// Batch launches count their roots in the termination word's "enqueued"
// field from the start (init_kernel), like items of the worklist: a claim
// takes the root index from next_root and then leaves "waiting" and
// "enqueued" in one atomic, so termination is simply waiting == registered
// with nothing enqueued, and a failed claim undoes nothing.  At a worker's
// start only the first `roots` workers to register try to claim roots (the
// others join in after two polls): with one root and 7,168 workers every
// worker's claim attempt on the same words cost more than the search of a
// small graph.
template <bool SERVE, class N>
__device__ Acquired acquire(const SolveArgs &a, const N &w, uint32_t slot_bytes, bool counted, uint32_t rid)
{
    Acquired r = {0u, kNone, 0u, kNone, 0u};
    // A cancelled or late launch hands out no more work.  Checked on every
    // call, not only when the worklist is empty: a worker that cycles
    // through queued units shorter than kFlushEvery nodes never reaches the
    // per-node check, and kept a launch running far past its time limit.
    uint32_t late = 0;
    if (lane_id() == 0) late = ld_rlx(&a.glob->cancel) || past_deadline(a);
    if (uni(late)) return r;
    if (!counted && lane_id() == 0) fetch_add(&a.glob->counter, kWaitOne);
    uint32_t spins = 0;
    for (;;) {
        if (a.version == 1) {
            uint32_t g, d, v;
            if (try_dequeue<SERVE>(a, w, g, d, v, slot_bytes)) {
                r.kind = 1;
                r.g = g;
                r.d = d;
                r.v = v;
                return r;
            }
        }
        if constexpr (!SERVE) {
            if ((rid < a.roots || spins >= 2u) && uni(ld_rlx(&a.glob->next_root)) < a.roots) {
                uint32_t idx = kNone;
                if (lane_id() == 0) {
                    idx = fetch_add(&a.glob->next_root, 1u);
                    if (idx < a.roots) fetch_add(&a.glob->counter, (unsigned long long)(0ull - (kEnqOne + kWaitOne)));
                }
                idx = uni(idx);
                if (idx < a.roots) {
                    r.kind = 2;
                    r.root = idx;
                    return r;
                }
            }
        } else {
            const uint32_t avail = roots_available<SERVE>(a);
            if (uni(ld_rlx(&a.glob->next_root)) < avail) {
                uint32_t idx = kNone;
                if (lane_id() == 0) {
                    fetch_add(&a.glob->counter, (unsigned long long)(0ull - kWaitOne));
                    // never move the counter past the published tail, or a
                    // request published later would be skipped
                    uint32_t cur = ld_rlx(&a.glob->next_root);
                    while (cur < avail) {
                        if (__hip_atomic_compare_exchange_strong(&a.glob->next_root, &cur, cur + 1u, __ATOMIC_RELAXED,
                                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                            idx = cur;
                            break;
                        }
                    }
                    if (idx >= avail) fetch_add(&a.glob->counter, kWaitOne);
                }
                idx = uni(idx);
                if (idx < avail) {
                    // the host wrote the ring entry, descriptor and graph to
                    // pinned memory before publishing the tail
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
                    uint32_t slot = 0;
                    if (lane_id() == 0)
                        slot = __hip_atomic_load(&a.req_ring[idx & a.ring_mask], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_SYSTEM);
                    r.kind = 2;
                    r.root = uni(slot);
                    load_request(a, r.root);
                    return r;
                }
            }
        }
        unsigned long long c = 0;
        if (lane_id() == 0) c = ld_rlx(&a.glob->counter);
        uint32_t lo = uni((uint32_t)c), hi = uni((uint32_t)(c >> 32));
        uint32_t waiting = lo & 0xFFFFu, reg = lo >> 16;
        // batch: unclaimed roots are in `hi`; serving: no root is coming once
        // the host has shut down and every published request is taken
        bool no_more_roots = true;
        if constexpr (SERVE)
            no_more_roots = uni(ld_rlx(&a.glob->host_shutdown)) != 0u &&
                            uni(ld_rlx(&a.glob->next_root)) >= roots_available<SERVE>(a);
        if (waiting == reg && hi == 0u && no_more_roots) {
            if (lane_id() == 0) st_rlx(&a.glob->done_flag, 1u);
            return r;
        }
        if (uni(ld_rlx(&a.glob->done_flag)) || uni(ld_rlx(&a.glob->cancel))) return r;
        if (lane_id() == 0) late = past_deadline(a);
        if (uni(late)) return r;
        backoff_idle(++spins, a.idle_cap, waiting);
    }
}

// Staged start.  A launch with few roots (one graph, a handful) cannot use
// thousands of workers at first, and idle workers are not free: every one
// polls the launch-wide words, and a 5,000-node search took 4.2 ms on 7,168
// workers but 0.5 ms on 256.  So only workers wid < active0 start at once;
// the others sleep, then look at the pool: once the launch has run for
// 8 x dormant_ticks they join (register and look for work) while at most
// one in eight registered workers waits, and otherwise sleep again, for
// max(dormant_ticks, elapsed / 8) with a jitter.  In between, every
// dormant_ticks / 4 (jittered), they only look whether the launch is over
// (its own cache line, not the pool's counter) and leave if so: a search
// that ends early then waits about dormant_ticks / 4 for them instead of up
// to 1.5 x max(dormant_ticks, elapsed / 8).  Measured: BA n = 400 MVC
// (29 nodes) 0.34 -> 0.22 ms, 3-regular n = 120 MVC 10-14% faster, the
// one-graph-per-launch solve of 256 batch graphs 123 -> 128 graphs/s.  The
// join decisions keep their schedule: reading the pool's counter as often
// as this made a 36k-node search 20% slower (contention on that line, and
// workers joining where they only added more).  A search that runs longer
// or widens later still gets every worker.  Only the staged kernels contain
// it: in the batch kernel even this prologue, inlined, shifted register
// allocation enough to cost 0.7% of a batched step (as a call, 11%).
__device__ bool dormant_join(const SolveArgs &a, uint32_t wid)
{
    const uint64_t start = wall_clock64();
    uint64_t interval = a.dormant_ticks;
    uint32_t h = wid * 2654435761u;
    uint64_t next = start + (interval * (128u + (h >> 24))) / 256u;
    const uint64_t quantum = a.dormant_ticks / 4u;
    uint64_t look = start + quantum;
    for (;;) {
        for (;;) {
            const uint64_t now = wall_clock64();
            if ((int64_t)(now - next) >= 0) break;
            if ((int64_t)(now - look) >= 0) {
                uint32_t over = 0;
                if (lane_id() == 0) over = ld_rlx(&a.glob->done_flag) | ld_rlx(&a.glob->cancel);
                if (uni(over)) return false;
                h = h * 1664525u + 1013904223u;
                look = now + (quantum * (128u + (h >> 24))) / 256u;
            }
            __builtin_amdgcn_s_sleep(127);
        }
        unsigned long long c = 0;
        uint32_t over = 0;
        if (lane_id() == 0) {
            c = ld_rlx(&a.glob->counter);
            over = ld_rlx(&a.glob->done_flag) | ld_rlx(&a.glob->cancel);
        }
        if (uni(over)) return false;
        const uint32_t lo = uni((uint32_t)c), enq = uni((uint32_t)(c >> 32));
        const uint32_t waiting = lo & 0xFFFFu, reg = lo >> 16;
        const uint64_t now = wall_clock64();
        const uint64_t elapsed = now - start;
        if (elapsed >= 8u * a.dormant_ticks && (8u * waiting <= reg || enq > waiting)) return true;
        if (elapsed / 8u > interval) interval = elapsed / 8u;
        h = h * 1664525u + 1013904223u;
        next = now + (interval * (128u + (h >> 24))) / 256u;
    }
}

// Copies count 32-bit words from HBM into 16-bit LDS entries with kInFlight
// loads per lane issued before the first store: a loop of one load and one store
// per lane waits a full memory round trip for every 64 entries (BA n = 2000,
// m = 3, 12k adjacency entries: a PVC answer 0.485 -> 0.465 ms and its MVC
// solve 0.77 -> 0.70 ms with 8 in flight).
template <uint32_t kInFlight>
__device__ __forceinline__ void copy_words_to_u16(uint16_t *dst, const uint32_t *src, uint32_t count)
{
    for (uint32_t i0 = 0; i0 < count; i0 += kInFlight * kWave) {
        uint32_t x[kInFlight];
#pragma unroll
        for (uint32_t j = 0; j < kInFlight; ++j) {
            const uint32_t i = i0 + j * kWave + lane_id();
            x[j] = i < count ? src[i] : 0u;
        }
#pragma unroll
        for (uint32_t j = 0; j < kInFlight; ++j) {
            const uint32_t i = i0 + j * kWave + lane_id();
            if (i < count) dst[i] = (uint16_t)x[j];
        }
    }
}

// A graph's descriptor with every field made wave-uniform (scalar
// registers): loop bounds such as n then compile to scalar compares instead
// of per-lane compares with exec-mask bookkeeping.
__device__ __forceinline__ GDesc load_desc(const GDesc *p)
{
    GDesc d = *p;
    d.n = uni(d.n);
    d.m = uni(d.m);
    d.off_src = uni(d.off_src);
    d.off_dst = uni(d.off_dst);
    d.off_deg = uni(d.off_deg);
    d.off_cover = uni(d.off_cover);
    d.off_bits = uni(d.off_bits);
    d.instance = uni(d.instance);
    d.k = uni(d.k);
    d.U = uni(d.U);
    d.seq = uni(d.seq);
    d.max_nodes = ((uint64_t)uni((uint32_t)(d.max_nodes >> 32)) << 32) | uni((uint32_t)d.max_nodes);
    return d;
}

// ---------------------------------------------------------------------------
// the persistent solve kernel
// ---------------------------------------------------------------------------

// The kernel body takes the launch arguments by reference from a one-line
// kernel: hipcc then allocates 55 instead of 81 VGPRs for the int16 LDS
// variant (same source otherwise), so the LDS budget alone sets occupancy.
template <typename DegT, bool HBM, bool SERVE, bool OFF16, bool STAGED, bool ADJ = false>
__device__ __forceinline__ void solve_body(const SolveArgs &a)
{
    extern __shared__ int4 smem4[];
    const uint32_t wib = threadIdx.x / kWave;
    const uint32_t wid = blockIdx.x * (blockDim.x / kWave) + wib;
    if (wid >= a.workers) return;
    if constexpr (STAGED)
        if (wid >= a.active0 && !dormant_join(a, wid)) return;

    // node storage: node state (npad * node_bytes_per_vertex) | list [lcap u32] | mark [mwords u32] |
    // xs [64 i32] | cand [npad]; a stack or worklist slot holds the node state
    char *store = HBM ? a.hbm_nodes + (size_t)wid * a.worker_bytes
                      : reinterpret_cast<char *>(smem4) + (size_t)wib * a.worker_bytes;
    const uint32_t slot_bytes = a.npad * node_bytes_per_vertex<DegT>();
    static_assert(!ADJ || OFF16, "the LDS adjacency sits after the LDS offsets");
    using NodeT = Node<DegT, HBM, OFF16, ADJ>;
    NodeT w;
    w.base = store;
    if constexpr (NodeT::kPacked) {
        w.dw = reinterpret_cast<uint32_t *>(store);
        w.deg32 = nullptr;
        w.s1w = nullptr;
        w.s2 = reinterpret_cast<typename NodeT::s2_t *>(store + (size_t)a.npad * 4u);
    } else {
        w.dw = nullptr;
        w.deg32 = reinterpret_cast<int32_t *>(store);
        w.s1w = reinterpret_cast<uint32_t *>(store + (size_t)a.npad * 4u);
        w.s2 = reinterpret_cast<typename NodeT::s2_t *>(store + (size_t)a.npad * 8u);
    }
    w.list = reinterpret_cast<typename NodeT::cand_t *>(store + slot_bytes);
    w.mark = reinterpret_cast<uint32_t *>(w.list + a.lcap);  // lcap is a multiple of 4
    w.xs = reinterpret_cast<int *>(w.mark + a.mwords);
    w.cand = reinterpret_cast<typename NodeT::cand_t *>(w.xs + kWave);
    w.cand3 = w.cand + a.ncap;
    w.ncap = a.ncap;
    w.src = nullptr;
    w.off16 = OFF16 ? reinterpret_cast<uint16_t *>(w.cand3 + a.ncap) : nullptr;
    // the adjacency follows the npad + 2 offset entries
    w.adj16 = ADJ ? w.off16 + a.npad + 2u : nullptr;
    w.dst = nullptr;
    w.bits = nullptr;
    w.n = 0;
    w.lcap = a.lcap;
    for (uint32_t i = lane_id(); i < a.mwords; i += kWave) w.mark[i] = 0u;
    w.sync();

    char *stk = static_cast<char *>(a.stk) + (size_t)wid * a.depth * slot_bytes;
    uint2 *stk_meta = a.stk_meta + (size_t)wid * a.depth;

    // register as a worker that waits for work, in one atomic; rid = the
    // number of workers registered before this one
    uint32_t rid = 0;
    if (lane_id() == 0) {
        rid = (uint32_t)(fetch_add(&a.glob->counter, kRegOne + kWaitOne) >> 16) & 0xFFFFu;
        // the time limit counts from the first worker's start, not from the
        // launch: a persistent kernel can wait for CUs held by another one
        uint64_t zero = 0;
        if (ld_rlx(&a.glob->t0) == 0u)
            __hip_atomic_compare_exchange_strong(&a.glob->t0, &zero, (uint64_t)wall_clock64(),
                                                 __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    rid = uni(rid);
    bool counted = true;  // the first acquire: already counted as waiting

    uint32_t g = kNone;
    GDesc gd = {};
    int top = -1, bot = 0;  // local stack entries bot..top (bot moves up on donation)
    bool have = false;
    uint32_t d = 0;
    uint32_t path = 0, descend = 0;
    Cands c;            // rule candidates listed since the last passes
    uint32_t off16_g = kNone;  // graph whose offsets the LDS copy holds
    uint64_t pending = 0, total_nodes = 0, idle_ticks = 0, donations = 0, steals = 0;
    // slot copies of the current graph's node state since the last flush:
    // stack pushes and pops, steals (ring slot in) and donations (stack slot
    // to ring slot, two); flush_nodes adds their HBM bytes to the launch total
    uint32_t copies = 0;
    uint32_t flush_every = kFlushEvery;  // see flush_interval

    for (;;) {
        if (!have) {
            if (top < bot) {
                top = -1;
                bot = 0;
            }
            if (top >= 0) {
                // pop (popStack, src/stack.cuh:10-20), then build the stored
                // right child
                const uint32_t ts = (uint32_t)top % a.depth;
                const uint2 m = stk_meta[ts];  // in flight during the copy
                if constexpr (NodeT::kPacked && !HBM)
                    copy_state_to_lds<NodeT>(w.base, stk + (size_t)ts * slot_bytes, a.npad, w.n);
                else
                    copy_state<NodeT>(w.base, stk + (size_t)ts * slot_bytes, a.npad, w.n);
                ++copies;
                w.sync();
                d = uni(m.x);
                uint32_t v = uni(m.y);
                --top;
                cands_reset(c, false);
                d += take_neighbours(w, v, c);
                have = true;
            } else {
                if (g != kNone) {
                    total_nodes += pending;
                    flush_nodes<STAGED>(a, gd, g, pending, copies, copy_bytes<NodeT>(w.n), flush_every, true);
                    finish_unit<SERVE>(a, g, gd);
                    g = kNone;
                }
                if (lane_id() == 0 && ld_rlx(&a.glob->cancel)) st_rlx(&a.glob->done_flag, 1u);
                const uint64_t t_idle = wall_clock64();
                Acquired q = acquire<SERVE>(a, w, slot_bytes, counted, rid);
                counted = false;
                idle_ticks += wall_clock64() - t_idle;
                if (q.kind == 0) break;
                steals += q.kind == 1;
                if (q.kind == 1) {
                    g = q.g;
                    copies = 1u;
                } else {
                    g = a.version == 1 ? q.root : (q.root >> a.start_depth);
                }
                gd = load_desc(a.gdesc + g);
                {
                    uint32_t fe = 0;
                    if (lane_id() == 0)
                        fe = flush_interval<STAGED>(a, gd, g, gd.max_nodes ? ld_rlx(&a.gstate[g].nodes) : 0u);
                    flush_every = uni(fe);
                }
                w.n = gd.n;
                w.src = a.gbuf + gd.off_src;
                if (OFF16 && g != off16_g) {
                    // a graph this worker has not held last: its offsets into LDS
                    if constexpr (ADJ) {
                        // and its adjacency (2m entries, vertex ids below 2^15)
                        const uint32_t e2 = uni(w.src[w.n]);
                        copy_words_to_u16<8>(w.off16, w.src, w.n + 1u);
                        copy_words_to_u16<8>(w.adj16, a.gbuf + gd.off_dst, e2);
                    } else {
                        for (uint32_t i = lane_id(); i <= w.n; i += kWave) w.off16[i] = (uint16_t)w.src[i];
                    }
                    w.sync();
                    off16_g = g;
                }
                w.dst = a.gbuf + gd.off_dst;
                w.bits = gd.off_bits ? a.gbuf + gd.off_bits : nullptr;
                if (q.kind == 1) {
                    d = q.d;
                    cands_reset(c, q.v == kNone);
                    if (q.v != kNone) d += take_neighbours(w, q.v, c);
                    descend = 0;
                } else {
                    init_root(w, reinterpret_cast<const int32_t *>(a.gbuf + gd.off_deg), a.npad);
                    cands_reset(c, true);
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
        // the stop flag and the incumbent in flight together (one round
        // trip to L2 per node instead of two)
        const uint32_t stop_now = ld_rlx(&gs->stop);
        const uint32_t best_now = (uint32_t)ld_rlx(&gs->best);
        if (uni(stop_now)) {
            // graph finished by someone else (PVC found, budget): drop local work
            top = -1;
            bot = 0;
            have = false;
            continue;
        }
        if (++pending >= flush_every) {
            total_nodes += pending;
            flush_nodes<STAGED>(a, gd, g, pending, copies, copy_bytes<NodeT>(w.n), flush_every, false);
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
        uint32_t T = mvc ? uni(best_now) : gd.k + 1u;
        Summary s;
        d = reduce_node(w, d, T, &s, c);

        // the live incumbent (another worker may have lowered it while the
        // rules ran with T); the bound uses both (bound_prunes)
        const uint32_t B = mvc ? uni((uint32_t)ld_rlx(&gs->best)) : T;
        if (bound_prunes(s.edges, d, T, B)) {
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
                take_vertex(w, v, c);
                d += 1;
            } else {
                d += take_neighbours(w, v, c);
            }
            continue;
        }
        // Branch: the right child (N(v) into the cover) is stored as its
        // parent state plus v and built when it is popped.
        // The local stack is a ring of `depth` slots holding entries
        // bot..top.  With depth = min(n, max(U, k+1)) + 1 it never fills
        // (an entry pushed at cover size d sits at index <= d); a smaller
        // configured depth hands the oldest entry to the worklist instead.
        if (top - bot + 1 >= (int)a.depth) {
            bool moved = false;
            if (a.version == 1) {
                for (uint32_t spins = 0;; ++spins) {
                    const uint32_t bs = (uint32_t)bot % a.depth;
                    uint2 m = stk_meta[bs];
                    if (try_enqueue(a, w, stk + (size_t)bs * slot_bytes, g, uni(m.x), uni(m.y), slot_bytes)) {
                        ++bot;
                        ++donations;
                        copies += 2u;
                        moved = true;
                        break;
                    }
                    uint32_t late = 0;
                    if (lane_id() == 0) late = ld_rlx(&a.glob->cancel) || ld_rlx(&gs->stop) || past_deadline(a);
                    if (uni(late)) break;
                    backoff(spins < 4 ? spins : 4);
                }
            }
            if (!moved) {
                // hybrid: the wait ended because the graph stopped or the
                // launch is cancelled; stack-only: the depth is exact, so
                // this is an internal error
                if (a.version != 1 && lane_id() == 0) st_rlx(&a.glob->error, kErrStackOverflow);
                have = false;
                continue;
            }
        }
        ++top;
        const uint32_t ts = (uint32_t)top % a.depth;
        copy_state<NodeT>(stk + (size_t)ts * slot_bytes, w.base, a.npad, w.n);
        if (lane_id() == 0) stk_meta[ts] = make_uint2(d, v);
        ++copies;
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
                const uint32_t bs = (uint32_t)bot % a.depth;
                uint2 m = stk_meta[bs];
                if (try_enqueue(a, w, stk + (size_t)bs * slot_bytes, g, uni(m.x), uni(m.y), slot_bytes)) {
                    ++bot;
                    ++donations;
                    copies += 2u;
                }
            }
        }
        take_vertex(w, v, c);
        d += 1;
    }
    if (lane_id() == 0) {
        WorkerStats ws;
        ws.nodes = total_nodes;
        ws.idle_ticks = idle_ticks;
        ws.donations = donations;
        ws.steals = steals;
        ws.exit_tick = wall_clock64();
        a.wstats[wid] = ws;
        if constexpr (SERVE) {
            // a server's statistics and error word live in pinned host
            // memory: stopping a server then needs no copy, which could wait
            // for another persistent kernel on the device
            const uint32_t e = ld_rlx(&a.glob->error);
            if (e) __hip_atomic_store(&a.host->error, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Per-launch initialisation with plain stores (no memset nodes, so the whole
// solve can be captured in a HIP graph).  units = outstanding subtrees per
// graph at start: 1 root (hybrid) or 2^start_depth paths (static split).
// Results of a solve, copied by a kernel into mapped pinned host memory.
// hipMemcpyAsync would do the same, but such a copy has been measured to
// wait until a server's persistent kernel on the device ends, while a kernel
// on the same stream is scheduled beside it like the solve itself.
struct Publish {
    const uint32_t *src[4];
    uint32_t *dst[4];
    uint32_t words[4];
};

__global__ void publish_kernel(Publish p)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (int k = 0; k < 4; ++k)
        for (uint32_t i = tid; i < p.words[k]; i += stride) p.dst[k][i] = p.src[k][i];
}

// One graph's outcome from its search state after the launch (glob = the
// launch's counters), for the host read-back of vc_solve and the device
// read-back of vc_solve_noalloc alike.
__host__ __device__ inline vc_result result_of(const GState &s, uint32_t instance, const Global &glob,
                                               uint32_t wall_khz)
{
    vc_result r;
    r.minimum = 0;
    r.k_found = 0;
    r.solved = 0;
    r.cover_from_gpu = 0;
    r.tree_nodes = s.nodes;
    r.done_us = 0;
    const bool finished = s.done != 0u && !glob.error;
    if (instance == VC_MVC) {
        r.minimum = (uint32_t)s.best;
        r.solved = finished && !s.budget_hit;
        r.cover_from_gpu = s.cover_size >= 0 && s.cover_size == s.best;
    } else {
        r.k_found = s.found;
        r.minimum = s.found ? (uint32_t)s.cover_size : 0u;
        r.solved = finished && (s.found || !s.budget_hit);
        r.cover_from_gpu = s.found != 0u;
    }
    r.status = glob.error ? VC_EDEVICE
                          : (r.solved ? VC_OK : (glob.cancel && !glob.timed_out ? VC_ECANCELLED : VC_ETIMEOUT));
    if (s.done && s.done_tick >= glob.t0 && glob.t0 != 0u && wall_khz > 0) {
        const uint64_t us = (s.done_tick - glob.t0) * 1000u / (uint64_t)wall_khz;
        r.done_us = (uint32_t)(us < 0xFFFFFFFFull ? us : 0xFFFFFFFFull);
    }
    return r;
}

// Resets a launch's scratch.  Batch launches (preload_roots = 1) start
// with every root counted as enqueued in the termination word (acquire).
__global__ void init_kernel(SolveArgs a, uint32_t units, uint32_t preload_roots)
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
        Global z = {};  // t0 = 0: the first solve worker to start sets it
        if (preload_roots) z.counter = (unsigned long long)a.roots * kEnqOne;
        *a.glob = z;
    }
}

// Read-back of vc_solve_noalloc, written inside its fixed scratch: launch
// totals, one vc_result per graph, and every cover bitmap expanded to one
// byte per vertex (graph g's bytes start at 32 x its first bitmap word).
struct NoAllocHead {
    unsigned long long nodes, idle_ticks, donations, steals, last_exit, t0, state_bytes;
    uint32_t error, timed_out, cancel, pad[15];
};
static_assert(sizeof(NoAllocHead) == 128, "NoAllocHead is two 64-byte lines");

// One workgroup of 256 threads sums the worker statistics; every thread
// converts results and expands cover words.
__global__ __launch_bounds__(256) void noalloc_results_kernel(SolveArgs a, uint32_t wall_khz, uint32_t cover_words,
                                                              NoAllocHead *head, vc_result *out, uint8_t *cover_bytes)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t g = tid; g < a.num_graphs; g += stride)
        out[g] = result_of(a.gstate[g], a.gdesc[g].instance, *a.glob, wall_khz);
    for (uint32_t i = tid; i < cover_words * 32u; i += stride)
        cover_bytes[i] = (uint8_t)((a.covers[i >> 5] >> (i & 31u)) & 1u);
    if (blockIdx.x != 0) return;
    // sums of nodes, idle ticks, donations, steals, state bytes; max of exit ticks
    __shared__ unsigned long long part[6][256];
    unsigned long long acc[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t w = threadIdx.x; w < a.workers; w += 256u) {
        const WorkerStats &x = a.wstats[w];
        acc[0] += x.nodes;
        acc[1] += x.idle_ticks;
        acc[2] += x.donations;
        acc[3] += x.steals;
        acc[5] = x.exit_tick > acc[5] ? x.exit_tick : acc[5];
    }
    for (int k = 0; k < 6; ++k) part[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (uint32_t h = 128u; h > 0u; h >>= 1) {
        if (threadIdx.x < h) {
            for (int k = 0; k < 5; ++k) part[k][threadIdx.x] += part[k][threadIdx.x + h];
            const unsigned long long o = part[5][threadIdx.x + h];
            if (o > part[5][threadIdx.x]) part[5][threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        NoAllocHead r = {};
        r.nodes = part[0][0];
        r.idle_ticks = part[1][0];
        r.donations = part[2][0];
        r.steals = part[3][0];
        r.state_bytes = a.glob->state_bytes;
        r.last_exit = part[5][0];
        r.t0 = a.glob->t0;
        r.error = a.glob->error;
        r.timed_out = a.glob->timed_out;
        r.cancel = a.glob->cancel;
        *head = r;
    }
}

template <typename DegT, bool HBM, bool SERVE, bool STAGED = false>
__global__ __launch_bounds__(1024) void solve_kernel(SolveArgs a)
{
    solve_body<DegT, HBM, SERVE, false, STAGED>(a);
}

template __global__ void solve_kernel<int16_t, false, false>(SolveArgs);
// Batch launches whose graphs all have 2m < 2^16 keep each worker's current
// srcPtr in LDS (OFF16).  Asking for 7 waves per SIMD caps the body at 72
// VGPRs (about 40 spilled, mostly off the per-node path); with 64-entry
// candidate lists a worker's LDS at n = 400 lets 28 workers share a CU.
// Measured per batched step: 2.87 s at 7 waves against 3.16 s at 6 (80
// VGPRs) and 3.38 s at 8 (64 VGPRs, 68 spilled).
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(7))) void solve_kernel_off16(SolveArgs a)
{
    solve_body<int16_t, false, false, true, false>(a);
}
// The same with a staged start (launches with few roots, dormant_join).
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(7))) void solve_kernel_off16_staged(SolveArgs a)
{
    solve_body<int16_t, false, false, true, true>(a);
}
// Staged, with each worker's current graph's adjacency in LDS as well: for
// launches of few graphs whose adjacency fits beside the node state (small
// searches are chains of dependent rule passes, each waiting for adjacency
// reads; from LDS those take ~100 cycles instead of an L2 or MALL trip).
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(7))) void solve_kernel_off16_adj(SolveArgs a)
{
    solve_body<int16_t, false, false, true, true, true>(a);
}
template __global__ void solve_kernel<int16_t, true, false>(SolveArgs);
template __global__ void solve_kernel<int32_t, true, false>(SolveArgs);
template __global__ void solve_kernel<int16_t, false, true>(SolveArgs);
template __global__ void solve_kernel<int16_t, true, true>(SolveArgs);
template __global__ void solve_kernel<int32_t, true, true>(SolveArgs);
template __global__ void solve_kernel<int16_t, false, false, true>(SolveArgs);
template __global__ void solve_kernel<int16_t, true, false, true>(SolveArgs);
template __global__ void solve_kernel<int32_t, true, false, true>(SolveArgs);

}  // namespace vcg

// Launch scratch of vc_solve_noalloc: stacks, worklist, counters, worker
// statistics and the read-back area, in device memory that the library
// holds from the moment its code is loaded on a GPU until the process ends
// (one copy per GPU, at an address that hipGetSymbolAddress gives per
// device).  Host code never passes this name to a kernel or a copy: its
// host-side symbol is not a device address.
constexpr size_t kNoAllocScratchBytes = 128ull << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

// ===========================================================================
// host launcher and C ABI (include/vcgpu.h)
// ===========================================================================

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "vcgpu.h"

using vcg::GDesc;
using vcg::GState;
using vcg::Global;
using vcg::SolveArgs;

struct vc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;  // solve start / end / results on host
    int cu_count = 0;
    int wall_khz = 100000;  // device wall-clock rate
    size_t total_mem = 0;
    size_t lds_per_cu = 0;
    // launch scratch, grown on demand and reused across solves
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // autotuner cache: (vertex-count bucket, density bucket) -> chosen params
    std::map<std::pair<int, int>, vc_params> tuned;
    // hipFree / hipHostFree wait for the whole device, including a running
    // server's persistent kernel; while a server runs, frees are deferred
    // until the last one stops (or the context is destroyed).
    std::mutex mem_mu;
    int active_servers = 0;
    std::vector<std::pair<void *, bool>> deferred;  // (pointer, pinned host memory)
};

void free_now(void *p, bool host)
{
    if (host)
        (void)hipHostFree(p);
    else
        (void)hipFree(p);
}

// Frees p now, or later if a server of this context is running.
void ctx_release(vc_ctx *ctx, void *p, bool host)
{
    if (!p) return;
    if (ctx) {
        std::lock_guard<std::mutex> lk(ctx->mem_mu);
        if (ctx->active_servers > 0) {
            ctx->deferred.emplace_back(p, host);
            return;
        }
    }
    free_now(p, host);
}

// Releases the deferred frees once no server runs.
void ctx_flush_deferred(vc_ctx *ctx)
{
    std::vector<std::pair<void *, bool>> d;
    {
        std::lock_guard<std::mutex> lk(ctx->mem_mu);
        if (ctx->active_servers > 0) return;
        d.swap(ctx->deferred);
    }
    for (auto &x : d) free_now(x.first, x.second);
}

namespace {

// vc_solve_noalloc's scratch per device: the device address of `scratch` on
// each GPU, taken with that device current (when the first context on the
// device is created, else at the first vc_solve_noalloc call there), and a
// lock that serialises this process's calls sharing that GPU's array.
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxDevices] = {};
std::mutex g_noalloc_call_mu[kMaxDevices];

// The calling thread's current device must be `device`.  Returns the
// device address of that GPU's scratch, or nullptr if it cannot be found.
void *noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess && p) g_noalloc_ptr[device] = p;
    }
    return g_noalloc_ptr[device];
}

}  // namespace

// What a captured dispatch depends on; a different value means re-capture.
struct DispatchKey {
    int variant;
    uint32_t num_blocks, block_dim, version, start_depth, qsize, qthreshold, depth, want_covers;
    uint32_t off16, adj, active0, roots, flush_share;  // kernel choice and staged start
    size_t smem;
    uint64_t limit_ticks;
    const void *scratch;
    const void *h_ws;
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
    double mean_n = 0, mean_density = 0;  // autotuner bucket
    // pinned readback buffers (stable addresses, so a captured dispatch can
    // copy into them on every replay)
    GState *h_state = nullptr;
    Global *h_glob = nullptr;
    vcg::WorkerStats *h_ws = nullptr;
    uint32_t ws_cap = 0;
    uint32_t *h_cov = nullptr;
    // captured dispatch: init kernel + solve kernel + readback copies
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DispatchKey key{};
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
    // undirected: every listed edge is listed from its other end too (the
    // neighbour sketch and the degree updates rely on it)
    for (uint32_t v = 0; v < n; ++v)
        for (uint32_t x = g.srcPtr[v]; x < g.srcPtr[v + 1]; ++x) {
            const uint32_t u = g.dst[x];
            if (!std::binary_search(g.dst + g.srcPtr[u], g.dst + g.srcPtr[u + 1], v)) return false;
        }
    return true;
}

// Kernel variants: degree type x node storage.  int16 degrees (32-bit
// sketch sums) while every graph has n < 32768, else int32 (64-bit square
// sums); node storage in LDS when a worker fits, else in HBM.  int32 nodes
// need at least 512 KB at n = 32768, so they are never LDS-resident.
enum Variant { kLds16 = 0, kHbm16 = 1, kHbm32 = 2 };

const void *kernel_of(int variant, bool serve, bool off16 = false, bool adj = false)
{
    if (adj) return reinterpret_cast<const void *>(vcg::solve_kernel_off16_adj);
    if (serve) {
        switch (variant) {
        case kLds16: return reinterpret_cast<const void *>(vcg::solve_kernel<int16_t, false, true>);
        case kHbm16: return reinterpret_cast<const void *>(vcg::solve_kernel<int16_t, true, true>);
        default: return reinterpret_cast<const void *>(vcg::solve_kernel<int32_t, true, true>);
        }
    }
    switch (variant) {
    case kLds16:
        return off16 ? reinterpret_cast<const void *>(vcg::solve_kernel_off16)
                     : reinterpret_cast<const void *>(vcg::solve_kernel<int16_t, false, false>);
    case kHbm16: return reinterpret_cast<const void *>(vcg::solve_kernel<int16_t, true, false>);
    default: return reinterpret_cast<const void *>(vcg::solve_kernel<int32_t, true, false>);
    }
}

struct Plan {
    int variant;
    uint32_t npad, lcap, ncap, mwords, worker_bytes, deg_bytes, slot_bytes, wpb, block_dim, blocks_per_cu,
        num_blocks, workers;
    uint32_t depth, qsize, qthreshold, roots, units;
    uint32_t off16;  // workers keep their graph's srcPtr in LDS as uint16
    uint32_t adj;    // ... and its dst (solve_kernel_off16_adj, a staged launch)
    size_t smem;
    size_t off_glob, off_stk, off_meta, off_qdeg, off_qmeta, off_tickets, off_wnodes, off_hbm, total;
};

// entries per rule-candidate list: small enough that 28 workers' LDS fits a
// CU at n = 400 (a longer run of candidates makes the next pass scan all)
constexpr uint32_t kCandCap = 64;
// The same for plans whose occupancy the lists do not set: node state in
// HBM, or the adjacency in LDS (launches of few, often large graphs).  On
// BA n = 2,000 a hub's removal and the triangle rule's deletions list
// hundreds of leaf candidates at once, and with 64 entries every such leaf
// pass scanned all 2,000 vertices (80% of the search's busy time).
constexpr uint32_t kCandCapLarge = 1024;
constexpr uint32_t kListCap = 1024;  // marks per rule pass (the rest go to the next pass)
// graphs up to this many vertices get an adjacency bitmap (128 KB at most),
// while the graph set stays below 2^31 words
constexpr uint32_t kAdjBitsMaxN = 1024;
constexpr uint64_t kAdjBitsMaxWords = 1ull << 31;

// Per-worker node storage for a degree type: node state [npad] | list [lcap] | mark [mwords] |
// xs [64] | cand [ncap] | cand3 [ncap] | (off16) srcPtr [npad + 2] uint16 | (adjacency in LDS)
// dst [adj_entries] uint16.  A stack / worklist slot is the node state.
void size_storage(Plan &pl, uint32_t nmax, uint32_t deg_bytes, bool off16 = false, uint32_t adj_entries = 0,
                  uint32_t cand_cap = kCandCap)
{
    pl.deg_bytes = deg_bytes;
    pl.npad = round_up(nmax, 64u / deg_bytes);
    pl.slot_bytes = pl.npad * (deg_bytes == 2u ? vcg::node_bytes_per_vertex<int16_t>()
                                               : vcg::node_bytes_per_vertex<int32_t>());
    pl.lcap = std::min<uint32_t>(round_up(nmax, 4u), kListCap);
    // candidate lists: a longer run of candidates between two passes of a
    // rule makes that pass scan every vertex instead
    pl.ncap = std::min<uint32_t>(round_up(nmax, 4u), adj_entries ? kCandCapLarge : cand_cap);
    const uint32_t id_bytes = deg_bytes == 2u ? 2u : 4u;  // mark and candidate list entries
    pl.mwords = round_up((nmax + 31u) / 32u, 4u);
    const uint32_t cand_bytes = 2u * pl.ncap * id_bytes;
    pl.worker_bytes =
        round_up(pl.slot_bytes + id_bytes * pl.lcap + 4u * pl.mwords + 4u * vcg::kWave + cand_bytes +
                     (off16 ? 2u * (pl.npad + 2u) : 0u) + 2u * round_up(adj_entries, 4u),
                 64u);
}

// Launch shape shared by the batch solve and the server: node storage,
// workgroup, grid from the occupancy query, worklist, and the memory plan.
constexpr uint32_t kServeWorkersPerCU = 2;
constexpr uint32_t kMaxBlockDim = 1024;   // the solve kernels' launch bound
constexpr uint32_t kMaxWorkers = 0xFFFFu; // 16-bit worker counts in the termination word

vc_status plan_launch(vc_ctx *ctx, uint32_t nmax_in, uint32_t depth, const vc_params *p, Plan &pl, bool serve,
                      bool off16_ok = false, size_t mem_budget = 0, uint32_t adj_entries = 0)
{
    const uint32_t version = p->version;
    if (version != VC_HYBRID && version != VC_STACK_ONLY) return VC_EINVAL;
    if (p->use_global_memory < -1 || p->use_global_memory > 1) return VC_EINVAL;
    // -b: threads per workgroup, 64..1024 as the reference accepts
    // (src/auxFunctions.cpp:409-444); a worker is one wave, so a workgroup
    // holds block_dim / 64 independent workers
    uint32_t block_dim = p->block_dim ? p->block_dim : 256u;
    if (block_dim % vcg::kWave != 0 || block_dim > kMaxBlockDim) return VC_EINVAL;
    if (version == VC_STACK_ONLY && p->start_depth > 20u) return VC_EINVAL;
    const uint32_t nmax = std::max<uint32_t>(nmax_in, 1u);
    const size_t lds_limit = ctx->lds_per_cu ? ctx->lds_per_cu : 160u * 1024u;

    // Node storage in LDS when the workgroup's workers fit in one CU's LDS
    // (an automatic workgroup shrinks until they do), otherwise in HBM, as
    // the reference falls back to its _global_ kernels (-g 1 forces HBM,
    // -g 0 forbids it); int16 degrees while n < 32768 (VC_FLAG_WIDE forces
    // int32).
    const bool wide = nmax >= 32768u || (p->flags & VC_FLAG_WIDE) != 0u;
    size_storage(pl, nmax, wide ? 4u : 2u);
    uint32_t wpb = block_dim / vcg::kWave;
    bool hbm = wide || p->use_global_memory == 1;
    if (!hbm) {
        if (!p->block_dim)
            while (wpb > 1 && (size_t)wpb * pl.worker_bytes > lds_limit) wpb /= 2;
        hbm = (size_t)wpb * pl.worker_bytes > lds_limit;
    }
    if (hbm && p->use_global_memory == 0) return VC_ENOTSUP;
    pl.variant = wide ? kHbm32 : hbm ? kHbm16 : kLds16;
    if (hbm) size_storage(pl, nmax, wide ? 4u : 2u, false, 0u, kCandCapLarge);
    // LDS copy of the current graph's offsets (batch launches whose graphs
    // all have 2m < 2^16; a server reuses graph slots for new requests)
    pl.off16 = 0u;
    pl.adj = 0u;
    if (!serve && off16_ok && pl.variant == kLds16) {
        Plan t = pl;
        size_storage(t, nmax, 2u, true);
        if ((size_t)wpb * t.worker_bytes <= lds_limit) {
            pl = t;
            pl.off16 = 1u;
        }
        // and the adjacency (adj_entries = the largest 2m of the launch),
        // when asked for; an automatic workgroup shrinks to fit
        if (pl.off16 && adj_entries) {
            Plan ta = pl;
            size_storage(ta, nmax, 2u, true, adj_entries);
            uint32_t wa = wpb;
            if (!p->block_dim)
                while (wa > 1 && (size_t)wa * ta.worker_bytes > lds_limit) wa /= 2;
            if ((size_t)wa * ta.worker_bytes <= lds_limit) {
                pl = ta;
                pl.off16 = 1u;
                pl.adj = 1u;
                wpb = wa;
            }
        }
    }
    pl.wpb = wpb;
    pl.block_dim = wpb * vcg::kWave;
    pl.smem = hbm ? 0 : (size_t)pl.wpb * pl.worker_bytes;
    if (pl.smem > lds_limit) return VC_ENOTSUP;

    int bpc = 0;
    if (!hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kernel_of(pl.variant, serve, pl.off16 != 0u, pl.adj != 0u),
                                                             (int)pl.block_dim, pl.smem)))
        return VC_EDEVICE;
    if (bpc < 1) return VC_ENOTSUP;
    pl.blocks_per_cu = (uint32_t)bpc;
    uint32_t auto_blocks = (uint32_t)bpc * (uint32_t)ctx->cu_count;
    if (serve) {
        // A server is idle most of the time, and every idle wave polls the
        // launch-wide words.  kServeWorkersPerCU workers per CU served the
        // mixed request stream (1,000 requests/s) faster than the occupancy
        // grid (856 vs 554 graphs/s; small requests p50 0.26 vs 4.3 ms, p99
        // 5.8 vs 37 ms), and the CUs left free run other kernels -- with the
        // occupancy grid a vc_solve issued beside a server waited for it.
        auto_blocks = std::min<uint32_t>(
            auto_blocks, std::max<uint32_t>(1u, kServeWorkersPerCU * (uint32_t)ctx->cu_count / pl.wpb));
    }
    pl.num_blocks = p->num_blocks ? p->num_blocks : auto_blocks;
    // the termination word counts waiting and registered workers in 16-bit
    // fields (vcg::Global::counter)
    if ((uint64_t)pl.num_blocks * pl.wpb > kMaxWorkers) return VC_EINVAL;

    pl.depth = std::max<uint32_t>(depth, 1u);

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
    const size_t slot = pl.slot_bytes;
    const size_t budget = mem_budget ? mem_budget : ctx->total_mem / 10 * 7;
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
        pl.off_hbm = take(pl.variant != kLds16 ? (size_t)pl.workers * pl.worker_bytes : 0);
        pl.total = off;
        if (pl.total <= budget || pl.num_blocks <= 1) break;
        pl.num_blocks = std::max<uint32_t>(1u, pl.num_blocks / 2u);
    }
    if (pl.total > budget) return VC_ENOMEM;

    pl.units = version == VC_HYBRID ? 1u : (1u << p->start_depth);
    pl.roots = 0;
    return VC_OK;
}

vc_status make_plan(vc_ctx *ctx, const vc_graphs *set, const vc_job *jobs, const vc_params *p, Plan &pl,
                    size_t mem_budget = 0)
{
    // Stack depth: the stack never holds more entries than the cover size of
    // the current node, which stays below max(U, k+1) (prune at d >= U for
    // MVC, d > k for PVC), and never more than n.
    uint32_t need = 1;
    for (int i = 0; i < set->count; ++i) {
        uint32_t lim = jobs[i].instance == VC_MVC ? jobs[i].upper_bound : std::max(jobs[i].upper_bound, jobs[i].k + 1u);
        lim = std::min(lim, set->h_desc[i].n);
        need = std::max(need, lim);
    }
    bool off16_ok = true;
    uint32_t e2max = 0;
    for (int i = 0; i < set->count; ++i) {
        off16_ok &= 2ull * set->h_desc[i].m < 65536ull;
        e2max = std::max<uint32_t>(e2max, (uint32_t)std::min<uint64_t>(2ull * set->h_desc[i].m, 0xFFFFFFFFull));
    }
    vc_status st = plan_launch(ctx, set->nmax, need + 1u, p, pl, false, off16_ok, mem_budget);
    if (st != VC_OK) return st;
    // A launch of few roots on the automatic grid (the staged start) also
    // tries the adjacency in LDS (solve_kernel_off16_adj).  Taken when it
    // keeps the worker count, or for graphs of 1,024 vertices or more (no
    // adjacency bitmap: the triangle rule's pair test is a binary search of
    // the adjacency) -- a single BA n = 2,000 PVC solve, a chain of
    // dependent rule passes, ran 1.3 ms with global adjacency.
    const uint64_t roots_est = (uint64_t)set->count * pl.units;
    if (!p->num_blocks && pl.off16 && e2max > 0 && 2ull * roots_est < pl.workers) {
        Plan pa;
        if (plan_launch(ctx, set->nmax, need + 1u, p, pa, false, off16_ok, mem_budget, e2max) == VC_OK && pa.adj &&
            (pa.workers >= pl.workers || set->nmax >= kAdjBitsMaxN))
            pl = pa;
    }
    // roots (graphs x static paths) are counted in the termination word's
    // 32-bit enqueued field
    if ((uint64_t)set->count * pl.units >= (1ull << 31)) return VC_EINVAL;
    pl.roots = (uint32_t)set->count * pl.units;
    return VC_OK;
}

// Fills the launch arguments that depend only on the plan and the scratch base.
void fill_args(SolveArgs &a, const Plan &pl, char *base, uint32_t version, uint32_t start_depth, uint64_t limit_ticks)
{
    a.off16 = pl.off16;
    a.glob = reinterpret_cast<Global *>(base + pl.off_glob);
    a.stk = base + pl.off_stk;
    a.stk_meta = reinterpret_cast<uint2 *>(base + pl.off_meta);
    a.q_deg = base + pl.off_qdeg;
    a.q_meta = reinterpret_cast<uint4 *>(base + pl.off_qmeta);
    a.tickets = reinterpret_cast<uint32_t *>(base + pl.off_tickets);
    a.wstats = reinterpret_cast<vcg::WorkerStats *>(base + pl.off_wnodes);
    a.npad = pl.npad;
    a.lcap = pl.lcap;
    a.ncap = pl.ncap;
    a.mwords = pl.mwords;
    a.worker_bytes = pl.worker_bytes;
    a.depth = pl.depth;
    a.qsize = pl.qsize;
    a.qthreshold = pl.qthreshold;
    a.workers = pl.workers;
    a.version = version;
    a.start_depth = version == VC_STACK_ONLY ? start_depth : 0u;
    a.roots = pl.roots;
    a.limit_ticks = limit_ticks;
    a.hbm_nodes = pl.variant != kLds16 ? base + pl.off_hbm : nullptr;
    a.idle_cap = 6u;  // about 14 us between polls of an idle worker
}

template <bool SERVE>
void launch_variant(const Plan &pl, hipStream_t stream, const SolveArgs &a)
{
    if (!SERVE && pl.adj) {
        hipLaunchKernelGGL(vcg::solve_kernel_off16_adj, dim3(pl.num_blocks), dim3(pl.block_dim), pl.smem, stream, a);
        return;
    }
    if (!SERVE && a.active0 != 0u) {
        // staged start: the same kernels with the dormant_join prologue
        switch (pl.variant) {
        case kLds16:
            if (pl.off16)
                hipLaunchKernelGGL(vcg::solve_kernel_off16_staged, dim3(pl.num_blocks), dim3(pl.block_dim), pl.smem,
                                   stream, a);
            else
                hipLaunchKernelGGL((vcg::solve_kernel<int16_t, false, false, true>), dim3(pl.num_blocks),
                                   dim3(pl.block_dim), pl.smem, stream, a);
            break;
        case kHbm16:
            hipLaunchKernelGGL((vcg::solve_kernel<int16_t, true, false, true>), dim3(pl.num_blocks),
                               dim3(pl.block_dim), 0, stream, a);
            break;
        default:
            hipLaunchKernelGGL((vcg::solve_kernel<int32_t, true, false, true>), dim3(pl.num_blocks),
                               dim3(pl.block_dim), 0, stream, a);
            break;
        }
        return;
    }
    switch (pl.variant) {
    case kLds16:
        if (!SERVE && pl.off16)
            hipLaunchKernelGGL(vcg::solve_kernel_off16, dim3(pl.num_blocks), dim3(pl.block_dim), pl.smem, stream,
                               a);
        else
            hipLaunchKernelGGL((vcg::solve_kernel<int16_t, false, SERVE>), dim3(pl.num_blocks), dim3(pl.block_dim),
                               pl.smem, stream, a);
        break;
    case kHbm16:
        hipLaunchKernelGGL((vcg::solve_kernel<int16_t, true, SERVE>), dim3(pl.num_blocks), dim3(pl.block_dim), 0,
                           stream, a);
        break;
    default:
        hipLaunchKernelGGL((vcg::solve_kernel<int32_t, true, SERVE>), dim3(pl.num_blocks), dim3(pl.block_dim), 0,
                           stream, a);
        break;
    }
}

bool launch_solve(const Plan &pl, hipStream_t stream, const SolveArgs &a)
{
    if (a.serving)
        launch_variant<true>(pl, stream, a);
    else
        launch_variant<false>(pl, stream, a);
    return hip_ok(hipGetLastError());
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
    case VC_EAGAIN: return "request not finished yet";
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
        !hip_ok(hipEventCreate(&c->ev1)) || !hip_ok(hipEventCreate(&c->ev2))) {
        delete c;
        return VC_EDEVICE;
    }
    // the device address of vc_solve_noalloc's fixed scratch on this GPU
    // (a failure here is reported by that entry point, which tries again)
    (void)noalloc_scratch(device);
    *out = c;
    return VC_OK;
}

vc_status vc_ctx_destroy(vc_ctx *c)
{
    if (!c) return VC_EINVAL;
    (void)hipSetDevice(c->device);
    if (c->scratch) (void)hipFree(c->scratch);
    for (auto &x : c->deferred) free_now(x.first, x.second);
    c->deferred.clear();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return VC_OK;
}

vc_status vc_graphs_upload(vc_ctx *ctx, int count, const vc_csr *graphs, vc_graphs **out)
{
    if (!ctx || !out || count < 1 || !graphs) return VC_EINVAL;
    *out = nullptr;
    if (!hip_ok(hipSetDevice(ctx->device))) return VC_EDEVICE;
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
        // adjacency bitmap for the triangle rule's pair test (one load
        // instead of a binary search) while it stays small
        const uint64_t bit_words = ((uint64_t)g.vertexNum * g.vertexNum + 31u) / 32u;
        d.off_bits = 0;
        if (g.vertexNum <= kAdjBitsMaxN && words + bit_words < kAdjBitsMaxWords) {
            d.off_bits = (uint32_t)words;
            words += (bit_words + 3) / 4 * 4;
        }
        d.off_cover = cover_words;
        cover_words += round_up((g.vertexNum + 31u) / 32u, 2u);
        s->nmax = std::max(s->nmax, g.vertexNum);
        if (words >= 0xFFFFFFF0ull) {
            delete s;
            return VC_ENOMEM;
        }
    }
    s->cover_words = std::max<uint32_t>(cover_words, 2u);
    for (int i = 0; i < count; ++i) {
        const double n = graphs[i].vertexNum, m = graphs[i].edgeNum;
        s->mean_n += n / count;
        s->mean_density += (n > 1 ? 2.0 * m / (n * (n - 1)) : 0.0) / count;
    }
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
        if (d.off_bits)
            for (uint32_t v = 0; v < g.vertexNum; ++v)
                for (uint32_t x = g.srcPtr[v]; x < g.srcPtr[v + 1]; ++x) {
                    const uint64_t bit = (uint64_t)v * g.vertexNum + g.dst[x];
                    host[d.off_bits + (bit >> 5)] |= 1u << (bit & 31u);
                }
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
    if (ctx) (void)hipSetDevice(ctx->device);
    ctx_release(ctx, s->d_words, false);
    ctx_release(ctx, s->d_desc, false);
    ctx_release(ctx, s->d_covers, false);
    ctx_release(ctx, s->d_state, false);
    ctx_release(ctx, s->h_state, true);
    ctx_release(ctx, s->h_glob, true);
    ctx_release(ctx, s->h_ws, true);
    ctx_release(ctx, s->h_cov, true);
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    delete s;
    return VC_OK;
}

vc_status vc_occupancy(vc_ctx *ctx, const vc_graphs *set, const vc_params *params, int *blocks_per_cu, int *block_dim,
                       int *use_global_memory)
{
    if (!ctx || !set || !params) return VC_EINVAL;
    if (!hip_ok(hipSetDevice(ctx->device))) return VC_EDEVICE;
    std::vector<vc_job> jobs(set->count);
    for (int i = 0; i < set->count; ++i) {
        jobs[i] = vc_job{VC_MVC, 0u, set->h_desc[i].n, 0u, 0ull};
    }
    Plan pl;
    vc_status st = make_plan(ctx, set, jobs.data(), params, pl);
    if (st != VC_OK) return st;
    if (blocks_per_cu) *blocks_per_cu = (int)pl.blocks_per_cu;
    if (block_dim) *block_dim = (int)pl.block_dim;
    if (use_global_memory) *use_global_memory = pl.variant != kLds16;
    return VC_OK;
}

}  // extern "C"

namespace {

// Checks the jobs, plans the launch within mem_budget bytes of scratch (0 =
// 70% of device memory) and writes the jobs into the set's descriptors.
vc_status prepare_solve(vc_ctx *ctx, vc_graphs *set, const vc_job *jobs, const vc_params *params, Plan &pl,
                        size_t mem_budget)
{
    if (!hip_ok(hipSetDevice(ctx->device))) return VC_EDEVICE;
    for (int i = 0; i < set->count; ++i) {
        const vc_job &j = jobs[i];
        if (j.instance != VC_MVC && j.instance != VC_PVC) return VC_EINVAL;
        if (j.upper_bound > set->h_desc[i].n) return VC_EINVAL;
    }
    vc_status st = make_plan(ctx, set, jobs, params, pl, mem_budget);
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
    return VC_OK;
}

// Launch arguments of a batch solve of `set` with scratch at base.
// Workers that start at once in a staged launch: two per CU.  Kernel time
// with 128 / 256 / 512 starting workers (the rest joining as described at
// dormant_join; measured before its end-only look, which took the 512 case
// of the first graph to 0.64-0.66 ms): 3-regular n = 120 MVC (5,453 nodes) 0.73-0.74 / 0.66-0.67 /
// 0.72-0.77 ms; BA n = 2,000 PVC answer 0.57 / 0.57 / 0.46 ms; ER
// G(200, 0.05) MVC ~210 ms for all three.  256 helps only searches under a
// millisecond (four 3-regular n = 120 graphs: 12-15%) and hurts longer ones
// that still end before the others join (3-regular n = 160, 36k nodes:
// 2.25 -> 3.3 ms; ER G(120, 0.08), 199k nodes: 14.9 -> 16.3 ms).  With every
// worker started (7,168) the n = 120 search took 2.4-3.0 ms.
constexpr uint32_t kActiveStart = 512;

SolveArgs solve_args(const vc_ctx *ctx, const vc_graphs *set, const vc_params *params, const Plan &pl, char *base)
{
    SolveArgs a;
    std::memset(&a, 0, sizeof(a));
    a.gbuf = set->d_words;
    a.gdesc = set->d_desc;
    a.gstate = set->d_state;
    a.covers = set->d_covers;
    a.num_graphs = (uint32_t)set->count;
    // deadline on the device wall clock (every wait loop checks it)
    const uint64_t ms = params->time_limit_ms ? params->time_limit_ms : 3600ull * 1000ull;
    fill_args(a, pl, base, params->version, params->start_depth, ms * (uint64_t)ctx->wall_khz);
    a.flush_share = std::max<uint32_t>(1u, pl.workers / std::max<uint32_t>(1u, a.num_graphs));
    // staged start (vcg::dormant_join) when the grid is automatic and the
    // roots could keep fewer than half the workers busy; the adjacency
    // kernel is always staged (active0 >= workers: every worker starts)
    if (pl.adj || (!params->num_blocks && 2ull * pl.roots < pl.workers && kActiveStart < pl.workers)) {
        a.active0 = std::max<uint32_t>(kActiveStart, 2u * pl.roots);
        a.dormant_ticks = (uint64_t)ctx->wall_khz / 10u;  // 100 us
    }
    return a;
}

void fill_info(vc_launch_info *info, const vc_ctx *ctx, const Plan &pl, float ms, uint64_t nodes, uint64_t idle,
               uint64_t donations, uint64_t steals, uint64_t t0, uint64_t last_exit, uint64_t state_bytes)
{
    if (!info) return;
    std::memset(info, 0, sizeof(*info));
    info->num_blocks = pl.num_blocks;
    info->block_dim = pl.block_dim;
    info->workers = pl.workers;
    info->blocks_per_cu = pl.blocks_per_cu;
    info->use_global_memory = pl.variant != kLds16;
    info->list_size = pl.qsize;
    info->list_threshold = pl.qthreshold;
    info->stack_depth = pl.depth;
    info->kernel_ms = ms;
    info->cu_count = (uint32_t)ctx->cu_count;
    info->tree_nodes = nodes;
    info->donations = donations;
    info->steals = steals;
    // share of worker time spent waiting for work
    const double span = (double)ms * 1e-3 * (double)ctx->wall_khz * 1e3 * (double)pl.workers;
    info->idle_fraction = span > 0 ? (float)std::min(1.0, (double)idle / span) : 0.f;
    info->active_ms = (t0 && last_exit > t0 && ctx->wall_khz > 0) ? (float)((double)(last_exit - t0) / ctx->wall_khz)
                                                                 : 0.f;
    info->state_bytes = state_bytes;
}

}  // namespace

extern "C" {

vc_status vc_solve(vc_ctx *ctx, const vc_graphs *set_c, const vc_job *jobs, const vc_params *params,
                   vc_result *results, uint8_t *const *covers, vc_launch_info *info)
{
    if (!ctx || !set_c || !jobs || !params || !results) return VC_EINVAL;
    vc_graphs *set = const_cast<vc_graphs *>(set_c);
    Plan pl;
    vc_status st = prepare_solve(ctx, set, jobs, params, pl, 0);
    if (st != VC_OK) return st;

    if (ctx->scratch_bytes < pl.total) {
        ctx_release(ctx, ctx->scratch, false);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        if (!hip_ok(hipMalloc(&ctx->scratch, pl.total))) return VC_ENOMEM;
        ctx->scratch_bytes = pl.total;
    }
    const SolveArgs a = solve_args(ctx, set, params, pl, static_cast<char *>(ctx->scratch));

    bool want_covers = false;
    if (covers)
        for (int i = 0; i < set->count; ++i) want_covers |= covers[i] != nullptr;
    // pinned readback buffers, mapped: publish_kernel writes them
    constexpr unsigned kPinned = hipHostMallocCoherent | hipHostMallocMapped;
    bool ok = true;
    if (!set->h_state)
        ok = hip_ok(hipHostMalloc(reinterpret_cast<void **>(&set->h_state), sizeof(GState) * set->count, kPinned)) &&
             hip_ok(hipHostMalloc(reinterpret_cast<void **>(&set->h_glob), sizeof(Global), kPinned)) &&
             hip_ok(hipHostMalloc(reinterpret_cast<void **>(&set->h_cov), sizeof(uint32_t) * set->cover_words, kPinned));
    if (ok && set->ws_cap < pl.workers) {
        ctx_release(ctx, set->h_ws, true);
        set->h_ws = nullptr;
        ok = hip_ok(hipHostMalloc(reinterpret_cast<void **>(&set->h_ws), sizeof(vcg::WorkerStats) * pl.workers, kPinned));
        set->ws_cap = ok ? pl.workers : 0;
    }
    if (!ok) return VC_ENOMEM;

    // the device work of one solve, in stream order
    auto enqueue_dispatch = [&](bool timed) -> bool {
        hipLaunchKernelGGL(vcg::init_kernel, dim3(256), dim3(256), 0, ctx->stream, a, pl.units, 1u);
        if (!hip_ok(hipGetLastError())) return false;
        if (timed && !hip_ok(hipEventRecord(ctx->ev0, ctx->stream))) return false;
        if (!launch_solve(pl, ctx->stream, a)) return false;
        if (timed && !hip_ok(hipEventRecord(ctx->ev1, ctx->stream))) return false;
        vcg::Publish pub;
        std::memset(&pub, 0, sizeof(pub));
        pub.src[0] = reinterpret_cast<const uint32_t *>(set->d_state);
        pub.dst[0] = reinterpret_cast<uint32_t *>(set->h_state);
        pub.words[0] = (uint32_t)(sizeof(GState) / 4u * (size_t)set->count);
        pub.src[1] = reinterpret_cast<const uint32_t *>(a.glob);
        pub.dst[1] = reinterpret_cast<uint32_t *>(set->h_glob);
        pub.words[1] = (uint32_t)(sizeof(Global) / 4u);
        pub.src[2] = reinterpret_cast<const uint32_t *>(a.wstats);
        pub.dst[2] = reinterpret_cast<uint32_t *>(set->h_ws);
        pub.words[2] = (uint32_t)(sizeof(vcg::WorkerStats) / 4u * (size_t)pl.workers);
        if (want_covers) {
            pub.src[3] = set->d_covers;
            pub.dst[3] = set->h_cov;
            pub.words[3] = (uint32_t)set->cover_words;
        }
        hipLaunchKernelGGL(vcg::publish_kernel, dim3(64), dim3(256), 0, ctx->stream, pub);
        return hip_ok(hipGetLastError());
    };

    if (params->flags & VC_FLAG_GRAPH) {
        // HIP graph: capture once per (set, launch shape), replay afterwards.
        // The solve is bracketed by events outside the graph.
        DispatchKey key;
        std::memset(&key, 0, sizeof(key));
        key.variant = pl.variant;
        key.num_blocks = pl.num_blocks;
        key.block_dim = pl.block_dim;
        key.version = a.version;
        key.start_depth = a.start_depth;
        key.qsize = pl.qsize;
        key.qthreshold = pl.qthreshold;
        key.depth = pl.depth;
        key.want_covers = want_covers;
        key.off16 = pl.off16;
        key.adj = pl.adj;
        key.active0 = a.active0;
        key.roots = a.roots;
        key.flush_share = a.flush_share;
        key.smem = pl.smem;
        key.limit_ticks = a.limit_ticks;
        key.scratch = ctx->scratch;
        key.h_ws = set->h_ws;
        if (!set->exec || std::memcmp(&key, &set->key, sizeof(key)) != 0) {
            if (set->exec) (void)hipGraphExecDestroy(set->exec);
            if (set->graph) (void)hipGraphDestroy(set->graph);
            set->exec = nullptr;
            set->graph = nullptr;
            if (!hip_ok(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal))) return VC_EDEVICE;
            bool cap = enqueue_dispatch(false);
            hipGraph_t g = nullptr;
            bool ended = hip_ok(hipStreamEndCapture(ctx->stream, &g));
            if (!cap || !ended || !g) {
                if (g) (void)hipGraphDestroy(g);
                return VC_EDEVICE;
            }
            if (!hip_ok(hipGraphInstantiate(&set->exec, g, nullptr, nullptr, 0))) {
                (void)hipGraphDestroy(g);
                return VC_EDEVICE;
            }
            set->graph = g;
            set->key = key;
        }
        if (!hip_ok(hipEventRecord(ctx->ev0, ctx->stream)) || !hip_ok(hipGraphLaunch(set->exec, ctx->stream)) ||
            !hip_ok(hipEventRecord(ctx->ev1, ctx->stream)))
            return VC_EDEVICE;
    } else {
        if (!enqueue_dispatch(true)) return VC_EDEVICE;
    }
    // Wait for this solve's own completion event, not for the stream: when
    // two streams share a hardware queue, a stream-level wait can also wait
    // for another stream's persistent kernel (a running server).
    if (!hip_ok(hipEventRecord(ctx->ev2, ctx->stream)) || !hip_ok(hipEventSynchronize(ctx->ev2))) return VC_EDEVICE;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) ms = 0.f;
    const GState *hs = set->h_state;
    const Global &hg = *set->h_glob;
    const vcg::WorkerStats *wn = set->h_ws;
    const uint32_t *bitmaps = set->h_cov;

    uint64_t total_nodes = 0, idle = 0, donations = 0, steals = 0, last_exit = 0;
    for (uint32_t wi = 0; wi < pl.workers; ++wi) {
        const vcg::WorkerStats &x = wn[wi];
        total_nodes += x.nodes;
        idle += x.idle_ticks;
        donations += x.donations;
        steals += x.steals;
        last_exit = std::max<uint64_t>(last_exit, x.exit_tick);
    }
    vc_status overall = hg.error ? VC_EDEVICE : (hg.timed_out ? VC_ETIMEOUT : VC_OK);
    for (int i = 0; i < set->count; ++i) {
        vc_result &r = results[i];
        r = vcg::result_of(hs[i], jobs[i].instance, hg, (uint32_t)ctx->wall_khz);
        if (covers && covers[i] && r.cover_from_gpu) {
            const uint32_t *bm = &bitmaps[set->h_desc[i].off_cover];
            for (uint32_t v = 0; v < set->h_desc[i].n; ++v) covers[i][v] = (bm[v >> 5] >> (v & 31u)) & 1u;
        }
    }
    fill_info(info, ctx, pl, ms, total_nodes, idle, donations, steals, hg.t0, last_exit, hg.state_bytes);
    return overall;
}

vc_status vc_solve_noalloc(vc_ctx *ctx, const vc_graphs *set_c, const vc_job *jobs, const vc_params *params,
                           vc_result *results, uint8_t *const *covers, vc_launch_info *info)
{
    if (!ctx || !set_c || !jobs || !params || !results) return VC_EINVAL;
    if (params->flags & VC_FLAG_GRAPH) return VC_ENOTSUP;  // a captured dispatch allocates its graph
    vc_graphs *set = const_cast<vc_graphs *>(set_c);
    if (!hip_ok(hipSetDevice(ctx->device))) return VC_EDEVICE;
    char *const base = static_cast<char *>(noalloc_scratch(ctx->device));
    if (!base) return VC_EDEVICE;
    // read-back area at the end of the scratch; the launch plan gets the rest
    // and shrinks its grid to fit (VC_ENOMEM if one workgroup does not)
    auto up256 = [](size_t x) { return (x + 255u) / 256u * 256u; };
    const size_t res_bytes = up256(sizeof(vc_result) * (size_t)set->count);
    const size_t cov_bytes = up256((size_t)set->cover_words * 32u);
    static_assert(sizeof(vcg::NoAllocHead) <= 256u, "read-back header fits its 256 bytes");
    const size_t tail = 256u + res_bytes + cov_bytes;
    if (tail >= kNoAllocScratchBytes) return VC_ENOMEM;
    const size_t tail_at = kNoAllocScratchBytes - tail;  // a multiple of 256
    std::lock_guard<std::mutex> lk(g_noalloc_call_mu[ctx->device]);
    Plan pl;
    vc_status st = prepare_solve(ctx, set, jobs, params, pl, tail_at);
    if (st != VC_OK) return st;
    if (pl.total > tail_at) return VC_ENOMEM;
    const SolveArgs a = solve_args(ctx, set, params, pl, base);
    auto *d_head = reinterpret_cast<vcg::NoAllocHead *>(base + tail_at);
    auto *d_res = reinterpret_cast<vc_result *>(base + tail_at + 256u);
    auto *d_cov = reinterpret_cast<uint8_t *>(base + tail_at + 256u + res_bytes);

    hipLaunchKernelGGL(vcg::init_kernel, dim3(256), dim3(256), 0, ctx->stream, a, pl.units, 1u);
    if (!hip_ok(hipGetLastError()) || !hip_ok(hipEventRecord(ctx->ev0, ctx->stream)) ||
        !launch_solve(pl, ctx->stream, a) || !hip_ok(hipEventRecord(ctx->ev1, ctx->stream)))
        return VC_EDEVICE;
    hipLaunchKernelGGL(vcg::noalloc_results_kernel, dim3(64), dim3(256), 0, ctx->stream, a,
                       (uint32_t)ctx->wall_khz, set->cover_words, d_head, d_res, d_cov);
    vcg::NoAllocHead head;
    if (!hip_ok(hipGetLastError()) ||
        !hip_ok(hipMemcpyAsync(&head, d_head, sizeof(head), hipMemcpyDeviceToHost, ctx->stream)) ||
        !hip_ok(hipMemcpyAsync(results, d_res, sizeof(vc_result) * (size_t)set->count, hipMemcpyDeviceToHost,
                               ctx->stream)) ||
        !hip_ok(hipEventRecord(ctx->ev2, ctx->stream)) || !hip_ok(hipEventSynchronize(ctx->ev2)))
        return VC_EDEVICE;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) ms = 0.f;
    if (covers) {
        bool any = false;
        for (int i = 0; i < set->count; ++i) {
            if (!covers[i] || !results[i].cover_from_gpu || set->h_desc[i].n == 0) continue;
            if (!hip_ok(hipMemcpyAsync(covers[i], d_cov + (size_t)set->h_desc[i].off_cover * 32u, set->h_desc[i].n,
                                       hipMemcpyDeviceToHost, ctx->stream)))
                return VC_EDEVICE;
            any = true;
        }
        if (any && (!hip_ok(hipEventRecord(ctx->ev2, ctx->stream)) || !hip_ok(hipEventSynchronize(ctx->ev2))))
            return VC_EDEVICE;
    }
    fill_info(info, ctx, pl, ms, head.nodes, head.idle_ticks, head.donations, head.steals, head.t0, head.last_exit,
              head.state_bytes);
    return head.error ? VC_EDEVICE : (head.timed_out ? VC_ETIMEOUT : VC_OK);
}

}  // extern "C"

// ===========================================================================
// autotuner
// ===========================================================================

extern "C" vc_status vc_autotune(vc_ctx *ctx, const vc_graphs *set, const vc_job *jobs, const vc_params *base,
                                 vc_params *best, float *best_ms)
{
    if (!ctx || !set || !jobs || !base || !best) return VC_EINVAL;
    // bucket: log2 of the mean vertex count, log2 of the mean density in half steps
    const int nb = (int)std::floor(std::log2(std::max(1.0, set->mean_n)));
    const int db = (int)std::floor(std::log2(std::max(1e-6, set->mean_density)) * 2.0);
    const std::pair<int, int> key(nb, db);
    auto hit = ctx->tuned.find(key);
    if (hit != ctx->tuned.end()) {
        *best = hit->second;
        best->time_limit_ms = base->time_limit_ms;
        best->flags = base->flags;
        if (best_ms) *best_ms = 0.f;
        return VC_OK;
    }
    // candidates: workgroup of 1, 2 or 4 worker waves; node storage in LDS or
    // HBM; worklist sized for the grid or 4x larger.  LDS shapes go first:
    // where they fit they are the fast ones, and each later candidate runs
    // under a time limit of twice the best kernel time so far, so a slow
    // shape ends early instead of running the whole sample.
    std::vector<vc_params> cands;
    for (int gm : {0, 1})
        for (uint32_t bd : {64u, 128u, 256u})
            for (uint32_t qmul : {0u, 4u}) {
                vc_params p = *base;
                p.block_dim = bd;
                p.use_global_memory = gm;
                p.list_size = 0;
                p.list_threshold = 0.f;
                p.flags = 0;
                if (qmul) {
                    Plan pl;
                    if (make_plan(ctx, set, jobs, &p, pl) != VC_OK) continue;
                    p.list_size = next_pow2(pl.qsize * qmul);
                }
                cands.push_back(p);
            }
    std::vector<vc_result> res(set->count);
    float best_t = 0.f, best_kernel_ms = 0.f;
    int best_i = -1;
    for (size_t i = 0; i < cands.size(); ++i) {
        Plan pl;
        if (make_plan(ctx, set, jobs, &cands[i], pl) != VC_OK) continue;
        if (best_i >= 0) {
            const uint32_t cap = (uint32_t)std::max(1000.f, 2.f * best_kernel_ms);
            if (!cands[i].time_limit_ms || cap < cands[i].time_limit_ms) cands[i].time_limit_ms = cap;
        }
        vc_launch_info info;
        vc_status st = vc_solve(ctx, set, jobs, &cands[i], res.data(), nullptr, &info);
        if (st != VC_OK && st != VC_ETIMEOUT) continue;
        // time per solved graph, so a shape that finishes more graphs within
        // the jobs' node budgets is not penalised for it
        uint32_t solved = 0;
        for (const vc_result &r : res) solved += r.solved;
        const float t = info.kernel_ms / (float)std::max<uint32_t>(solved, 1u);
        if (best_i < 0 || t < best_t) {
            best_t = t;
            best_kernel_ms = info.kernel_ms;
            best_i = (int)i;
        }
    }
    if (best_i < 0) return VC_ENOTSUP;
    *best = cands[best_i];
    best->time_limit_ms = base->time_limit_ms;
    best->flags = base->flags;
    ctx->tuned[key] = *best;
    if (best_ms) *best_ms = best_t;
    return VC_OK;
}

// ===========================================================================
// serving: continuous batching over one persistent launch
// ===========================================================================
//
// The host never calls the HIP runtime while the server kernel runs (small
// copies may be executed as blit kernels, which could not be scheduled next
// to a persistent kernel that holds every CU).  Requests are written into
// pinned host memory -- graph words into the host arena, descriptor into the
// host descriptor table, slot id into the request ring -- and published by a
// release store of the ring tail.  The worker that claims a request copies
// it into HBM; completions and cover bitmaps come back the same way.

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace {

// First-fit allocator over 64-byte units of the arena.
class ArenaAlloc {
  public:
    void init(uint64_t units)
    {
        free_.clear();
        if (units) free_[0] = units;
    }
    bool alloc(uint64_t units, uint64_t &off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < units) continue;
            off = it->first;
            uint64_t rest = it->second - units;
            free_.erase(it);
            if (rest) free_[off + units] = rest;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t units)
    {
        auto it = free_.emplace(off, units).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<uint64_t, uint64_t> free_;
};

struct ServerRequest {
    uint64_t ticket = 0;
    vc_job job{};
    uint32_t n = 0, m = 0;
    std::vector<uint32_t> words;  // srcPtr[n+1] | dst[2m] | degree[n] | pad | cover bitmap
    uint32_t off_dst = 0, off_deg = 0, off_cover = 0, cover_words = 0;
};

struct ServerSlot {
    bool busy = false;
    uint64_t ticket = 0;
    uint32_t seq = 0;
    uint64_t arena_off = 0, arena_units = 0;
    uint32_t n = 0, cover_word = 0;
    vc_job job{};
};

struct ServerDone {
    vc_result r{};
    std::vector<uint8_t> cover;
    uint64_t done_ns = 0;
};

uint64_t now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

constexpr uint64_t kUnit = 64;

}  // namespace

struct vc_server {
    vc_ctx *ctx = nullptr;
    bool counted = false;  // included in ctx->active_servers
    vc_server_config cfg{};
    Plan pl{};
    SolveArgs a{};
    hipStream_t kstream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void *scratch = nullptr;
    GDesc *d_desc = nullptr;
    GState *d_state = nullptr;
    char *d_arena = nullptr;
    // pinned host memory shared with the kernel
    vcg::HostCtrl *host = nullptr;
    vcg::Completion *comp = nullptr;
    vcg::WorkerStats *h_ws = nullptr;  // per-worker statistics, written at kernel exit
    GDesc *h_desc = nullptr;
    uint32_t *h_ring = nullptr;
    uint32_t *h_arena = nullptr;
    uint32_t ring_size = 0;

    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::unique_ptr<ServerRequest>> queue;  // submitted, not admitted
    std::vector<ServerSlot> slots;
    std::vector<uint32_t> free_slots;
    std::vector<uint32_t> busy_list;
    ArenaAlloc arena;
    std::unordered_map<uint64_t, ServerDone> done;
    uint64_t next_ticket = 1;
    uint32_t tail = 0;
    uint32_t seq = 0;
    bool stopping = false;
    bool dead = false;  // the kernel exited (deadline) or failed
    std::thread th;
};

namespace {

// Admission: move queued requests into free slots while the arena has room,
// write them to pinned memory, publish the new tail.  Returns the number
// admitted.
uint32_t server_admit(vc_server *s)
{
    uint32_t admitted = 0;
    std::lock_guard<std::mutex> lk(s->mu);
    while (!s->queue.empty() && !s->free_slots.empty()) {
        ServerRequest &rq = *s->queue.front();
        const uint64_t units = (rq.words.size() * sizeof(uint32_t) + kUnit - 1) / kUnit;
        uint64_t off = 0;
        if (!s->arena.alloc(units, off)) break;  // admission by memory footprint
        const uint32_t slot = s->free_slots.back();
        s->free_slots.pop_back();
        ServerSlot &sl = s->slots[slot];
        sl.busy = true;
        sl.ticket = rq.ticket;
        sl.seq = ++s->seq;
        if (sl.seq == 0) sl.seq = ++s->seq;
        sl.arena_off = off;
        sl.arena_units = units;
        sl.n = rq.n;
        sl.job = rq.job;
        const uint32_t base = (uint32_t)(off * (kUnit / 4));
        sl.cover_word = base + rq.off_cover;
        std::memcpy(s->h_arena + base, rq.words.data(), rq.words.size() * sizeof(uint32_t));
        GDesc d;
        std::memset(&d, 0, sizeof(d));
        d.n = rq.n;
        d.m = rq.m;
        d.off_src = base;
        d.off_dst = base + rq.off_dst;
        d.off_deg = base + rq.off_deg;
        d.off_cover = sl.cover_word;
        d.instance = rq.job.instance;
        d.k = rq.job.k;
        d.U = rq.job.upper_bound;
        d.seq = sl.seq;
        d.max_nodes = rq.job.max_nodes;
        s->h_desc[slot] = d;
        s->h_ring[s->tail & (s->ring_size - 1u)] = slot;
        ++s->tail;
        s->busy_list.push_back(slot);
        s->queue.pop_front();
        ++admitted;
    }
    // publish: every write above precedes the tail the kernel acquires
    if (admitted) __atomic_store_n(&s->host->tail, s->tail, __ATOMIC_RELEASE);
    return admitted;
}

// Collects finished requests.  Returns how many finished.
uint32_t server_collect(vc_server *s)
{
    uint32_t finished = 0;
    const uint64_t t_done = now_ns();
    std::lock_guard<std::mutex> lk(s->mu);
    for (size_t i = 0; i < s->busy_list.size();) {
        const uint32_t slot = s->busy_list[i];
        ServerSlot &sl = s->slots[slot];
        if (__atomic_load_n(&s->comp[slot].seq, __ATOMIC_ACQUIRE) != sl.seq) {
            ++i;
            continue;
        }
        const vcg::Completion &c = s->comp[slot];
        ServerDone dn;
        vc_result &r = dn.r;
        std::memset(&r, 0, sizeof(r));
        r.tree_nodes = c.nodes;
        const bool budget = (c.flags & 1u) != 0u, cancelled = (c.flags & 2u) != 0u;
        r.minimum = c.minimum;
        if (sl.job.instance == VC_MVC) {
            r.solved = !budget && !cancelled;
        } else {
            r.k_found = c.k_found;
            r.solved = c.k_found || (!budget && !cancelled);
        }
        r.cover_from_gpu = (c.flags & 4u) != 0u;
        r.status = r.solved ? VC_OK : VC_ETIMEOUT;
        if (r.cover_from_gpu) {
            dn.cover.resize(sl.n);
            const uint32_t *bm = s->h_arena + sl.cover_word;
            for (uint32_t v = 0; v < sl.n; ++v) dn.cover[v] = (bm[v >> 5] >> (v & 31u)) & 1u;
        }
        dn.done_ns = t_done;
        s->done[sl.ticket] = std::move(dn);
        s->arena.release(sl.arena_off, sl.arena_units);
        sl.busy = false;
        s->free_slots.push_back(slot);
        s->busy_list[i] = s->busy_list.back();
        s->busy_list.pop_back();
        ++finished;
    }
    if (finished) s->cv.notify_all();
    return finished;
}

void server_loop(vc_server *s)
{
    (void)hipSetDevice(s->ctx->device);
    uint32_t idle_rounds = 0;
    for (;;) {
        uint32_t work = server_admit(s) + server_collect(s);
        // a cheap liveness check every few idle rounds (a stream query is not
        // a copy and needs no CU)
        bool alive = true;
        if (!work && ++idle_rounds % 64 == 0) alive = hipEventQuery(s->ev1) == hipErrorNotReady;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (!alive) s->dead = true;
            if (s->dead || (s->stopping && s->queue.empty() && s->busy_list.empty())) break;
        }
        if (!work) std::this_thread::sleep_for(std::chrono::microseconds(10));
    }
    // let the kernel drain and exit (its own end event: a stream-level wait
    // could also wait for another server's kernel on a shared hardware queue)
    __atomic_store_n(&s->host->shutdown, 1u, __ATOMIC_RELEASE);
    (void)hipEventSynchronize(s->ev1);
    server_collect(s);
    std::lock_guard<std::mutex> lk(s->mu);
    auto fail = [&](uint64_t ticket) {
        ServerDone dn;
        std::memset(&dn.r, 0, sizeof(dn.r));
        dn.r.status = VC_ETIMEOUT;
        dn.done_ns = now_ns();
        s->done[ticket] = std::move(dn);
    };
    for (uint32_t slot : s->busy_list) fail(s->slots[slot].ticket);
    s->busy_list.clear();
    for (auto &rq : s->queue) fail(rq->ticket);
    s->queue.clear();
    s->dead = true;
    s->cv.notify_all();
}

void server_free(vc_server *s)
{
    vc_ctx *ctx = s->ctx;
    if (s->counted) {
        std::lock_guard<std::mutex> lk(ctx->mem_mu);
        --ctx->active_servers;
    }
    ctx_release(ctx, s->scratch, false);
    ctx_release(ctx, s->d_desc, false);
    ctx_release(ctx, s->d_state, false);
    ctx_release(ctx, s->d_arena, false);
    ctx_release(ctx, s->host, true);
    ctx_release(ctx, s->comp, true);
    ctx_release(ctx, s->h_ws, true);
    ctx_release(ctx, s->h_desc, true);
    ctx_release(ctx, s->h_ring, true);
    ctx_release(ctx, s->h_arena, true);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->kstream) (void)hipStreamDestroy(s->kstream);
    delete s;
    ctx_flush_deferred(ctx);
}

template <class T>
bool pinned_alloc(T **p, size_t bytes)
{
    return hip_ok(hipHostMalloc(reinterpret_cast<void **>(p), bytes, hipHostMallocCoherent | hipHostMallocMapped));
}

}  // namespace

extern "C" {

uint64_t vc_now_ns(void) { return now_ns(); }

vc_status vc_server_start(vc_ctx *ctx, const vc_server_config *cfg, vc_server **out)
{
    if (!ctx || !cfg || !out || cfg->max_vertices == 0) return VC_EINVAL;
    *out = nullptr;
    if (!hip_ok(hipSetDevice(ctx->device))) return VC_EDEVICE;
    vc_server *s = new vc_server();
    s->ctx = ctx;
    s->cfg = *cfg;
    {
        std::lock_guard<std::mutex> lk(ctx->mem_mu);
        ++ctx->active_servers;
        s->counted = true;
    }
    const uint32_t slots = cfg->slots ? cfg->slots : 4096u;
    const uint64_t arena_bytes = cfg->arena_bytes ? cfg->arena_bytes : (1ull << 30);
    if (arena_bytes / 4 >= 0xFFFFFFF0ull) {
        server_free(s);
        return VC_EINVAL;
    }
    vc_params p;
    std::memset(&p, 0, sizeof(p));
    p.version = VC_HYBRID;
    p.block_dim = cfg->block_dim;
    p.num_blocks = cfg->num_blocks;
    p.use_global_memory = -1;
    const uint32_t depth = cfg->stack_depth ? cfg->stack_depth : cfg->max_vertices + 1u;
    vc_status st = plan_launch(ctx, cfg->max_vertices, depth, &p, s->pl, true);
    if (st != VC_OK) {
        server_free(s);
        return st;
    }
    s->ring_size = next_pow2(std::max<uint32_t>(2u * slots, 1024u));
    bool ok = hip_ok(hipStreamCreateWithFlags(&s->kstream, hipStreamNonBlocking)) &&
              hip_ok(hipEventCreate(&s->ev0)) && hip_ok(hipEventCreate(&s->ev1)) &&
              hip_ok(hipMalloc(&s->scratch, s->pl.total)) && hip_ok(hipMalloc(&s->d_desc, sizeof(GDesc) * slots)) &&
              hip_ok(hipMalloc(&s->d_state, sizeof(GState) * slots)) && hip_ok(hipMalloc(&s->d_arena, arena_bytes)) &&
              pinned_alloc(&s->host, sizeof(vcg::HostCtrl)) && pinned_alloc(&s->comp, sizeof(vcg::Completion) * slots) &&
              pinned_alloc(&s->h_ws, sizeof(vcg::WorkerStats) * s->pl.workers) &&
              pinned_alloc(&s->h_desc, sizeof(GDesc) * slots) && pinned_alloc(&s->h_ring, sizeof(uint32_t) * s->ring_size) &&
              pinned_alloc(&s->h_arena, arena_bytes);
    if (!ok) {
        server_free(s);
        return VC_ENOMEM;
    }
    std::memset(s->host, 0, sizeof(vcg::HostCtrl));
    std::memset(s->comp, 0, sizeof(vcg::Completion) * slots);
    std::memset(s->h_desc, 0, sizeof(GDesc) * slots);
    s->slots.resize(slots);
    for (uint32_t i = 0; i < slots; ++i) s->free_slots.push_back(slots - 1 - i);
    s->arena.init(arena_bytes / kUnit);

    SolveArgs &a = s->a;
    std::memset(&a, 0, sizeof(a));
    a.gbuf = reinterpret_cast<const uint32_t *>(s->d_arena);
    a.covers = reinterpret_cast<uint32_t *>(s->d_arena);
    a.gdesc = s->d_desc;
    a.gstate = s->d_state;
    a.num_graphs = 0;  // slot state is set by the worker that claims a request
    const uint64_t ms = cfg->time_limit_ms ? cfg->time_limit_ms : 3600ull * 1000ull;
    fill_args(a, s->pl, static_cast<char *>(s->scratch), VC_HYBRID, 0u, ms * (uint64_t)ctx->wall_khz);
    a.serving = 1;
    a.ring_mask = s->ring_size - 1u;
    a.req_ring = s->h_ring;
    a.host_desc = s->h_desc;
    a.host_arena = s->h_arena;
    a.host = s->host;
    a.comp = s->comp;
    a.wstats = s->h_ws;
    a.poll_ticks = std::max<uint64_t>(1, (uint64_t)ctx->wall_khz * 2u / 1000u);  // about 2 microseconds
    // A server is idle most of the time: its idle waves poll about every
    // 220 us, and with thousands of them polling at random phases a new
    // request is still seen within the 2 us host-poll interval.
    a.idle_cap = 10u;
    hipLaunchKernelGGL(vcg::init_kernel, dim3(256), dim3(256), 0, s->kstream, a, 1u, 0u);
    ok = hip_ok(hipGetLastError()) && hip_ok(hipEventRecord(s->ev0, s->kstream)) &&
         launch_solve(s->pl, s->kstream, a) && hip_ok(hipEventRecord(s->ev1, s->kstream));
    if (!ok) {
        (void)hipStreamSynchronize(s->kstream);
        server_free(s);
        return VC_EDEVICE;
    }
    s->th = std::thread(server_loop, s);
    *out = s;
    return VC_OK;
}

vc_status vc_server_submit(vc_server *s, const vc_csr *g, const vc_job *job, uint64_t *ticket)
{
    if (!s || !g || !job || !ticket) return VC_EINVAL;
    if (g->vertexNum > s->cfg.max_vertices || g->vertexNum == 0) return VC_EINVAL;
    if (job->instance != VC_MVC && job->instance != VC_PVC) return VC_EINVAL;
    if (job->upper_bound > g->vertexNum) return VC_EINVAL;
    // Admission by stack footprint: a worker's stack must hold every entry of
    // one root-to-leaf path of this request, min(n, max(U, k+1)) + 1 slots,
    // so the local stacks never overflow into a full worklist.
    {
        const uint32_t t = job->instance == VC_PVC ? std::max(job->upper_bound, job->k + 1u) : job->upper_bound;
        if (std::min(g->vertexNum, t) + 1u > s->pl.depth) return VC_ENOTSUP;
    }
    if (!csr_is_valid(*g)) return VC_EBADGRAPH;
    auto rq = std::make_unique<ServerRequest>();
    rq->job = *job;
    const uint32_t n = g->vertexNum, m = g->edgeNum;
    rq->n = n;
    rq->m = m;
    rq->off_dst = n + 1u;
    rq->off_deg = rq->off_dst + 2u * m;
    rq->off_cover = round_up(rq->off_deg + n, 16u);
    rq->cover_words = round_up((n + 31u) / 32u, 2u);
    rq->words.assign((size_t)rq->off_cover + rq->cover_words, 0u);
    std::memcpy(rq->words.data(), g->srcPtr, sizeof(uint32_t) * (n + 1u));
    if (m) std::memcpy(rq->words.data() + rq->off_dst, g->dst, sizeof(uint32_t) * 2ull * m);
    std::memcpy(rq->words.data() + rq->off_deg, g->degree, sizeof(int32_t) * n);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->dead || s->stopping) return VC_ECANCELLED;
    rq->ticket = s->next_ticket++;
    *ticket = rq->ticket;
    s->queue.push_back(std::move(rq));
    return VC_OK;
}

static vc_status take_done(vc_server *s, uint64_t ticket, vc_result *result, uint8_t *cover, uint64_t *done_ns)
{
    auto it = s->done.find(ticket);
    if (it == s->done.end()) return VC_EAGAIN;
    if (result) *result = it->second.r;
    if (cover && !it->second.cover.empty()) std::memcpy(cover, it->second.cover.data(), it->second.cover.size());
    if (done_ns) *done_ns = it->second.done_ns;
    s->done.erase(it);
    return VC_OK;
}

vc_status vc_server_poll(vc_server *s, uint64_t ticket, vc_result *result, uint8_t *cover, uint64_t *done_ns)
{
    if (!s) return VC_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    return take_done(s, ticket, result, cover, done_ns);
}

vc_status vc_server_wait(vc_server *s, uint64_t ticket, uint32_t timeout_ms, vc_result *result, uint8_t *cover,
                         uint64_t *done_ns)
{
    if (!s) return VC_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(timeout_ms);
    for (;;) {
        vc_status st = take_done(s, ticket, result, cover, done_ns);
        if (st != VC_EAGAIN) return st;
        if (s->cv.wait_until(lk, deadline) == std::cv_status::timeout)
            return take_done(s, ticket, result, cover, done_ns);
    }
}

vc_status vc_server_stop(vc_server *s, vc_launch_info *info)
{
    if (!s) return VC_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    if (s->th.joinable()) s->th.join();
    vc_status st = hip_ok(hipEventSynchronize(s->ev1)) ? VC_OK : VC_EDEVICE;
    // the kernel has ended: its statistics and error word are in pinned host
    // memory (a copy here could wait for another server's persistent kernel)
    if (__atomic_load_n(&s->host->error, __ATOMIC_ACQUIRE) != 0u) st = VC_EDEVICE;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        uint64_t idle = 0;
        for (uint32_t i = 0; i < s->pl.workers; ++i) {
            const vcg::WorkerStats &x = s->h_ws[i];
            info->tree_nodes += x.nodes;
            idle += x.idle_ticks;
            info->donations += x.donations;
            info->steals += x.steals;
        }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s->ev0, s->ev1) != hipSuccess) ms = 0.f;
        info->kernel_ms = ms;
        double span = (double)ms * 1e-3 * (double)s->ctx->wall_khz * 1e3 * (double)s->pl.workers;
        info->idle_fraction = span > 0 ? (float)std::min(1.0, (double)idle / span) : 0.f;
        info->num_blocks = s->pl.num_blocks;
        info->block_dim = s->pl.block_dim;
        info->workers = s->pl.workers;
        info->blocks_per_cu = s->pl.blocks_per_cu;
        info->use_global_memory = s->pl.variant != kLds16;
        info->list_size = s->pl.qsize;
        info->list_threshold = s->pl.qthreshold;
        info->stack_depth = s->pl.depth;
        info->cu_count = (uint32_t)s->ctx->cu_count;
    }
    server_free(s);
    return st;
}

}  // extern "C"
