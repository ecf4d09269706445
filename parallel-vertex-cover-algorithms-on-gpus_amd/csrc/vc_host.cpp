// vc_host.cpp -- host-side helpers of the reference driver, behind the C ABI:
// graph file loader, graph check and the two upper-bound approximations.
// They run before the device solve (src/main.cu:49-73).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vcgpu.h"

namespace {

struct OwnedCsr {
    std::vector<uint32_t> src, dst;
    std::vector<int32_t> deg;
};

// CSR build of createCSRGraphFromFile (src/auxFunctions.cpp:158-206):
// degree count, prefix offsets, each edge stored in both directions, then
// every adjacency sorted ascending.
vc_status build_csr(uint32_t n, uint32_t m, const uint32_t *eu, const uint32_t *ev, vc_csr_owned *out)
{
    for (uint32_t i = 0; i < m; ++i)
        if (eu[i] >= n || ev[i] >= n) return VC_EBADGRAPH;
    OwnedCsr *o = new OwnedCsr();
    o->src.assign((size_t)n + 1, 0u);
    o->dst.assign(2ull * m + 1, 0u);
    o->deg.assign((size_t)n + 1, 0);
    for (uint32_t i = 0; i < m; ++i) {
        o->deg[eu[i]]++;
        o->deg[ev[i]]++;
    }
    std::vector<uint32_t> fill((size_t)n + 1);
    uint32_t next = 0;
    for (uint32_t v = 0; v < n; ++v) {
        o->src[v] = next;
        fill[v] = next;
        next += (uint32_t)o->deg[v];
    }
    o->src[n] = 2u * m;
    for (uint32_t i = 0; i < m; ++i) {
        o->dst[fill[eu[i]]++] = ev[i];
        o->dst[fill[ev[i]]++] = eu[i];
    }
    for (uint32_t v = 0; v < n; ++v) std::sort(o->dst.begin() + o->src[v], o->dst.begin() + o->src[v + 1]);
    out->storage = o;
    out->csr.vertexNum = n;
    out->csr.edgeNum = m;
    out->csr.srcPtr = o->src.data();
    out->csr.dst = o->dst.data();
    out->csr.degree = o->deg.data();
    return VC_OK;
}

// Working copy for the approximations.  The reference edits the graph in
// place (CSRGraph::deleteVertex, src/CSRGraphRep.cpp:46-64): a deleted
// vertex is removed from the live prefix of each neighbour's list by
// overwriting it with that list's last live entry.  The approximations read
// the lists after such edits (full original ranges, unsorted prefixes), so
// the same edit is kept here to obtain the same U.
//
// The reference's rule passes scan every vertex (leafReductionRule repeats
// full scans until one finds no degree-1 vertex; triangleReductionRule
// rescans from vertex 0 after every triangle it applies), which makes both
// approximations quadratic in n.  Here each rule keeps a set of candidate
// vertices and visits only those, in the same index order, with the same
// per-vertex test: a vertex outside the set is one whose test cannot fire.
// A test reads the vertex's degree, the degrees of the vertices in its
// adjacency range, and (triangle) the range of one of them; all of these
// change only in remove(), which lists every vertex that can be affected:
// each vertex whose degree changed, and the degree-2 vertices in the
// adjacency range of each (a live vertex stays in the range of every
// neighbour, since an edit only overwrites the deleted vertex).  So the
// rules delete the same vertices in the same order as the full scans.
struct Approx {
    uint32_t n;
    std::vector<uint32_t> src, dst;
    std::vector<int32_t> deg;
    std::vector<uint64_t> leaf_set, tri_set;  // candidate bitsets
    uint32_t leaf_n = 0, tri_n = 0;           // their members
    // RemoveMax only: a max tree over the degrees (leaves at size + v) so
    // the highest-degree vertex, lowest index first, is found in O(log n)
    // instead of by a scan of every vertex per deletion
    uint32_t size = 0;
    std::vector<int32_t> maxt;

    explicit Approx(const vc_csr &g) : n(g.vertexNum)
    {
        src.assign(g.srcPtr, g.srcPtr + n + 1);
        dst.assign(g.dst, g.dst + 2ull * g.edgeNum);
        deg.assign(g.degree, g.degree + n);
        leaf_set.assign((n + 63u) / 64u, 0ull);
        tri_set.assign((n + 63u) / 64u, 0ull);
        for (uint32_t v = 0; v < n; ++v) {
            if (deg[v] == 1) add(leaf_set, leaf_n, v);
            if (deg[v] == 2) add(tri_set, tri_n, v);
        }
    }

    static void add(std::vector<uint64_t> &s, uint32_t &count, uint32_t v)
    {
        const uint64_t bit = 1ull << (v & 63u);
        if (!(s[v >> 6] & bit)) {
            s[v >> 6] |= bit;
            ++count;
        }
    }
    static void take(std::vector<uint64_t> &s, uint32_t &count, uint32_t v)
    {
        s[v >> 6] &= ~(1ull << (v & 63u));
        --count;
    }
    // lowest member >= from, or n when there is none
    uint32_t next(const std::vector<uint64_t> &s, uint32_t count, uint32_t from) const
    {
        if (count == 0 || from >= n) return n;
        uint32_t w = from >> 6;
        uint64_t bits = s[w] & (~0ull << (from & 63u));
        while (!bits) {
            if (++w >= s.size()) return n;
            bits = s[w];
        }
        return std::min<uint32_t>(n, (w << 6) + (uint32_t)__builtin_ctzll(bits));
    }

    void track_max()
    {
        size = 1;
        while (size < n) size <<= 1;
        maxt.assign(2ull * size, -1);
        for (uint32_t v = 0; v < n; ++v) maxt[size + v] = deg[v];
        for (uint32_t i = size - 1; i >= 1; --i) maxt[i] = std::max(maxt[2 * i], maxt[2 * i + 1]);
    }
    void set_max(uint32_t v)
    {
        uint32_t i = size + v;
        maxt[i] = deg[v];
        for (i >>= 1; i >= 1; i >>= 1) maxt[i] = std::max(maxt[2 * i], maxt[2 * i + 1]);
    }
    // the first vertex of maximum degree (the reference's scan with a strict
    // >), and that degree; degree 0 when no vertex has an edge left
    uint32_t max_vertex(int32_t &d) const
    {
        d = maxt[1];
        if (d <= 0) return 0;
        uint32_t i = 1;
        while (i < size) i = maxt[2 * i] == d ? 2 * i : 2 * i + 1;
        return i - size;
    }

    // x's degree changed: x and the degree-2 vertices of its range may now
    // pass a rule's test
    void touched(uint32_t x)
    {
        if (size) set_max(x);
        if (deg[x] == 1) add(leaf_set, leaf_n, x);
        if (deg[x] == 2) add(tri_set, tri_n, x);
        for (uint32_t j = src[x]; j < src[x + 1]; ++j)
            if (deg[dst[j]] == 2) add(tri_set, tri_n, dst[j]);
    }

    void remove(uint32_t v)
    {
        const int32_t vb = (int32_t)src[v], ve = (int32_t)src[v] + deg[v];
        for (int32_t i = vb; i < ve; ++i) {
            const uint32_t nb = dst[i];
            const int32_t last = (int32_t)src[nb] + deg[nb];
            for (int32_t j = (int32_t)src[nb]; j < last; ++j) {
                if (dst[j] == v) {
                    dst[j] = dst[last - 1];
                    if (deg[nb] != -1) deg[nb]--;
                }
            }
        }
        deg[v] = -1;
        touched(v);
        for (int32_t i = vb; i < ve; ++i) touched(dst[i]);
    }

    // Membership search over a range that may be unsorted after edits
    // (auxBinarySearch, src/auxFunctions.cpp:25-42).  The reference's
    // unsigned r = mid - 1 would wrap at mid == 0 and read out of bounds;
    // here the search stops there instead.
    bool search(uint32_t l, uint32_t r, uint32_t x) const
    {
        while (l <= r) {
            const uint32_t mid = l + (r - l) / 2;
            if (dst[mid] == x) return true;
            if (dst[mid] < x) {
                l = mid + 1;
            } else {
                if (mid == 0) return false;
                r = mid - 1;
            }
        }
        return false;
    }

    // leafReductionRule(CSRGraph&, unsigned&) (src/auxFunctions.cpp:44-69):
    // passes in index order; a vertex listed behind a pass's position is
    // visited by the next pass, as the next full scan would find it
    void leaf(uint32_t &count)
    {
        while (leaf_n) {
            for (uint32_t i = next(leaf_set, leaf_n, 0); i < n; i = next(leaf_set, leaf_n, i + 1)) {
                take(leaf_set, leaf_n, i);
                if (deg[i] != 1) continue;
                for (uint32_t j = src[i]; j < src[i + 1]; ++j) {
                    if (deg[dst[j]] != -1) {
                        remove(dst[j]);
                        ++count;
                    }
                }
            }
        }
    }

    // triangleReductionRule(CSRGraph&, unsigned&) (src/auxFunctions.cpp:71-127)
    // (each application restarts at the lowest candidate, as the reference's
    // scan restarts at vertex 0)
    void triangle(uint32_t &count)
    {
        bool changed;
        do {
            changed = false;
            for (uint32_t i = next(tri_set, tri_n, 0); i < n; i = next(tri_set, tri_n, i + 1)) {
                take(tri_set, tri_n, i);
                if (deg[i] != 2) continue;
                uint32_t a = 0, b = 0;
                bool have_a = false, keep = false;
                for (uint32_t e = src[i]; e < src[i + 1]; ++e) {
                    const uint32_t u = dst[e];
                    const int32_t du = deg[u];
                    if (du <= 0) continue;
                    if (du == 1 || (du == 2 && u < i)) {
                        keep = true;
                        break;
                    }
                    if (!have_a) {
                        have_a = true;
                        a = u;
                    } else {
                        b = u;
                        break;
                    }
                }
                if (!keep && search(src[b], src[b + 1] - 1, a)) {
                    remove(a);
                    remove(b);
                    count += 2;
                    changed = true;
                    break;
                }
            }
        } while (changed);
    }

    void cover(uint8_t *out) const
    {
        if (!out) return;
        for (uint32_t v = 0; v < n; ++v) out[v] = deg[v] == -1;
    }
};

}  // namespace

extern "C" {

vc_status vc_csr_from_edges(uint32_t n, uint32_t m, const uint32_t *eu, const uint32_t *ev, vc_csr_owned *out)
{
    if (!out || (m && (!eu || !ev))) return VC_EINVAL;
    std::memset(out, 0, sizeof(*out));
    return build_csr(n, m, eu, ev, out);
}

// createCSRGraphFromFile (src/auxFunctions.cpp:129-209): header "n m", then
// m lines "u v", 1-indexed.  The file is read whole and its fields parsed
// directly (whitespace-separated unsigned decimals, as "%u" reads a
// well-formed file); a short or malformed file is VC_EBADGRAPH instead of the
// reference's reads of stale values.
vc_status vc_load_graph_file(const char *path, vc_csr_owned *out)
{
    return vc_load_graph_file_indexed(path, 1u, out);
}

// index_base 1 is the reference's format; index_base 0 reads the files the
// reference ships 0-indexed (test/test_graphs/formated_*, lastfm-asia.txt,
// data/GraphDatabase/Extra/*-cell.txt), which its own loader shifts to -1.
vc_status vc_load_graph_file_indexed(const char *path, uint32_t index_base, vc_csr_owned *out)
{
    if (!path || !out || index_base > 1u) return VC_EINVAL;
    std::memset(out, 0, sizeof(*out));
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return VC_EINVAL;
    std::vector<char> buf;
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), fp)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    std::fclose(fp);
    const char *p = buf.data(), *end = buf.data() + buf.size();
    auto next = [&](uint32_t &x) -> bool {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
        if (p < end && *p == '+') ++p;
        if (p >= end || *p < '0' || *p > '9') return false;
        uint64_t v = 0;
        while (p < end && *p >= '0' && *p <= '9') {
            v = v * 10u + (uint64_t)(*p - '0');
            if (v > 0xFFFFFFFFull) return false;
            ++p;
        }
        x = (uint32_t)v;
        return true;
    };
    uint32_t n = 0, m = 0;
    if (!next(n) || !next(m)) return VC_EBADGRAPH;
    std::vector<uint32_t> eu(m), ev(m);
    for (uint32_t i = 0; i < m; ++i) {
        uint32_t a, b;
        if (!next(a) || !next(b) || a < index_base || b < index_base) return VC_EBADGRAPH;
        eu[i] = a - index_base;  // shift of src/auxFunctions.cpp:151-153
        ev[i] = b - index_base;
    }
    return build_csr(n, m, eu.data(), ev.data(), out);
}

void vc_csr_free(vc_csr_owned *g)
{
    if (!g) return;
    delete static_cast<OwnedCsr *>(g->storage);
    std::memset(g, 0, sizeof(*g));
}

// check_graph (src/auxFunctions.cpp:308-354): no vertex lists a neighbour
// twice (in any order) and none lists itself.  The reference marks an n x n
// matrix (O(n^2) memory and time); a last-seen marker per vertex gives the
// same answer in O(n + m).  Offsets that leave the edge array and indices
// out of range (which the reference would read past) also fail.
int vc_check_graph(const vc_csr *g)
{
    if (!g) return 0;
    const uint32_t n = g->vertexNum;
    if (n == 0) return 1;
    if (!g->srcPtr || (g->srcPtr[n] > 0 && !g->dst)) return 0;
    const uint64_t slots = 2ull * g->edgeNum;
    std::vector<uint32_t> seen(n, 0xFFFFFFFFu);
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t b = g->srcPtr[v], e = g->srcPtr[v + 1];
        if (e < b || e > slots) return 0;
        for (uint32_t x = b; x < e; ++x) {
            const uint32_t u = g->dst[x];
            if (u >= n || u == v || seen[u] == v) return 0;
            seen[u] = v;
        }
    }
    return 1;
}

// RemoveMaxApproximateMVC (src/auxFunctions.cpp:211-256).  The reference's
// rule functions return the exit value of their own loop (always false), so
// leaf and triangle run once per outer step before the max-degree deletion.
uint32_t vc_approx_remove_max(const vc_csr *g, uint8_t *cover)
{
    if (!g) return 0;
    Approx a(*g);
    a.track_max();
    uint32_t count = 0;
    for (;;) {
        a.leaf(count);
        a.triangle(count);
        int32_t best_d = 0;
        const uint32_t best_v = a.max_vertex(best_d);
        if (best_d <= 0) break;
        a.remove(best_v);
        ++count;
    }
    a.cover(cover);
    return count;
}

// RemoveEdgeApproximateMVC (src/auxFunctions.cpp:265-306): walk the vertices
// in order, reduce, then put both ends of one random live edge of the vertex
// into the cover.  getRandom (src/auxFunctions.cpp:258-263) re-seeds with
// time(0) on every draw; the seed is explicit here.
uint32_t vc_approx_remove_edge(const vc_csr *g, uint32_t seed, uint8_t *cover)
{
    if (!g) return 0;
    Approx a(*g);
    uint32_t count = 0;
    uint32_t remaining = g->edgeNum;
    for (uint32_t v = 0; v < a.n && remaining > 0; ++v) {
        a.leaf(count);
        a.triangle(count);
        if (a.deg[v] <= 0) continue;
        const int lower = (int)a.src[v];
        const int upper = (int)(a.src[v] + (uint32_t)a.deg[v] - 1u);
        std::srand(seed);
        const uint32_t pick = (uint32_t)((std::rand() % (upper - lower + 1)) + lower);
        const uint32_t u = a.dst[pick];
        remaining -= (uint32_t)a.deg[v];
        remaining -= (uint32_t)a.deg[u];
        ++remaining;
        a.remove(v);
        a.remove(u);
        count += 2;
    }
    a.cover(cover);
    return count;
}

}  // extern "C"
