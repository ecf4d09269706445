/*
 * vc_oracle.c -- CPU oracle (test infrastructure only; see vc_oracle.h).
 *
 * Each function restates the reference function cited above it, step for
 * step, so that results, node counts and rule-call counts match the
 * reference build in oracle/_ref.  Deviations are limited to defined
 * behaviour where the reference has undefined behaviour, and each is marked
 * "UB guard".
 */
#define _POSIX_C_SOURCE 200809L
#include "vc_oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Mutation switch for the oracle's own self-test (tests/test_cpu_oracle.py
 * builds variants with -DORACLE_PLANT=N and checks that the golden vectors
 * reject each one).  The shipped library is built with 0. */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

/* ------------------------------------------------------------------ */
/* Graph construction                                                  */
/* ------------------------------------------------------------------ */

static int cmp_u32(const void *a, const void *b)
{
    /* comp() in src/auxFunctions.cpp:14-23 compares as int; all vertex ids
     * are < 2^31 so the order is the same. */
    int32_t f = *(const int32_t *)a, s = *(const int32_t *)b;
    return (f > s) - (f < s);
}

/* CSR build of src/auxFunctions.cpp:158-206: count degrees, prefix offsets,
 * store each edge in both directions in input order, sort each adjacency. */
int or_graph_from_edges(uint32_t n, uint32_t m, const uint32_t *eu, const uint32_t *ev, or_graph *g)
{
    memset(g, 0, sizeof(*g));
    for (uint32_t i = 0; i < m; ++i)
        if (eu[i] >= n || ev[i] >= n) return -3;
    g->n = n;
    g->m = m;
    g->srcPtr = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    g->dst = (uint32_t *)malloc(sizeof(uint32_t) * (2 * (size_t)m + 1));
    g->degree = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    uint32_t *fill = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    for (uint32_t i = 0; i < m; ++i) {
        g->degree[eu[i]]++;
        g->degree[ev[i]]++;
    }
    uint32_t next = 0;
    for (uint32_t v = 0; v < n; ++v) {
        g->srcPtr[v] = next;
        fill[v] = next;
        next += (uint32_t)g->degree[v];
    }
    g->srcPtr[n] = 2 * m;
    for (uint32_t i = 0; i < m; ++i) {
        g->dst[fill[eu[i]]++] = ev[i];
        g->dst[fill[ev[i]]++] = eu[i];
    }
    free(fill);
    for (uint32_t v = 0; v < n; ++v)
        qsort(&g->dst[g->srcPtr[v]], (size_t)g->degree[v], sizeof(uint32_t), cmp_u32);
    return 0;
}

/* createCSRGraphFromFile (src/auxFunctions.cpp:129-209): "%u%u" header then
 * m pairs shifted down by one (:151-153). */
int or_graph_load(const char *path, or_graph *g)
{
    memset(g, 0, sizeof(*g));
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    unsigned n, m;
    if (fscanf(fp, "%u%u", &n, &m) != 2) {
        fclose(fp);
        return -2;
    }
    uint32_t *eu = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)m + 1));
    uint32_t *ev = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)m + 1));
    int rc = 0;
    for (unsigned i = 0; i < m; ++i) {
        unsigned a, b;
        if (fscanf(fp, "%u%u", &a, &b) != 2) {
            rc = -2;
            break;
        }
        eu[i] = a - 1u;
        ev[i] = b - 1u;
    }
    fclose(fp);
    if (rc == 0) rc = or_graph_from_edges(n, m, eu, ev, g);
    free(eu);
    free(ev);
    return rc;
}

void or_graph_free(or_graph *g)
{
    free(g->srcPtr);
    free(g->dst);
    free(g->degree);
    memset(g, 0, sizeof(*g));
}

/* check_graph (src/auxFunctions.cpp:308-354).  The reference marks an n*n
 * matrix; adjacency lists are sorted here, so a duplicate is two equal
 * neighbours in a row -- same verdict in O(m). */
int or_check_graph(const or_graph *g)
{
    /* check_graph (src/auxFunctions.cpp:308-354): an n x n adjacency matrix
     * marks every listed (vertex, neighbour) pair; a pair marked twice (a
     * duplicate, in any list order) fails, then any mark on the diagonal (a
     * self-loop) fails.  Bits instead of the reference's unsigned ints; a
     * neighbour index >= n (which the reference would write past its
     * matrix with) also fails. */
    const uint64_t n = g->n;
    if (n == 0) return 1;
    uint8_t *mat = (uint8_t *)calloc((size_t)((n * n + 7) / 8), 1);
    if (!mat) return 0;
    int ok = 1;
    for (uint32_t v = 0; v < g->n && ok; ++v) {
        for (uint32_t e = g->srcPtr[v]; e < g->srcPtr[v + 1]; ++e) {
            const uint32_t u = g->dst[e];
            if (u >= g->n) {
                ok = 0;
                break;
            }
            const uint64_t bit = (uint64_t)v * n + u;
            if (mat[bit >> 3] & (1u << (bit & 7))) {
                ok = 0;
                break;
            }
            mat[bit >> 3] |= (uint8_t)(1u << (bit & 7));
        }
    }
    for (uint32_t v = 0; v < g->n && ok; ++v) {
        const uint64_t bit = (uint64_t)v * n + v;
        if (mat[bit >> 3] & (1u << (bit & 7))) ok = 0;
    }
    free(mat);
    return ok;
}

/* ------------------------------------------------------------------ */
/* Upper-bound approximations on a destructively edited copy           */
/* ------------------------------------------------------------------ */

typedef struct {
    uint32_t n, m;
    uint32_t *srcPtr;
    uint32_t *dst;
    int32_t *degree;
} agraph;

static void ag_copy(agraph *a, const or_graph *g)
{
    /* CSRGraph::copy (src/CSRGraphRep.cpp:26-44) */
    a->n = g->n;
    a->m = g->m;
    a->srcPtr = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)g->n + 1));
    a->dst = (uint32_t *)malloc(sizeof(uint32_t) * (2 * (size_t)g->m + 1));
    a->degree = (int32_t *)malloc(sizeof(int32_t) * ((size_t)g->n + 1));
    memcpy(a->srcPtr, g->srcPtr, sizeof(uint32_t) * ((size_t)g->n + 1));
    memcpy(a->dst, g->dst, sizeof(uint32_t) * 2 * (size_t)g->m);
    memcpy(a->degree, g->degree, sizeof(int32_t) * (size_t)g->n);
}

static void ag_free(agraph *a)
{
    free(a->srcPtr);
    free(a->dst);
    free(a->degree);
}

/* CSRGraph::deleteVertex (src/CSRGraphRep.cpp:46-64): remove v from the live
 * prefix of each neighbour's list by overwriting it with the last live entry,
 * shrink that neighbour's degree, then mark v deleted. */
static void ag_delete_vertex(agraph *a, uint32_t v)
{
    for (int32_t i = (int32_t)a->srcPtr[v]; i < (int32_t)a->srcPtr[v] + a->degree[v]; i++) {
        uint32_t nb = a->dst[i];
        int32_t last = (int32_t)a->srcPtr[nb] + a->degree[nb];
        for (int32_t j = (int32_t)a->srcPtr[nb]; j < last; j++) {
            if (a->dst[j] == v) {
                a->dst[j] = a->dst[last - 1];
                if (a->degree[nb] != -1) a->degree[nb]--;
            }
        }
    }
    a->degree[v] = -1;
}

/* auxBinarySearch (src/auxFunctions.cpp:25-42) with unsigned bounds. */
static int ag_bsearch(const uint32_t *arr, uint32_t l, uint32_t r, uint32_t x)
{
    while (l <= r) {
        uint32_t mid = l + (r - l) / 2;
        if (arr[mid] == x) return 1;
        if (arr[mid] < x)
            l = mid + 1;
        else {
            /* UB guard: the reference computes r = mid - 1 = UINT_MAX when
             * mid == 0 and then reads far out of bounds.  Stop instead. */
            if (mid == 0) return 0;
            r = mid - 1;
        }
    }
    return 0;
}

/* leafReductionRule(CSRGraph&, unsigned&) (src/auxFunctions.cpp:44-69).
 * Scans the full original range srcPtr[i]..srcPtr[i+1] (stale entries
 * included), exactly like the reference. */
static void ag_leaf(agraph *a, uint32_t *minimum)
{
    int changed;
    do {
        changed = 0;
        for (uint32_t i = 0; i < a->n; ++i) {
            if (a->degree[i] == 1) {
                changed = 1;
                for (uint32_t j = a->srcPtr[i]; j < a->srcPtr[i + 1]; ++j) {
                    if (a->degree[a->dst[j]] != -1) {
                        ag_delete_vertex(a, a->dst[j]);
                        ++*minimum;
                    }
                }
            }
        }
    } while (changed);
}

/* triangleReductionRule(CSRGraph&, unsigned&) (src/auxFunctions.cpp:71-127):
 * after each triangle the scan restarts from vertex 0. */
static void ag_triangle(agraph *a, uint32_t *minimum)
{
    int changed;
    do {
        changed = 0;
        for (uint32_t i = 0; i < a->n; ++i) {
            if (a->degree[i] != 2) continue;
            uint32_t n1 = 0, n2 = 0;
            int found1 = 0, keep = 0;
            for (uint32_t e = a->srcPtr[i]; e < a->srcPtr[i + 1]; ++e) {
                uint32_t nb = a->dst[e];
                int32_t nd = a->degree[nb];
                if (nd > 0) {
                    if (nd == 1 || (nd == 2 && nb < i)) {
                        keep = 1;
                        break;
                    } else if (!found1) {
                        found1 = 1;
                        n1 = nb;
                    } else {
                        n2 = nb;
                        break;
                    }
                }
            }
            if (!keep) {
                if (ag_bsearch(a->dst, a->srcPtr[n2], a->srcPtr[n2 + 1] - 1, n1)) {
                    changed = 1;
                    ag_delete_vertex(a, n1);
                    ag_delete_vertex(a, n2);
                    *minimum += 2;
                    break;
                }
            }
        }
    } while (changed);
}

/* RemoveMaxApproximateMVC (src/auxFunctions.cpp:211-256).  Both rule
 * functions of the reference return the exit value of their own do/while
 * (always false), so the inner do/while runs leaf then triangle exactly once
 * per outer step; then the lowest-index max-degree vertex is deleted. */
uint32_t or_remove_max_approx(const or_graph *g)
{
    agraph a;
    ag_copy(&a, g);
    uint32_t minimum = 0;
    for (;;) {
        ag_leaf(&a, &minimum);
        ag_triangle(&a, &minimum);
        uint32_t maxV = 0;
        int32_t maxD = 0;
        for (uint32_t i = 0; i < a.n; i++) {
            if (ORACLE_PLANT == 5 ? (a.degree[i] >= maxD && a.degree[i] > 0) : a.degree[i] > maxD) {
                maxV = i;
                maxD = a.degree[i];
            }
        }
        if (maxD == 0) break;
        ag_delete_vertex(&a, maxV);
        ++minimum;
    }
    ag_free(&a);
    return minimum;
}

/* getRandom (src/auxFunctions.cpp:258-263) with time(0) replaced by seed. */
static uint32_t ag_random(int lower, int upper, uint32_t seed)
{
    srand(seed);
    return (uint32_t)((rand() % (upper - lower + 1)) + lower);
}

/* RemoveEdgeApproximateMVC (src/auxFunctions.cpp:265-306). */
uint32_t or_remove_edge_approx(const or_graph *g, uint32_t seed)
{
    agraph a;
    ag_copy(&a, g);
    uint32_t minimum = 0;
    uint32_t remaining = a.m;
    for (uint32_t v = 0; v < a.n && remaining > 0; v++) {
        ag_leaf(&a, &minimum);
        ag_triangle(&a, &minimum);
        if (a.degree[v] > 0) {
            uint32_t re = ag_random((int)a.srcPtr[v], (int)(a.srcPtr[v] + (uint32_t)a.degree[v] - 1u), seed);
            uint32_t u = a.dst[re];
            remaining -= (uint32_t)a.degree[v];
            remaining -= (uint32_t)a.degree[u];
            ++remaining;
            ag_delete_vertex(&a, v);
            ag_delete_vertex(&a, a.dst[re]);
            minimum += 2;
        }
    }
    ag_free(&a);
    return minimum;
}

/* ------------------------------------------------------------------ */
/* Reduction rules on a degree array (src/helperFunctions.cpp)         */
/* ------------------------------------------------------------------ */

/* squareSequential (src/helperFunctions.cpp:3-6): int*int in the reference;
 * identical for |num| <= 46340, computed in 64 bits here (UB guard). */
static long long sq(int num) { return (long long)num * (long long)num; }

/* binarySearchSequential (src/helperFunctions.cpp:8-25). */
static int seq_bsearch(const uint32_t *arr, uint32_t l, uint32_t r, uint32_t x)
{
    return ag_bsearch(arr, l, r, x);
}

/* deleteVertex (src/helperFunctions.cpp:27-47). */
static void del_vertex(const or_graph *g, uint32_t v, int32_t *deg, uint32_t *nd)
{
    if (deg[v] < 0) return;
    for (uint32_t i = g->srcPtr[v]; i < g->srcPtr[v] + (uint32_t)g->degree[v]; ++i) {
        uint32_t nb = g->dst[i];
        if (deg[nb] != -1) --deg[nb];
    }
    deg[v] = -1;
    ++*nd;
}

/* leafReductionRule (src/helperFunctions.cpp:49-75). */
static int rule_leaf(const or_graph *g, int32_t *deg, uint32_t *nd)
{
    int deleted = 0, changed;
    do {
        changed = 0;
        for (uint32_t i = 0; i < g->n; ++i) {
            if (deg[i] == 1) {
                changed = 1;
                for (uint32_t j = g->srcPtr[i]; j < g->srcPtr[i] + (uint32_t)g->degree[i]; ++j) {
                    if (deg[g->dst[j]] != -1) {
                        deleted = 1;
                        del_vertex(g, g->dst[j], deg, nd);
                    }
                }
            }
        }
    } while (changed);
    return deleted;
}

/* highDegreeReductionRule (src/helperFunctions.cpp:77-96); the comparison
 * is unsigned, as in the reference. */
static int rule_high(const or_graph *g, int32_t *deg, uint32_t *nd, int minimum)
{
    int deleted = 0, changed;
    do {
        changed = 0;
        for (uint32_t i = 0; i < g->n; ++i) {
            if (deg[i] > 0 && (ORACLE_PLANT == 4 ? (uint32_t)deg[i] + *nd > (uint32_t)minimum
                                                  : (uint32_t)deg[i] + *nd >= (uint32_t)minimum)) {
                changed = 1;
                deleted = 1;
                del_vertex(g, i, deg, nd);
            }
        }
    } while (changed);
    return deleted;
}

/* triangleReductionRule (src/helperFunctions.cpp:98-155): after each
 * triangle the scan restarts from vertex 0. */
static int rule_triangle(const or_graph *g, int32_t *deg, uint32_t *nd)
{
    int deleted = 0, changed;
    do {
        changed = 0;
        for (uint32_t i = 0; i < g->n; ++i) {
            if (deg[i] != 2) continue;
            uint32_t n1 = 0, n2 = 0;
            int found1 = 0, keep = 0;
            for (uint32_t e = g->srcPtr[i]; e < g->srcPtr[i] + (uint32_t)g->degree[i]; ++e) {
                uint32_t nb = g->dst[e];
                int32_t ndg = deg[nb];
                if (ndg > 0) {
                    if (ndg == 1 || (ndg == 2 && nb < i)) {
                        keep = 1;
                        break;
                    } else if (!found1) {
                        found1 = 1;
                        n1 = nb;
                    } else {
                        n2 = nb;
                        break;
                    }
                }
            }
            if (!keep) {
                if (seq_bsearch(g->dst, g->srcPtr[n2], g->srcPtr[n2 + 1] - 1, n1)) {
                    changed = 1;
                    deleted = 1;
                    del_vertex(g, n1, deg, nd);
                    if (ORACLE_PLANT != 3) del_vertex(g, n2, deg, nd);
                    break;
                }
            }
        }
    } while (changed);
    return deleted;
}

/* ------------------------------------------------------------------ */
/* The two DFS drivers                                                 */
/* ------------------------------------------------------------------ */

/* Shared body of Sequential (src/Sequential.cpp:3-104) and
 * SequentialParameterized (src/SequentialParameterized.cpp:4-111). */
static uint32_t seq_search(const or_graph *g, uint32_t minimum, int param, uint32_t k,
                           uint32_t *kFound, or_stats *st, uint8_t *cover, int32_t *cover_found)
{
    const uint32_t n = g->n;
    or_stats local;
    if (!st) {
        memset(&local, 0, sizeof(local));
        st = &local;
    }
    uint64_t cap = st->max_nodes;
    memset(st, 0, sizeof(*st));
    st->max_nodes = cap;
    st->completed = 1;
    if (kFound) *kFound = 0;
    if (cover_found) *cover_found = 0;

    /* Stack of n+1 slots (src/Sequential.cpp:5-15). */
    size_t slots = (size_t)n + 1;
    int32_t *stack = (int32_t *)malloc(sizeof(int32_t) * slots * (n ? n : 1));
    uint32_t *stackD = (uint32_t *)malloc(sizeof(uint32_t) * slots);
    int32_t *deg = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    for (uint32_t j = 0; j < n; ++j) stack[j] = g->degree[j];
    stackD[0] = 0;
    int top = 0;
    int popNext = 1;
    /* UB guard: the reference reads numDeletedVertices inside the copy loop,
     * so it is never set when n == 0. */
    uint32_t nd = 0;

    while (top != -1) {
        if (cap && st->nodes >= cap) {
            st->completed = 0;
            break;
        }
        st->nodes++;
        if (popNext) {
            memcpy(deg, &stack[(size_t)top * n], sizeof(int32_t) * n);
            nd = stackD[top];
            --top;
        }
        int tri, high;
        do {
            st->leaf_calls++;
            rule_leaf(g, deg, &nd);
            st->tri_calls++;
            tri = rule_triangle(g, deg, &nd);
            st->high_calls++;
            high = rule_high(g, deg, &nd, (int)minimum);
        } while (tri || high);

        uint32_t maxVertex = 0;
        int32_t maxDegree = 0;
        uint32_t numEdges = 0;
        for (uint32_t i = 0; i < n; ++i) {
            int32_t d = deg[i];
            if (ORACLE_PLANT == 2 ? (d >= maxDegree && d > 0) : d > maxDegree) {
                maxDegree = d;
                maxVertex = i;
            }
            if (d > 0) numEdges += (uint32_t)d;
        }
        numEdges /= 2;

        int prune;
        if (!param) {
            /* src/Sequential.cpp:67 */
            prune = nd >= minimum;
            if (!prune) {
                st->square_calls++;
                prune = (long long)numEdges >= sq((int)(minimum - nd - (ORACLE_PLANT == 1 ? 0u : 1u))) + 1;
            }
        } else {
            /* src/SequentialParameterized.cpp:70 */
            prune = nd > k;
            if (!prune) {
                st->square_calls++;
                prune = (long long)numEdges >= sq((int)(k - nd)) + 1;
            }
        }

        if (prune) {
            popNext = 1;
        } else if (maxDegree == 0) {
            if (!param) {
                minimum = nd; /* src/Sequential.cpp:73-77 */
                if (cover) {
                    for (uint32_t i = 0; i < n; ++i) cover[i] = deg[i] == -1;
                }
                if (cover_found) *cover_found = 1;
                popNext = 1;
            } else {
                if (k >= nd) { /* src/SequentialParameterized.cpp:76-84 */
                    if (kFound) *kFound = 1;
                    if (cover) {
                        for (uint32_t i = 0; i < n; ++i) cover[i] = deg[i] == -1;
                    }
                    break;
                }
                popNext = 1;
            }
        } else {
            /* Branch (src/Sequential.cpp:78-95): push a copy with every
             * original neighbour of maxVertex deleted, delete maxVertex on
             * the working copy and continue without popping. */
            popNext = 0;
            ++top;
            if (top > st->max_top) st->max_top = top;
            st->branches++;
            memcpy(&stack[(size_t)top * n], deg, sizeof(int32_t) * n);
            stackD[top] = nd;
            for (uint32_t i = g->srcPtr[maxVertex]; i < (uint32_t)g->degree[maxVertex] + g->srcPtr[maxVertex]; ++i) {
                st->delete_calls++;
                del_vertex(g, g->dst[i], &stack[(size_t)top * n], &stackD[top]);
            }
            st->delete_calls++;
            del_vertex(g, maxVertex, deg, &nd);
        }
    }

    free(stack);
    free(stackD);
    free(deg);
    return minimum;
}

uint32_t or_sequential(const or_graph *g, uint32_t minimum, or_stats *st, uint8_t *cover,
                       int32_t *cover_found)
{
    return seq_search(g, minimum, 0, 0, NULL, st, cover, cover_found);
}

uint32_t or_sequential_parameterized(const or_graph *g, uint32_t minimum, uint32_t k,
                                     uint32_t *kFound, or_stats *st, uint8_t *cover)
{
    uint32_t kf = 0;
    uint32_t r = seq_search(g, minimum, 1, k, &kf, st, cover, NULL);
    if (kFound) *kFound = kf;
    return r;
}

int or_cover_is_valid(const or_graph *g, const uint8_t *cover)
{
    for (uint32_t v = 0; v < g->n; ++v) {
        if (cover[v]) continue;
        for (uint32_t e = g->srcPtr[v]; e < g->srcPtr[v + 1]; ++e)
            if (!cover[g->dst[e]]) return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------ */
/* Multi-threaded CPU baseline                                         */
/* ------------------------------------------------------------------ */

typedef struct {
    int count;
    const or_graph *graphs;
    uint64_t max_nodes;
    uint32_t *results;
    uint64_t *nodes;
    int32_t *done;
    int next; /* guarded by lock */
    pthread_mutex_t lock;
} batch_job;

static void *batch_worker(void *arg)
{
    batch_job *job = (batch_job *)arg;
    for (;;) {
        pthread_mutex_lock(&job->lock);
        int i = job->next++;
        pthread_mutex_unlock(&job->lock);
        if (i >= job->count) break;
        const or_graph *g = &job->graphs[i];
        uint32_t U = or_remove_max_approx(g);
        or_stats st;
        memset(&st, 0, sizeof(st));
        st.max_nodes = job->max_nodes;
        job->results[i] = or_sequential(g, U, &st, NULL, NULL);
        job->nodes[i] = st.nodes;
        job->done[i] = st.completed;
    }
    return NULL;
}

double or_solve_batch_mt(int count, const or_graph *graphs, int threads, uint64_t max_nodes,
                         uint32_t *results, uint64_t *nodes, int32_t *done)
{
    batch_job job;
    job.count = count;
    job.graphs = graphs;
    job.max_nodes = max_nodes;
    job.results = results;
    job.nodes = nodes;
    job.done = done;
    job.next = 0;
    pthread_mutex_init(&job.lock, NULL);
    if (threads < 1) threads = 1;
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, batch_worker, &job);
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(tid);
    pthread_mutex_destroy(&job.lock);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
