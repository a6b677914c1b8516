/*
 * oracle_bfs.c -- CPU restatement of the reference's BFS path:
 * create_graph_from_edge_list (graph_base.h:335-436), hybrid_bfs
 * (hybrid_bfs.cc), sliding_queue (sliding_queue.h), worklist (worklist.h),
 * find_if(unroll) (emu_cxx_utils/find.h), lcg (lcg.h), pick_random_vertex
 * (hybrid_bfs_main.cc:136-144) and aggregate_stats (combined.cc:108-146).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Serial: with one worker the
 * reference's Cilk loops run in program order, which this file follows so the
 * parent array is the one a serial reference run would choose.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Planted mistakes, so the tests can show that their known answers catch
 * each kind of error (0 = the faithful restatement).
 *   1: BU loop continues only while awake > old_awake   (:178 uses >=)
 *   2: TD steps do not charge scout to edges_to_check  (:182)
 *   3: scout counts vertices instead of their degrees (:93 uses -deg)
 *   4: BU commits parents during the search pass        (:117-141 two-phase)
 *   5: check() does not test the parent's depth         (:356-360)        */
static int g_mutation = 0;
void oracle_set_mutation(int id) { g_mutation = id; }

static const int32_t *g_bu_head = NULL;
void oracle_set_bu_head(const int32_t *head) { g_bu_head = head; }

void oracle_head_index(int64_t nv, const int64_t *off, const int32_t *adj, int32_t *head) {
    const int64_t cap = ((int64_t)1 << 31) - 1;
    for (int64_t v = 0; v < nv; ++v) {
        int64_t best = -1, best_deg = -1;
        for (int64_t e = off[v]; e < off[v + 1]; ++e) {
            const int64_t u = adj[e];
            int64_t d = off[u + 1] - off[u];
            if (d > cap) d = cap;
            if (d > best_deg) {
                best_deg = d;
                best = u;
            }
        }
        head[v] = (int32_t)best;
    }
}

/* ------------------------------------------------------------ graph */

int oracle_build_csr(int64_t nv, int64_t m, const int64_t *src, const int64_t *dst,
                     int64_t *off, int32_t *adj) {
    /* graph_base.h:361-366 degree count (+1 for each endpoint) */
    memset(off, 0, sizeof(int64_t) * (size_t)(nv + 1));
    for (int64_t i = 0; i < m; ++i) {
        int64_t s = src[i], d = dst[i];
        if (s < 0 || s >= nv || d < 0 || d >= nv) return -1;
        off[s + 1]++;
        off[d + 1]++;
    }
    for (int64_t v = 0; v < nv; ++v) off[v + 1] += off[v];
    /* graph_base.h:404-427 carve, then insert both directions in list order */
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nv > 0 ? nv : 1));
    if (!pos) return -1;
    memcpy(pos, off, sizeof(int64_t) * (size_t)nv);
    for (int64_t i = 0; i < m; ++i) {
        adj[pos[src[i]]++] = (int32_t)dst[i];
        adj[pos[dst[i]]++] = (int32_t)src[i];
    }
    free(pos);
    return 0;
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

static int edge_exists(const int64_t *off, const int32_t *adj, int64_t s, int64_t d) {
    for (int64_t e = off[s]; e < off[s + 1]; ++e)
        if (adj[e] == d) return 1;
    return 0;
}

int oracle_check_graph(int64_t nv, int64_t m, const int64_t *src, const int64_t *dst,
                       const int64_t *off, const int32_t *adj) {
    int ok = 1;
    for (int64_t i = 0; i < m; ++i) {
        if (!edge_exists(off, adj, src[i], dst[i])) ok = 0;
        if (!edge_exists(off, adj, dst[i], src[i])) ok = 0;
    }
    /* graph_base.h:124-131 adjacent_find; lists are compared after a sort so
     * a duplicate anywhere in the list is caught */
    int64_t maxd = 0;
    for (int64_t v = 0; v < nv; ++v)
        if (off[v + 1] - off[v] > maxd) maxd = off[v + 1] - off[v];
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxd > 0 ? maxd : 1));
    if (!tmp) return 0;
    for (int64_t v = 0; v < nv; ++v) {
        int64_t dg = off[v + 1] - off[v];
        memcpy(tmp, adj + off[v], sizeof(int32_t) * (size_t)dg);
        qsort(tmp, (size_t)dg, sizeof(int32_t), cmp_i32);
        for (int64_t k = 1; k < dg; ++k)
            if (tmp[k] == tmp[k - 1]) ok = 0;
    }
    free(tmp);
    return ok;
}

/* ------------------------------------------------------------- BFS */

typedef struct {
    int64_t nv, m;
    const int64_t *off;
    const int32_t *adj;
    int64_t *parent;     /* hybrid_bfs.h:15 parent_     */
    int64_t *new_parent; /* hybrid_bfs.h:17 new_parent_ */
    /* sliding_queue.h:10-25: buffer, next_, start_, end_ */
    int64_t *buf;
    int64_t next, start, end;
    int32_t *depth;
    int64_t level;
    int64_t *wl_next;    /* worklist.h:20 next_vertex_ */
    int64_t e_bu;        /* neighbour ids examined by the last BU step */
} bfs_state;

static inline int64_t deg_of(const bfs_state *b, int64_t v) {
    return b->off[v + 1] - b->off[v];
}

/* sliding_queue.h:77-84 slide_window */
static void slide(bfs_state *b) {
    b->start = b->end;
    b->end = b->next;
}

static void push(bfs_state *b, int64_t v) {
    b->buf[b->next++] = v;
    if (b->depth) b->depth[v] = (int32_t)(b->level + 1);
}

/* hybrid_bfs.cc:433-443 clear */
static int state_init(bfs_state *b, int64_t nv, int64_t m, const int64_t *off,
                      const int32_t *adj, int64_t *parent, int32_t *depth) {
    memset(b, 0, sizeof(*b));
    b->nv = nv;
    b->m = m;
    b->off = off;
    b->adj = adj;
    b->parent = parent;
    b->depth = depth;
    size_t n = (size_t)(nv > 0 ? nv : 1);
    b->new_parent = (int64_t *)malloc(sizeof(int64_t) * n);
    b->buf = (int64_t *)malloc(sizeof(int64_t) * n);
    b->wl_next = (int64_t *)malloc(sizeof(int64_t) * n);
    if (!b->new_parent || !b->buf || !b->wl_next) return -1;
    for (int64_t v = 0; v < nv; ++v) {
        int64_t d = deg_of(b, v);
        parent[v] = d != 0 ? -d : -1;
        b->new_parent[v] = -1;
        if (depth) depth[v] = -1;
    }
    return 0;
}

static void state_free(bfs_state *b) {
    free(b->new_parent);
    free(b->buf);
    free(b->wl_next);
}

/* hybrid_bfs.cc:71-100.  forall_items appends the window in order onto the
 * LIFO worklist (worklist.h:79-92), so one worker drains it in reverse window
 * order, each vertex's edges in order (worklist.h:96-117). */
static int64_t top_down_step(bfs_state *b) {
    int64_t scout = 0;
    int64_t head = -1;
    for (int64_t i = b->start; i < b->end; ++i) {
        int64_t src = b->buf[i];
        b->wl_next[src] = head;
        head = src;
    }
    for (int64_t src = head; src >= 0; src = b->wl_next[src]) {
        for (int64_t e = b->off[src]; e < b->off[src + 1]; ++e) {
            int64_t dst = b->adj[e];
            int64_t curr = b->parent[dst];
            if (curr < 0) {
                b->parent[dst] = src;
                push(b, dst);
                scout += g_mutation == 3 ? 1 : -curr;
            }
        }
    }
    return scout;
}

/* find.h:14-45 unroll_p: the first (deg mod 4) items in order, then groups
 * of four visited last-to-first.  Returns the index of the hit or -1 and adds
 * the number of predicate calls to *examined. */
static int64_t find_parent_unrolled(const bfs_state *b, int64_t child, int64_t *examined) {
    int64_t beg = b->off[child], end = b->off[child + 1];
    int64_t p = beg;
    while ((end - p) % 4 != 0) {
        int64_t u = b->adj[p++];
        ++*examined;
        if (b->parent[u] >= 0) return u;
    }
    for (; p != end; p += 4) {
        for (int k = 3; k >= 0; --k) {
            int64_t u = b->adj[p + k];
            ++*examined;
            if (b->parent[u] >= 0) return u;
        }
    }
    return -1;
}

/* hybrid_bfs.cc:111-143: pass 1 writes only new_parent_, pass 2 commits, so a
 * vertex found in this step is never seen as a parent in the same step.
 * e_bu counts in-order early-exit scans (the GPU kernel's order), which is
 * what the bandwidth model uses; with a head index set, the head is read
 * first (oracle_set_bu_head). */
static int64_t bottom_up_step(bfs_state *b) {
    int64_t awake = 0;
    int64_t dummy = 0;
    b->e_bu = 0;
    for (int64_t child = 0; child < b->nv; ++child) {
        if (b->parent[child] >= 0) continue;
        int head_hit = 0;
        if (g_bu_head && g_bu_head[child] >= 0) { /* degree-0 vertices read nothing */
            b->e_bu++;
            head_hit = b->parent[g_bu_head[child]] >= 0;
        }
        for (int64_t e = b->off[child]; e < b->off[child + 1] && !head_hit; ++e) {
            b->e_bu++;
            if (b->parent[b->adj[e]] >= 0) break;
        }
        int64_t u = find_parent_unrolled(b, child, &dummy);
        if (u >= 0) b->new_parent[child] = u;
        if (u >= 0 && g_mutation == 4) {
            b->parent[child] = u;
            push(b, child);
            awake += 1;
        }
    }
    for (int64_t v = 0; v < b->nv; ++v) {
        if (b->parent[v] < 0 && b->new_parent[v] >= 0) {
            b->parent[v] = b->new_parent[v];
            push(b, v);
            awake += 1;
        }
    }
    return awake;
}

static void level_stats(const bfs_state *b, oracle_level_t *r) {
    r->frontier = b->end - b->start;
    r->frontier_deg = 0;
    r->max_frontier_deg = 0;
    for (int64_t i = b->start; i < b->end; ++i) {
        int64_t d = deg_of(b, b->buf[i]);
        r->frontier_deg += d;
        if (d > r->max_frontier_deg) r->max_frontier_deg = d;
    }
    r->unvisited = 0;
    r->unvisited_deg = 0;
    for (int64_t v = 0; v < b->nv; ++v)
        if (b->parent[v] < 0) {
            r->unvisited++;
            r->unvisited_deg += deg_of(b, v);
        }
}

static void level_finish(const bfs_state *b, oracle_level_t *r, int64_t first_new) {
    r->next = b->next - first_new;
    r->next_deg = 0;
    for (int64_t i = first_new; i < b->next; ++i) r->next_deg += deg_of(b, b->buf[i]);
}

/* hybrid_bfs.cc:152-191 run_beamer.  C division truncates toward zero like
 * the reference's long division (matters once edges_to_check goes negative). */
int64_t oracle_beamer(int64_t nv, int64_t m, const int64_t *off, const int32_t *adj,
                      int64_t source, int64_t alpha, int64_t beta, int64_t *parent,
                      int32_t *depth, oracle_level_t *log, int64_t log_cap) {
    if (source < 0 || source >= nv || alpha <= 0 || beta <= 0) return -1;
    bfs_state b;
    if (state_init(&b, nv, m, off, adj, parent, depth)) {
        state_free(&b);
        return -1;
    }
    int64_t nlog = 0;
    b.level = -1;
    push(&b, source); /* depth[source] = 0 */
    b.level = 0;
    slide(&b);
    parent[source] = source;

    int64_t edges_to_check = m * 2;
    int64_t scout_count = deg_of(&b, source);

    while (b.start != b.end) {
        if (scout_count > edges_to_check / alpha) {
            int64_t awake_count, old_awake_count;
            awake_count = b.end - b.start;
            do {
                old_awake_count = awake_count;
                oracle_level_t r;
                memset(&r, 0, sizeof(r));
                if (log) level_stats(&b, &r);
                int64_t first_new = b.next;
                awake_count = bottom_up_step(&b);
                slide(&b);
                if (log && nlog < log_cap) {
                    r.level = b.level;
                    r.dir = ORACLE_BU;
                    r.edges_to_check = edges_to_check;
                    r.result = awake_count;
                    r.edges_examined = b.e_bu;
                    level_finish(&b, &r, first_new);
                    log[nlog] = r;
                }
                ++nlog;
                b.level++;
            } while ((g_mutation == 1 ? awake_count > old_awake_count
                                      : awake_count >= old_awake_count) ||
                     (awake_count > nv / beta));
            scout_count = 1;
        } else {
            if (g_mutation != 2) edges_to_check -= scout_count;
            oracle_level_t r;
            memset(&r, 0, sizeof(r));
            if (log) level_stats(&b, &r);
            int64_t first_new = b.next;
            scout_count = top_down_step(&b);
            slide(&b);
            if (log && nlog < log_cap) {
                r.level = b.level;
                r.dir = ORACLE_TD;
                r.edges_to_check = edges_to_check;
                r.result = scout_count;
                r.edges_examined = r.frontier_deg;
                level_finish(&b, &r, first_new);
                log[nlog] = r;
            }
            ++nlog;
            b.level++;
        }
    }
    state_free(&b);
    return nlog;
}

/* hybrid_bfs.cc:197-213 run_with_migrating_threads */
int64_t oracle_td_only(int64_t nv, int64_t m, const int64_t *off, const int32_t *adj,
                       int64_t source, int64_t *parent, int32_t *depth) {
    if (source < 0 || source >= nv) return -1;
    bfs_state b;
    if (state_init(&b, nv, m, off, adj, parent, depth)) {
        state_free(&b);
        return -1;
    }
    b.level = -1;
    push(&b, source);
    b.level = 0;
    slide(&b);
    parent[source] = source;
    int64_t steps = 0;
    while (b.start != b.end) {
        top_down_step(&b);
        slide(&b);
        b.level++;
        ++steps;
    }
    state_free(&b);
    return steps;
}

/* hybrid_bfs.cc:284-307 */
void oracle_bfs_depth(int64_t nv, const int64_t *off, const int32_t *adj,
                      int64_t source, int32_t *depth) {
    for (int64_t v = 0; v < nv; ++v) depth[v] = -1;
    if (source < 0 || source >= nv) return;
    int64_t *q = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nv > 0 ? nv : 1));
    if (!q) return;
    int64_t qh = 0, qt = 0;
    q[qt++] = source;
    depth[source] = 0;
    while (qh < qt) {
        int64_t u = q[qh++];
        for (int64_t e = off[u]; e < off[u + 1]; ++e) {
            int64_t v = adj[e];
            if (depth[v] == -1) {
                depth[v] = depth[u] + 1;
                q[qt++] = v;
            }
        }
    }
    free(q);
}

/* hybrid_bfs.cc:281-379 (depth by a fresh serial BFS, then per vertex). */
int oracle_check(int64_t nv, const int64_t *off, const int32_t *adj, int64_t source,
                 const int64_t *parent) {
    int32_t *depth = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
    if (!depth) return 0;
    oracle_bfs_depth(nv, off, adj, source, depth);
    int correct = 1;
    for (int64_t u = 0; u < nv; ++u) {
        if (depth[u] >= 0 && parent[u] >= 0) {
            if (u == source) {
                if (!(parent[u] == u && depth[u] == 0)) {
                    correct = 0;
                    break;
                }
                continue;
            }
            int found = 0;
            for (int64_t e = off[u]; e < off[u + 1]; ++e) {
                int64_t v = adj[e];
                if (v == parent[u]) {
                    if (g_mutation != 5 && depth[v] != depth[u] - 1) break;
                    found = 1;
                    break;
                }
            }
            if (!found) {
                correct = 0;
                break;
            }
        } else if ((depth[u] < 0) != (parent[u] < 0)) {
            correct = 0;
            break;
        }
    }
    free(depth);
    return correct;
}

/* hybrid_bfs.cc:401-414 */
int64_t oracle_count_traversed(int64_t nv, const int64_t *off, const int64_t *parent) {
    int64_t sum = 0;
    for (int64_t v = 0; v < nv; ++v)
        if (parent[v] >= 0) sum += off[v + 1] - off[v];
    return sum / 2;
}

/* --------------------------------------------------------- driver */

/* lcg.h:12-27: skip-ahead by binary powering of the affine map x -> a x + c. */
uint64_t oracle_lcg_init(uint64_t step) {
    uint64_t mul_k = 6364136223846793005ULL, add_k = 1, ran = 1;
    for (uint64_t un = step; un; un >>= 1) {
        if (un & 1) ran = mul_k * ran + add_k;
        add_k *= (mul_k + 1);
        mul_k *= mul_k;
    }
    return ran;
}

/* lcg.h:29-32 */
uint64_t oracle_lcg_next(uint64_t *state) {
    *state = 6364136223846793005ULL * *state + 1ULL;
    return *state;
}

/* hybrid_bfs_main.cc:136-144 with lcg rng(0) (:159).  `rng() % V` is an
 * unsigned long modulo (the reference's long V converts to unsigned). */
void oracle_pick_roots(int64_t nv, const int64_t *off, int64_t n, int64_t *roots) {
    uint64_t st = oracle_lcg_init(0);
    for (int64_t i = 0; i < n; ++i) {
        int64_t s;
        do {
            s = (int64_t)(oracle_lcg_next(&st) % (uint64_t)nv);
        } while (off[s + 1] - off[s] == 0);
        roots[i] = s;
    }
}

/* combined.cc:115-145 */
void oracle_aggregate(const double *x, int64_t n, double out[6]) {
    double sum = 0, rsum = 0, sq = 0, mn = x[0], mx = x[0];
    for (int64_t i = 0; i < n; ++i) sum += x[i];
    double mean = sum / (double)n;
    for (int64_t i = 0; i < n; ++i) rsum += 1.0 / x[i];
    for (int64_t i = 0; i < n; ++i) {
        double e = fabs(x[i] - mean);
        sq += e * e;
        if (x[i] > mx) mx = x[i];
        if (x[i] < mn) mn = x[i];
    }
    out[0] = sum;
    out[1] = mean;
    out[2] = (double)n / rsum;
    out[3] = sqrt(sq / (double)(n - 1));
    out[4] = mn;
    out[5] = mx;
}
