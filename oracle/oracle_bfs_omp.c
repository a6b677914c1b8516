/* oracle_bfs_omp.c -- test infrastructure: run_beamer (hybrid_bfs.cc:152-191)
 * on every host core with OpenMP, as the CPU baseline beside the serial
 * restatement (oracle_bfs.c).  Not part of the product.
 *
 * Same algorithm and rules as oracle_beamer:
 *   top-down  (hybrid_bfs.cc:71-100): each frontier vertex claims its
 *             unreached neighbours with a compare-and-swap on parent[]; the
 *             scout count adds the claimed vertex's degree (-parent before
 *             the claim, as clear() encodes it, :433-443);
 *   bottom-up (hybrid_bfs.cc:111-143): every unreached vertex looks for a
 *             reached neighbour and records it in new_parent; a second sweep
 *             commits (two-phase, so a vertex found in this step is never a
 *             parent in it) and counts the awake vertices.
 * The set of vertices a step reaches does not depend on the thread
 * schedule, so scout and awake counts -- and with them the direction of
 * every step -- equal the serial restatement's; only the parent chosen for a
 * vertex may differ (any neighbour one level up, as check() allows). */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

enum { kFlush = 256 };

static void flush(int64_t *queue, int64_t *tail, const int64_t *buf, int n) {
    if (!n) return;
    int64_t at = __atomic_fetch_add(tail, (int64_t)n, __ATOMIC_RELAXED);
    memcpy(queue + at, buf, sizeof(int64_t) * (size_t)n);
}

/* hybrid_bfs.cc:71-100; appends the new frontier to next[*tail...] */
static int64_t td_step(const int64_t *off, const int32_t *adj, int64_t *parent, const int64_t *front,
                       int64_t nfront, int64_t *next, int64_t *tail) {
    int64_t scout = 0;
#pragma omp parallel reduction(+ : scout)
    {
        int64_t buf[kFlush];
        int n = 0;
#pragma omp for schedule(dynamic, 64) nowait
        for (int64_t i = 0; i < nfront; ++i) {
            const int64_t src = front[i];
            for (int64_t e = off[src]; e < off[src + 1]; ++e) {
                const int64_t dst = adj[e];
                int64_t curr = __atomic_load_n(&parent[dst], __ATOMIC_RELAXED);
                if (curr < 0 && __atomic_compare_exchange_n(&parent[dst], &curr, src, 0,
                                                            __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                    scout += -curr;
                    buf[n++] = dst;
                    if (n == kFlush) {
                        flush(next, tail, buf, n);
                        n = 0;
                    }
                }
            }
        }
        flush(next, tail, buf, n);
    }
    return scout;
}

/* hybrid_bfs.cc:111-143 */
static int64_t bu_step(int64_t nv, const int64_t *off, const int32_t *adj, int64_t *parent,
                       int64_t *new_parent, int64_t *next, int64_t *tail) {
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t v = 0; v < nv; ++v) {
        if (parent[v] >= 0) continue;
        int64_t found = -1;
        for (int64_t e = off[v]; e < off[v + 1]; ++e) {
            if (parent[adj[e]] >= 0) {
                found = adj[e];
                break;
            }
        }
        new_parent[v] = found;
    }
    int64_t awake = 0;
#pragma omp parallel reduction(+ : awake)
    {
        int64_t buf[kFlush];
        int n = 0;
#pragma omp for schedule(static) nowait
        for (int64_t v = 0; v < nv; ++v) {
            if (parent[v] < 0 && new_parent[v] >= 0) {
                parent[v] = new_parent[v];
                ++awake;
                buf[n++] = v;
                if (n == kFlush) {
                    flush(next, tail, buf, n);
                    n = 0;
                }
            }
        }
        flush(next, tail, buf, n);
    }
    return awake;
}

int64_t oracle_beamer_omp(int64_t nv, int64_t m, const int64_t *off, const int32_t *adj,
                          int64_t source, int64_t alpha, int64_t beta, int64_t *parent,
                          int8_t *dirs, int64_t dirs_cap) {
    if (source < 0 || source >= nv || alpha <= 0 || beta <= 0) return -1;
    int64_t *q[2] = {malloc(sizeof(int64_t) * (size_t)(nv + 1)),
                     malloc(sizeof(int64_t) * (size_t)(nv + 1))};
    int64_t *new_parent = malloc(sizeof(int64_t) * (size_t)(nv + 1));
    if (!q[0] || !q[1] || !new_parent) {
        free(q[0]);
        free(q[1]);
        free(new_parent);
        return -1;
    }
    /* clear() (:433-443) */
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < nv; ++v) {
        const int64_t d = off[v + 1] - off[v];
        parent[v] = d > 0 ? -d : -1;
    }
    parent[source] = source;
    int cur = 0;
    q[cur][0] = source;
    int64_t nfront = 1, steps = 0;
    int64_t edges_to_check = m * 2;
    int64_t scout = off[source + 1] - off[source];
    while (nfront > 0) {
        if (scout > edges_to_check / alpha) {
            int64_t awake = nfront, old_awake;
            do {
                old_awake = awake;
                int64_t tail = 0;
                awake = bu_step(nv, off, adj, parent, new_parent, q[cur ^ 1], &tail);
                cur ^= 1;
                nfront = tail;
                if (dirs && steps < dirs_cap) dirs[steps] = ORACLE_BU;
                ++steps;
            } while (awake >= old_awake || awake > nv / beta);
            scout = 1;
        } else {
            edges_to_check -= scout;
            int64_t tail = 0;
            scout = td_step(off, adj, parent, q[cur], nfront, q[cur ^ 1], &tail);
            cur ^= 1;
            nfront = tail;
            if (dirs && steps < dirs_cap) dirs[steps] = ORACLE_TD;
            ++steps;
        }
    }
    free(q[0]);
    free(q[1]);
    free(new_parent);
    return steps;
}

int oracle_omp_threads(void) { return omp_get_max_threads(); }
