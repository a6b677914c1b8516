/*
 * oracle_cc.c -- CPU restatement of the reference's connected components
 * (components.cc).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* components::run (components.cc:38-104), serially.  Every vertex starts as
 * its own component (:42-47).  An iteration passes over every directed edge
 * (src -> dst) and lowers label[src] to label[dst] when that is smaller
 * (:56-66), then shortcuts every label to its root (label[v] =
 * label[label[v]] until it stops changing, :74-78).  The loop ends after
 * the first edge pass that changes nothing (:70); the iteration count
 * includes that pass.  Labels only decrease and stay ids of the vertex's
 * own component, so at the end each component carries the smallest id in
 * it. */
int64_t oracle_cc(int64_t nv, const int64_t *off, const int32_t *adj, int64_t *label,
                  int64_t *ncomp) {
    for (int64_t v = 0; v < nv; ++v) label[v] = v;
    int64_t iters;
    for (iters = 1;; ++iters) {
        int changed = 0;
        for (int64_t src = 0; src < nv; ++src)
            for (int64_t e = off[src]; e < off[src + 1]; ++e) {
                const int64_t dst = adj[e];
                if (label[dst] < label[src]) {
                    label[src] = label[dst];
                    changed = 1;
                }
            }
        if (!changed) break;
        for (int64_t v = 0; v < nv; ++v)
            while (label[v] != label[label[v]]) label[v] = label[label[v]];
    }
    /* component sizes and the count of non-empty components (:84-101) */
    int64_t *size = (int64_t *)calloc((size_t)(nv > 0 ? nv : 1), sizeof(int64_t));
    for (int64_t v = 0; v < nv; ++v) size[label[v]] += 1;
    int64_t n = 0;
    for (int64_t c = 0; c < nv; ++c) n += size[c] > 0;
    free(size);
    if (ncomp) *ncomp = n;
    return iters;
}

/* components::check (components.cc:152-202): one source per label; a serial
 * BFS from it must see only that label, and every vertex must be reached.
 * 1 = PASS, 0 = FAIL. */
int oracle_cc_check(int64_t nv, const int64_t *off, const int32_t *adj, const int64_t *label) {
    if (nv == 0) return 1;
    char *visited = (char *)calloc((size_t)nv, 1);
    int64_t *source = (int64_t *)malloc(sizeof(int64_t) * (size_t)nv);
    int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * (size_t)nv);
    int ok = 1;
    for (int64_t c = 0; c < nv; ++c) source[c] = -1;
    for (int64_t v = 0; v < nv && ok; ++v) {
        if (label[v] < 0 || label[v] >= nv) ok = 0;
        else source[label[v]] = v; /* the last vertex with that label, as the map keeps */
    }
    for (int64_t c = 0; c < nv && ok; ++c) {
        if (source[c] < 0) continue;
        int64_t head = 0, tail = 0;
        visited[source[c]] = 1;
        queue[tail++] = source[c];
        while (head < tail && ok) {
            const int64_t u = queue[head++];
            for (int64_t e = off[u]; e < off[u + 1]; ++e) {
                const int64_t v = adj[e];
                if (label[v] != c) {
                    ok = 0;
                    break;
                }
                if (!visited[v]) {
                    visited[v] = 1;
                    queue[tail++] = v;
                }
            }
        }
    }
    for (int64_t v = 0; v < nv && ok; ++v)
        if (!visited[v]) ok = 0;
    free(visited);
    free(source);
    free(queue);
    return ok;
}
