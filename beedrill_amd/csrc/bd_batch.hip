// bd_batch.hip -- batched searches (64-root bit-parallel, many graphs).
#include "bd_internal.h"

using namespace bd;

extern "C" int bd_bfs_multi_root(bd_ctx *c, const int64_t *roots, int n, int32_t *depths,
                                 double *ms, int64_t *levels) {
    (void)roots; (void)n; (void)depths; (void)ms; (void)levels;
    return set_err(c, BD_ESTATE, "multi-root search not available");
}

extern "C" int bd_bfs_multi_graph(bd_ctx *c, int ngraphs, const int64_t *vbase,
                                  const int64_t *offsets, const int32_t *cols,
                                  const int64_t *roots, int32_t *depths, double *ms) {
    (void)ngraphs; (void)vbase; (void)offsets; (void)cols; (void)roots; (void)depths; (void)ms;
    return set_err(c, BD_ESTATE, "multi-graph search not available");
}
