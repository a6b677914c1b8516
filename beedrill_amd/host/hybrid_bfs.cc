// hybrid_bfs.cc -- the reference's hybrid_bfs API over the GPU C ABI.
#include "hybrid_bfs.h"

#include <assert.h>

hybrid_bfs::hybrid_bfs(graph &g) : g_(&g) {}

void hybrid_bfs::run(long source, long alpha, long beta, int policy) {
    bd_ctx *c = g_->context();
    bd_check(c, bd_bfs_run(c, source, alpha, beta, policy, &stats_), "bfs");
    have_ = true;
}

void hybrid_bfs::run_beamer(long source, long alpha, long beta) {
    assert(source < g_->num_vertices());
    run(source, alpha, beta, policy_);
}

void hybrid_bfs::run_with_migrating_threads(long source) {
    assert(source < g_->num_vertices());
    run(source, 15, 18, BD_POLICY_TD_ONLY);
}

void hybrid_bfs::run_with_remote_writes(long source) {
    assert(source < g_->num_vertices());
    run(source, 15, 18, BD_POLICY_REMOTE_WRITES);
}

void hybrid_bfs::run_with_remote_writes_hybrid(long source, long alpha, long beta) {
    assert(source < g_->num_vertices());
    run(source, alpha, beta, BD_POLICY_REMOTE_WRITES_HYBRID);
}

void hybrid_bfs::clear() { have_ = false; }

bool hybrid_bfs::check(long source) {
    int ok = 0;
    bd_check(g_->context(), bd_bfs_check(g_->context(), source, &ok), "check");
    return ok != 0;
}

std::vector<int32_t> hybrid_bfs::depth() {
    std::vector<int32_t> d(g_->num_vertices());
    bd_check(g_->context(), bd_bfs_fetch(g_->context(), d.data(), nullptr), "fetch");
    return d;
}

std::vector<int64_t> hybrid_bfs::parent() {
    std::vector<int64_t> p(g_->num_vertices());
    bd_check(g_->context(), bd_bfs_fetch(g_->context(), nullptr, p.data()), "fetch");
    return p;
}

std::vector<bd_level_rec> hybrid_bfs::levels() {
    int64_t n = 0;
    bd_check(g_->context(), bd_bfs_levels(g_->context(), nullptr, 0, &n), "levels");
    std::vector<bd_level_rec> r(n);
    bd_check(g_->context(), bd_bfs_levels(g_->context(), r.data(), n, &n), "levels");
    return r;
}

// hybrid_bfs.cc:381-399: each reached vertex and the chain of parents up to
// the root.
void hybrid_bfs::print_tree() {
    std::vector<int64_t> p = parent();
    for (long v = 0; v < (long)p.size(); ++v) {
        long par = p[v];
        if (par < 0) continue;
        printf("%4li", v);
        while (true) {
            printf(" <- %4li", par);
            if (par == -1) break;
            if (par == p[par]) break;
            par = p[par];
        }
        printf("\n");
    }
    fflush(stdout);
}

long hybrid_bfs::count_num_traversed_edges() {
    int64_t n = 0;
    bd_check(g_->context(), bd_bfs_count_traversed(g_->context(), &n), "count");
    return (long)n;
}
