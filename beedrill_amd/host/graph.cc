// graph.cc -- host-side graph view over the device CSR.
#include "graph.h"

#include <algorithm>

bd_handle bd_open(int device) {
    bd_ctx *c = nullptr;
    int rc = bd_create(&c, device);
    if (rc != BD_OK) throw std::runtime_error("bd_create failed: no usable GPU (status " +
                                              std::to_string(rc) + ")");
    return bd_handle(c, [](bd_ctx *p) { bd_destroy(p); });
}

void bd_check(bd_ctx *c, int rc, const char *what) {
    if (rc != BD_OK) throw std::runtime_error(std::string(what) + ": " + bd_last_error(c));
}

void dist_edge_list::fetch() const {
    if ((long)src_.size() == m_) return;
    src_.resize(m_);
    dst_.resize(m_);
    bd_check(ctx_.get(), bd_graph_edges_download(ctx_.get(), src_.data(), dst_.data()),
             "bd_graph_edges_download");
}

void dist_edge_list::dump() const {
    forall_edges([](long s, long d) { printf("%li -> %li\n", s, d); });
    fflush(stdout);
}

graph::graph(bd_handle ctx) : ctx_(std::move(ctx)) {
    int64_t nv = 0, nnz = 0, m = 0;
    bd_check(ctx_.get(), bd_graph_info(ctx_.get(), &nv, &nnz, &m), "bd_graph_info");
    nv_ = nv;
    nnz_ = nnz;
    m_ = m;
}

void graph::fetch() const {
    if ((long)off_.size() == nv_ + 1) return;
    off_.resize(nv_ + 1);
    cols_.resize(nnz_);
    bd_check(ctx_.get(), bd_graph_download(ctx_.get(), off_.data(), cols_.data()),
             "bd_graph_download");
}

long graph::out_degree(long v) const {
    fetch();
    return (long)(off_[v + 1] - off_[v]);
}

const int32_t *graph::out_neighbors(long v) const {
    fetch();
    return cols_.data() + off_[v];
}

bool graph::check(const dist_edge_list &el) const {
    fetch();
    bool ok = true;
    auto exists = [&](long s, long d) {
        const int32_t *b = out_edges_begin(s), *e = out_edges_end(s);
        return std::find(b, e, (int32_t)d) != e;
    };
    el.forall_edges([&](long s, long d) {
        if (!exists(s, d) || !exists(d, s)) {
            printf("Missing out edge for %li->%li\n", s, d);
            ok = false;
        }
    });
    std::vector<int32_t> tmp;
    for (long v = 0; v < nv_; ++v) {
        tmp.assign(out_edges_begin(v), out_edges_end(v));
        std::sort(tmp.begin(), tmp.end());
        auto pos = std::adjacent_find(tmp.begin(), tmp.end());
        if (pos != tmp.end()) {
            printf("Edge %li->%d is duplicated\n", v, *pos);
            ok = false;
        }
    }
    fflush(stdout);
    return ok;
}

void graph::dump() const {
    fetch();
    for (long v = 0; v < nv_; ++v) {
        if (out_degree(v) == 0) continue;
        printf("%li ->", v);
        for (const int32_t *e = out_edges_begin(v); e < out_edges_end(v); ++e) printf(" %d", *e);
        printf("\n");
    }
    fflush(stdout);
}
