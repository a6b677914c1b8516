// graph.h -- host-side drop-in for the reference's graph CSR view
// (graph.h:16-19, graph_base.h:231-305), dist_edge_list (dist_edge_list.h)
// and create_graph_from_edge_list (graph_base.h:335-436).  The adjacency
// lives on the GPU (int64 offsets, int32 neighbour ids); host accessors read
// a copy fetched on first use.
#pragma once

#include <stdint.h>
#include <stdio.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "beedrill_amd.h"
#include "emu_shallow.h"

// Shared owner of one GPU context (device, stream, resident graph).
using bd_handle = std::shared_ptr<bd_ctx>;
bd_handle bd_open(int device = 0);
// Throws std::runtime_error carrying bd_last_error() on a non-zero status.
void bd_check(bd_ctx *c, int rc, const char *what);

// dist_edge_list (dist_edge_list.h:15-89).  load_binary reads an el64 file
// (dist_edge_list.cc:50-143); the edge list stays resident on the GPU.
struct dist_edge_list {
    using handle = std::unique_ptr<emu::repl_shallow<dist_edge_list>>;
    dist_edge_list() = default;
    // dist_edge_list.h:40-45: shares the other's device edge list.
    dist_edge_list(const dist_edge_list &other, emu::shallow_copy)
        : ctx_(other.ctx_), nv_(other.nv_), m_(other.m_) {}
    static handle load_binary(const char *filename, int device = 0);
    long num_vertices() const { return nv_; }
    long num_edges() const { return m_; }
    void dump() const;
    // forall_edges (dist_edge_list.h:68-85), run on the host in list order.
    template <class Function>
    void forall_edges(Function worker) const {
        fetch();
        for (long i = 0; i < m_; ++i) worker(src_[i], dst_[i]);
    }
    bd_handle ctx() const { return ctx_; }

   private:
    void fetch() const;
    bd_handle ctx_;
    long nv_ = 0, m_ = 0;
    mutable std::vector<int64_t> src_, dst_;
};

class graph {
   public:
    explicit graph(bd_handle ctx);
    // graph_base.h:54-60: shares the other's device CSR.
    graph(const graph &other, emu::shallow_copy) : graph(other.ctx_) {}
    long num_vertices() const { return nv_; }
    long num_edges() const { return m_; }
    long out_degree(long v) const;
    const int32_t *out_neighbors(long v) const;
    const int32_t *out_edges_begin(long v) const { return out_neighbors(v); }
    const int32_t *out_edges_end(long v) const { return out_neighbors(v) + out_degree(v); }
    // graph_base::check (graph_base.h:110-134).
    bool check(const dist_edge_list &el) const;
    void dump() const;
    // One GPU: like NODELETS() == 1 (graph_base.h:159-161), nothing to plot.
    void print_distribution() const {}
    bd_ctx *context() const { return ctx_.get(); }
    bd_handle handle() const { return ctx_; }

   private:
    void fetch() const;
    bd_handle ctx_;
    long nv_ = 0, nnz_ = 0, m_ = 0;
    mutable std::vector<int64_t> off_;
    mutable std::vector<int32_t> cols_;
};

// graph_base.h:335-436.  The device CSR was built when the edge list was
// loaded; the graph shares that context.
template <class Graph>
std::unique_ptr<emu::repl_shallow<Graph>> create_graph_from_edge_list(dist_edge_list &el) {
    return emu::make_repl_shallow<Graph>(el.ctx());
}
