// dist_edge_list.cc -- el64 loading onto the GPU (dist_edge_list.cc:50-143).
#include "graph.h"

dist_edge_list::handle dist_edge_list::load_binary(const char *filename, int device) {
    printf("Opening %s...\n", filename);
    fflush(stdout);
    handle h = emu::make_repl_shallow<dist_edge_list>();
    h->ctx_ = bd_open(device);
    bd_check(h->ctx_.get(), bd_graph_load_el64(h->ctx_.get(), filename), "load_binary");
    int64_t nv = 0, nnz = 0, m = 0;
    bd_check(h->ctx_.get(), bd_graph_info(h->ctx_.get(), &nv, &nnz, &m), "bd_graph_info");
    h->nv_ = nv;
    h->m_ = m;
    return h;
}
