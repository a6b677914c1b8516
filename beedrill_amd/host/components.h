// components.h -- drop-in for the reference class of the same name
// (components.h:5-35): connected components of the resident graph, run on
// the GPU through the C ABI.
#pragma once

#include <vector>

#include "graph.h"

class components {
   public:
    explicit components(graph &g);
    // components.cc:24-31: a copy sharing the other's graph.
    components(const components &other, emu::shallow_copy) : g_(other.g_), have_(other.have_) {}

    // components.h:27-32
    struct stats {
        long num_components;  // components found in the graph
        long num_iters;       // iterations required for convergence
    };
    // components.cc:38-104.  Each component ends labelled by its smallest
    // vertex id.
    stats run();
    // components.cc:106-150: each component's vertices as ranges of ids.
    void dump();
    // components.cc:33-36.  Every run starts from fresh labels on the device.
    void clear();
    // components.cc:152-202.
    bool check();

    // ---- additive API
    std::vector<int64_t> component();
    double last_ms() const { return ms_; }

   private:
    graph *g_;
    bool have_ = false;
    double ms_ = 0;
};
