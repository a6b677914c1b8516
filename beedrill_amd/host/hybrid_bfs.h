// hybrid_bfs.h -- drop-in for the reference class of the same name
// (hybrid_bfs.h:11-47).  Every public method of the reference is kept with
// its signature; the search runs on the GPU through the C ABI.  Additive
// members (policy selection, results and per-step records) follow them.
#pragma once

#include <vector>

#include "graph.h"

class hybrid_bfs {
   public:
    explicit hybrid_bfs(graph &g);
    // hybrid_bfs.cc:24-32: a copy sharing the other's graph and device state
    // (emu::make_repl_shallow<hybrid_bfs>(g), hybrid_bfs_main.cc:246).
    hybrid_bfs(const hybrid_bfs &other, emu::shallow_copy)
        : g_(other.g_), policy_(other.policy_), stats_(other.stats_), have_(other.have_) {}

    // hybrid_bfs.cc:197-213: top-down steps only, vertices claimed by atomics.
    void run_with_migrating_threads(long source);
    // hybrid_bfs.cc:218-235: top-down steps with remote writes only (each a
    // candidate-write launch and a claim sweep).
    void run_with_remote_writes(long source);
    // hybrid_bfs.cc:243-277: Beamer's state machine, with remote-write steps
    // where Beamer goes bottom-up.
    void run_with_remote_writes_hybrid(long source, long alpha, long beta);
    // hybrid_bfs.cc:152-191.  The direction policy is Beamer's unless
    // set_policy() chose another.
    void run_beamer(long source, long alpha, long beta);

    // hybrid_bfs.cc:433-443.  Every run starts from a cleared state on the
    // device, so clear() only forgets the previous results.
    void clear();
    // hybrid_bfs.cc:281-379.
    bool check(long source);
    // hybrid_bfs.cc:381-399.
    void print_tree();
    // hybrid_bfs.cc:401-414.
    long count_num_traversed_edges();

    // ---- additive API
    void set_policy(int policy) { policy_ = policy; }
    int policy() const { return policy_; }
    const bd_run_stats &last_stats() const { return stats_; }
    std::vector<int32_t> depth();
    std::vector<int64_t> parent();
    std::vector<bd_level_rec> levels();

   private:
    void run(long source, long alpha, long beta, int policy);
    graph *g_;
    int policy_ = BD_POLICY_BEAMER;
    bd_run_stats stats_{};
    bool have_ = false;
};
