// emu_shallow.h -- the construction idiom the reference's callers use for
// graph, dist_edge_list and hybrid_bfs: emu::make_repl_shallow<T>(args...)
// returning std::unique_ptr<emu::repl_shallow<T>> (emu_cxx_utils/replicated.h,
// used at hybrid_bfs_main.cc:246, graph_base.h:335-343 and
// dist_edge_list.h:53).  On the Emu it builds T once and shallow-copies it
// to every nodelet; one GPU has one copy, so repl_shallow<T> is T itself and
// the callers' code compiles unchanged.
#pragma once

#include <memory>
#include <utility>

namespace emu {

// Tag selecting a type's shallow-copy constructor T(const T&, shallow_copy).
struct shallow_copy {};

template <class T>
class repl_shallow : public T {
   public:
    template <class... Args>
    explicit repl_shallow(Args &&...args) : T(std::forward<Args>(args)...) {}
    T &get_nth(long) { return *this; }
};

template <class T, class... Args>
std::unique_ptr<repl_shallow<T>> make_repl_shallow(Args &&...args) {
    return std::unique_ptr<repl_shallow<T>>(new repl_shallow<T>(std::forward<Args>(args)...));
}

}  // namespace emu
