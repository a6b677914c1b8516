// components.cc -- the reference's components API over the GPU C ABI.
#include "components.h"

#include <stdio.h>

components::components(graph &g) : g_(&g) {}

components::stats components::run() {
    bd_cc_stats st{};
    bd_check(g_->context(), bd_cc_run(g_->context(), &st), "components");
    have_ = true;
    ms_ = st.ms;
    return stats{(long)st.num_components, (long)st.num_iters};
}

void components::clear() { have_ = false; }

bool components::check() {
    int ok = 0;
    bd_check(g_->context(), bd_cc_check(g_->context(), &ok), "components check");
    return ok != 0;
}

std::vector<int64_t> components::component() {
    std::vector<int64_t> l(g_->num_vertices());
    bd_check(g_->context(), bd_cc_fetch(g_->context(), l.data()), "components fetch");
    return l;
}

// components.cc:106-150: for each label, its vertices as ranges ("a-b").
void components::dump() {
    const std::vector<int64_t> l = component();
    const long nv = (long)l.size();
    std::vector<std::vector<long>> members(nv);
    for (long v = 0; v < nv; ++v) members[l[v]].push_back(v);
    for (long c = 0; c < nv; ++c) {
        const std::vector<long> &m = members[c];
        if (m.empty()) continue;
        printf("Component %li: ", c);
        for (size_t i = 0; i < m.size();) {
            size_t j = i;
            while (j + 1 < m.size() && m[j + 1] == m[j] + 1) ++j;
            if (i) printf(", ");
            if (i == j)
                printf("%li", m[i]);
            else
                printf("%li-%li", m[i], m[j]);
            i = j + 1;
        }
        printf("\n");
    }
    fflush(stdout);
}
