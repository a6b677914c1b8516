"""The allocation-free entry point bd_bfs_multi_graph_static against
bd_bfs_multi_graph, through the C ABI: the same depths (bit for bit), the
same step records, parents that pass the oracle's check(), on two calls in a
row, the second with a smaller batch; sizes past its staging are rejected."""
import numpy as np
import pytest

import beedrill_amd as bd

pytestmark = pytest.mark.gpu

REC_KEYS = ("level", "dir", "frontier", "frontier_deg", "unvisited", "next", "next_deg",
            "edges_examined")


def batch(orc, n, seed):
    rng = np.random.default_rng(seed)
    graphs, roots = [], []
    for g in range(n):
        scale, ef = int(rng.integers(4, 12)), int(rng.integers(8, 33))
        nv, s, d = orc.make_rmat(scale, ef, seed=seed * 1000 + g)
        off, adj = orc.build_csr(nv, s, d)
        graphs.append((off, adj, len(s)))
        roots.append(int(orc.pick_roots(off, 1)[0]))
    return graphs, roots


@pytest.mark.parametrize("n,seed", [(40, 11), (7, 12)])
def test_static_multi_graph_matches_allocating_one(ctx, orc, n, seed):
    graphs, roots = batch(orc, n, seed)
    vbase = ctx.upload_union(graphs)
    dp, ms = ctx.multi_graph(vbase, roots)
    ref_lv = [{k: r[k] for k in REC_KEYS} for r in ctx.levels()]
    ms2 = ctx.multi_graph_static(vbase, roots)
    assert ms > 0 and ms2 > 0
    depth, parent = ctx.fetch()
    np.testing.assert_array_equal(depth, dp)
    assert [{k: r[k] for k in REC_KEYS} for r in ctx.levels()] == ref_lv
    for g, (off, adj, _) in enumerate(graphs):
        np.testing.assert_array_equal(depth[vbase[g]:vbase[g + 1]],
                                      orc.bfs_depth(off, adj, roots[g]), err_msg=str(g))
        p = parent[vbase[g]:vbase[g + 1]]
        assert orc.check(off, adj, roots[g], np.where(p >= 0, p - vbase[g], -1)), g


def test_static_multi_graph_rejects_oversized_batches(ctx):
    n = 4097
    vbase = np.arange(n + 1, dtype=np.int64)
    with pytest.raises(bd.BeedrillError) as e:
        ctx.multi_graph_static(vbase, np.zeros(n, np.int64))
    assert e.value.code == -6
