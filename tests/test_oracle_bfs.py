"""Pins the oracle's BFS restatement with checks that do not retype it:
library shortest paths, brute force, hand-derived Beamer direction sequences
and planted mistakes that those answers must catch."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components, shortest_path
from scipy.stats import hmean

from conftest import complete_graph, path_graph, star_graph, two_components

TD, BU = 0, 1


def csr_of(orc, nv, s, d):
    off, adj = orc.build_csr(nv, s, d)
    return off, adj, len(s)


def scipy_depth(nv, s, d, source):
    A = sp.coo_matrix((np.ones(len(s)), (s, d)), shape=(nv, nv)).tocsr()
    dist = shortest_path(A, directed=False, unweighted=True, indices=source)
    out = np.where(np.isinf(dist), -1, dist).astype(np.int32)
    return out


def dirs_of(levels):
    return [r["dir"] for r in levels]


# ------------------------------------------------------------------ CSR

def test_csr_neighbour_lists_in_edge_list_order(orc):
    nv, s, d = orc.make_rmat(8, 8, seed=1)
    off, adj = orc.build_csr(nv, s, d)
    assert off[-1] == 2 * len(s)
    expect = [[] for _ in range(nv)]
    for a, b in zip(s.tolist(), d.tolist()):
        expect[a].append(b)
        expect[b].append(a)
    for v in range(nv):
        assert adj[off[v]:off[v + 1]].tolist() == expect[v]
    assert orc.check_graph(nv, s, d, off, adj)


def test_csr_matches_scipy_and_check_graph_catches_corruption(orc):
    nv, s, d = orc.make_rmat(10, 16, seed=2)
    off, adj = orc.build_csr(nv, s, d)
    A = sp.coo_matrix((np.ones(2 * len(s)), (np.r_[s, d], np.r_[d, s])), shape=(nv, nv)).tocsr()
    A.sort_indices()
    np.testing.assert_array_equal(np.diff(off), np.diff(A.indptr))
    for v in range(0, nv, 37):
        assert sorted(adj[off[v]:off[v + 1]].tolist()) == A.indices[A.indptr[v]:A.indptr[v + 1]].tolist()
    bad = adj.copy()
    v = int(np.argmax(np.diff(off)))
    bad[off[v]] = bad[off[v] + 1]          # duplicate + a missing edge
    assert not orc.check_graph(nv, s, d, off, bad)


def test_csr_rejects_out_of_range(orc):
    with pytest.raises(ValueError):
        orc.build_csr(4, np.array([0, 1]), np.array([1, 4]))


# ---------------------------------------------------------------- depth

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_depth_matches_scipy_on_rmat(orc, seed):
    nv, s, d = orc.make_rmat(11, 8, seed=seed)
    off, adj, m = csr_of(orc, nv, s, d)
    for root in orc.pick_roots(off, 3):
        ref = scipy_depth(nv, s, d, root)
        np.testing.assert_array_equal(orc.bfs_depth(off, adj, root), ref)
        _, depth, _ = orc.beamer(off, adj, m, root)
        np.testing.assert_array_equal(depth, ref)
        _, depth_td = orc.td_only(off, adj, m, root)
        np.testing.assert_array_equal(depth_td, ref)


def floyd_warshall_depth(nv, s, d, source):
    INF = 10 ** 9
    D = np.full((nv, nv), INF, np.int64)
    np.fill_diagonal(D, 0)
    for a, b in zip(s, d):
        D[a, b] = D[b, a] = 1
    for k in range(nv):
        D = np.minimum(D, D[:, [k]] + D[[k], :])
    row = D[source]
    return np.where(row >= INF, -1, row).astype(np.int32)


def test_depth_brute_force_tiny_graphs(orc):
    rng = np.random.default_rng(7)
    for trial in range(40):
        nv = int(rng.integers(2, 13))
        pairs = {(int(a), int(b)) for a, b in rng.integers(0, nv, size=(int(rng.integers(1, 20)), 2))
                 if a != b}
        if not pairs:
            continue
        s = np.array([p[0] for p in pairs], np.int64)
        d = np.array([p[1] for p in pairs], np.int64)
        off, adj, m = csr_of(orc, nv, s, d)
        src = int(s[0])
        ref = floyd_warshall_depth(nv, s, d, src)
        _, depth, _ = orc.beamer(off, adj, m, src, alpha=int(rng.integers(1, 20)),
                                 beta=int(rng.integers(1, 20)))
        np.testing.assert_array_equal(depth, ref)


# ------------------------------------------- hand-derived direction sequences

def test_beamer_sequence_path(orc):
    # 2M = 198.  TD while 2 <= edges_to_check/15 (levels 0..84), then BU while
    # each step wakes one vertex (awake 1 >= old 1), ending with an empty step.
    off, adj, m = csr_of(orc, *path_graph(100))
    _, depth, lv = orc.beamer(off, adj, m, 0)
    assert dirs_of(lv) == [TD] * 85 + [BU] * 15
    assert depth.tolist() == list(range(100))
    assert lv[85]["edges_to_check"] == 29


def test_beamer_sequence_star(orc):
    off, adj, m = csr_of(orc, *star_graph(100))
    _, depth, lv = orc.beamer(off, adj, m, 1)          # from a leaf
    assert dirs_of(lv) == [TD, BU, BU]
    assert [r["result"] for r in lv] == [99, 98, 0]
    assert depth[0] == 1 and depth[1] == 0 and (depth[2:] == 2).all()
    _, depth, lv = orc.beamer(off, adj, m, 0)          # from the centre
    assert dirs_of(lv) == [BU, BU]
    assert [r["result"] for r in lv] == [99, 0]


def test_beamer_sequence_complete_and_two_components(orc):
    off, adj, m = csr_of(orc, *complete_graph(20))
    _, _, lv = orc.beamer(off, adj, m, 0)
    assert dirs_of(lv) == [TD, BU]
    assert lv[0]["result"] == 19 * 19
    off, adj, m = csr_of(orc, *two_components())
    parent, depth, lv = orc.beamer(off, adj, m, 0)
    assert dirs_of(lv) == [BU, BU, BU]          # 2M/15 = 0, so scout 1 > 0
    assert depth.tolist() == [0, 1, 2, -1, -1]
    assert parent[3] == -1 and parent[4] == -1  # clear(): -deg = -1


def test_planted_mistakes_are_caught(orc):
    off, adj, m = csr_of(orc, *path_graph(100))
    soff, sadj, sm = csr_of(orc, *star_graph(100))
    with orc.mutation(1):   # awake > old instead of >=
        assert dirs_of(orc.beamer(off, adj, m, 0)[2]) != [TD] * 85 + [BU] * 15
    with orc.mutation(2):   # edges_to_check not charged
        assert dirs_of(orc.beamer(off, adj, m, 0)[2]) != [TD] * 85 + [BU] * 15
    with orc.mutation(3):   # scout counts vertices
        assert dirs_of(orc.beamer(soff, sadj, sm, 1)[2]) != [TD, BU, BU]
    with orc.mutation(4):   # single-phase bottom-up
        assert orc.beamer(off, adj, m, 0)[1].tolist() != list(range(100))
    # check(): parent a neighbour but at the wrong depth
    parent, _, _ = orc.beamer(soff, sadj, sm, 1)
    bad = parent.copy()
    bad[0] = 2                       # centre's parent = a depth-2 leaf
    assert not orc.check(soff, sadj, 1, bad)
    with orc.mutation(5):
        assert orc.check(soff, sadj, 1, bad)


# ------------------------------------------------------- check / counters

def test_check_accepts_valid_and_rejects_corrupt(orc):
    nv, s, d = orc.make_rmat(10, 16, seed=3)
    off, adj, m = csr_of(orc, nv, s, d)
    root = int(orc.pick_roots(off, 1)[0])
    parent, depth, _ = orc.beamer(off, adj, m, root)
    assert orc.check(off, adj, root, parent)
    p = parent.copy()
    v = int(np.flatnonzero(depth == 2)[0])
    p[v] = root if root not in adj[off[v]:off[v + 1]] else p[v]
    non_nb = [u for u in range(nv) if u not in set(adj[off[v]:off[v + 1]].tolist()) and u != v][0]
    p[v] = non_nb
    assert not orc.check(off, adj, root, p)
    p = parent.copy()
    p[root] = int(adj[off[root]])
    assert not orc.check(off, adj, root, p)
    unreached = np.flatnonzero(depth < 0)
    if len(unreached):
        p = parent.copy()
        p[unreached[0]] = root
        assert not orc.check(off, adj, root, p)


def test_count_traversed_is_component_edge_count(orc):
    nv, s, d = orc.make_rmat(10, 16, seed=4)
    off, adj, m = csr_of(orc, nv, s, d)
    A = sp.coo_matrix((np.ones(len(s)), (s, d)), shape=(nv, nv))
    _, lab = connected_components(A, directed=False)
    for root in orc.pick_roots(off, 3):
        parent, _, _ = orc.beamer(off, adj, m, root)
        comp = lab == lab[root]
        assert orc.count_traversed(off, parent) == int(comp[s].sum())


def test_level_log_is_consistent(orc):
    nv, s, d = orc.make_rmat(12, 16, seed=0)
    off, adj, m = csr_of(orc, nv, s, d)
    root = int(orc.pick_roots(off, 1)[0])
    parent, depth, lv = orc.beamer(off, adj, m, root)
    deg = np.diff(off)
    for r in lv:
        lvl = r["level"]
        assert r["frontier"] == int((depth == lvl).sum())
        assert r["frontier_deg"] == int(deg[depth == lvl].sum())
        assert r["next"] == int((depth == lvl + 1).sum())
        assert r["next_deg"] == int(deg[depth == lvl + 1].sum())
        assert r["result"] == (r["next_deg"] if r["dir"] == TD else r["next"])
    assert 2 * orc.count_traversed(off, parent) == int(deg[depth >= 0].sum())


def test_pick_roots_skip_isolated(orc):
    # 2^24 vertices would be too large here; check the rule on a small graph
    nv = 1 << 12
    s = np.array([9797422 % nv, 5], np.int64)
    d = np.array([1, 6], np.int64)
    off, adj = orc.build_csr(nv, s, d)
    first = [int(v) % nv for v in orc.lcg_outputs(0, 20000)]
    roots = orc.pick_roots(off, 3)
    deg = np.diff(off)
    expect = [v for v in first if deg[v] > 0][:3]
    assert roots.tolist() == expect


def test_aggregate_matches_scipy(orc):
    x = np.array([1.5e9, 2.0e9, 0.7e9, 3.3e9])
    g = orc.aggregate(x)
    assert g["hmean"] == pytest.approx(hmean(x), rel=1e-12)
    assert g["stddev"] == pytest.approx(np.std(x, ddof=1), rel=1e-12)
    assert g["mean"] == pytest.approx(x.mean()) and g["min"] == x.min() and g["max"] == x.max()


def test_bench_aggregate_is_aggregate_stats(orc):
    """The aggregate the benchmark line reports (aggregate_stats::compute,
    combined.cc:108-146): closed forms on three values, then the oracle's
    restatement on 64 values spread like per-root GTEPS."""
    import bench
    g = bench.aggregate([1.0, 2.0, 4.0])
    assert g["n"] == 3 and g["min"] == 1.0 and g["max"] == 4.0
    assert g["hmean"] == pytest.approx(12 / 7, rel=1e-15)
    assert g["mean"] == pytest.approx(7 / 3, rel=1e-15)
    assert g["stddev"] == pytest.approx(np.sqrt(7 / 3), rel=1e-15)  # sum of squares 14/3, N-1 = 2
    assert bench.hmean([2.0, 2.0]) == 2.0
    x = np.random.default_rng(3).uniform(0.3e9, 3.0e9, 64)
    b, o = bench.aggregate(x), orc.aggregate(x)
    for k in ("hmean", "mean", "stddev", "min", "max"):
        assert b[k] == pytest.approx(o[k], rel=1e-12), k


# ------------------------------------------------------ bottom-up head index

def test_head_index_matches_lexsort(orc):
    """head[v] = highest-degree neighbour, first in CSR order among equals:
    restated with numpy's lexsort over (vertex, -degree, position)."""
    nv, s, d = orc.make_rmat(12, 16, seed=4)
    s = np.concatenate([s, [7, 7]])      # a self-loop and a duplicate stay in the CSR
    d = np.concatenate([d, [7, int(d[0])]])
    off, adj = orc.build_csr(nv, s, d)
    deg = np.diff(off)
    head = orc.head_index(off, adj)
    owner = np.repeat(np.arange(nv), deg)
    order = np.lexsort((np.arange(len(adj)), -deg[adj], owner))
    expect = np.full(nv, -1, np.int64)
    has = deg > 0
    expect[has] = adj[order][off[:-1][has]]
    np.testing.assert_array_equal(head, expect)
    # brute force on a few vertices
    for v in np.flatnonzero(has)[:50]:
        nb = adj[off[v]:off[v + 1]]
        assert head[v] == nb[int(np.argmax(deg[nb]))]


def test_head_first_scan_changes_only_edges_examined(orc):
    nv, s, d = orc.make_rmat(14, 16, seed=2)
    off, adj, m = csr_of(orc, nv, s, d)
    head = orc.head_index(off, adj)
    for root in orc.pick_roots(off, 3):
        p0, d0, l0 = orc.beamer(off, adj, m, int(root))
        p1, d1, l1 = orc.beamer(off, adj, m, int(root), head=head)
        np.testing.assert_array_equal(d0, d1)
        np.testing.assert_array_equal(p0, p1)
        assert len(l0) == len(l1)
        for a, b in zip(l0, l1):
            for k in a:
                if k != "edges_examined" or a["dir"] == TD:
                    assert a[k] == b[k], k
        # closed form of the head-first count for every bottom-up step:
        # per scanned vertex 1 for the head, plus the in-order scan when the
        # head was not yet visited
        deg = np.diff(off)
        for rec in l1:
            if rec["dir"] != BU:
                continue
            L = rec["level"]
            scanned = (deg > 0) & ((d1 > L) | (d1 < 0))
            hv = scanned & (d1[np.maximum(head, 0)] <= L) & (d1[np.maximum(head, 0)] >= 0)
            want = int(scanned.sum())
            for v in np.flatnonzero(scanned & ~hv):
                nb = adj[off[v]:off[v + 1]]
                seen = (d1[nb] >= 0) & (d1[nb] <= L)
                want += int(np.argmax(seen)) + 1 if seen.any() else len(nb)
            assert rec["edges_examined"] == want


# ------------------------------------------------- all-core CPU baseline

@pytest.mark.parametrize("scale,ef,seed", [(11, 8, 0), (13, 16, 4), (12, 2, 7)])
def test_parallel_beamer_matches_serial(orc, scale, ef, seed):
    # the all-core restatement takes the serial one's direction at every
    # step, builds a valid tree and reaches the same vertices
    nv, s, d = orc.make_rmat(scale, ef, seed=seed)
    off, adj, m = csr_of(orc, nv, s, d)
    for root in orc.pick_roots(off, 4):
        parent, dirs = orc.beamer_omp(off, adj, m, root)
        ref_parent, _, levels = orc.beamer(off, adj, m, root)
        assert dirs.tolist() == [lv["dir"] for lv in levels]
        assert orc.check(off, adj, root, parent)
        assert orc.count_traversed(off, parent) == orc.count_traversed(off, ref_parent)
        np.testing.assert_array_equal(parent >= 0, ref_parent >= 0)


def test_parallel_beamer_on_special_graphs(orc):
    for nv, s, d in (path_graph(200), star_graph(50), complete_graph(12), two_components()):
        off, adj = orc.build_csr(nv, s, d)
        m = len(s)
        parent, dirs = orc.beamer_omp(off, adj, m, 0)
        _, _, levels = orc.beamer(off, adj, m, 0)
        assert dirs.tolist() == [lv["dir"] for lv in levels]
        assert orc.check(off, adj, 0, parent)
