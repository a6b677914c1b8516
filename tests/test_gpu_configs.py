"""GPU parity at the workloads' own sizes, through the C ABI.

RMAT scale 24 (the metric's graph), scale 22, the 64-root bit-parallel
search at scale 23, the 256-graph batch (scales 10-16, edge factors 8-32) and
the scale-26 stress graphs (RMAT, Erdos-Renyi, the 4096 x 4096 lattice).
Graphs come from the device generator, which the generator tests pin to the
oracle's (and the oracle's to the reference generator) bit for bit; at these
sizes the oracle's search functions check the device's results: depth
bit-exact against the textbook BFS, step records against the oracle's
run_beamer restatement, parents through the oracle's check().  Where the
oracle would be too slow (scale 26) the checks are the size-independent
ones: hybrid_bfs::check on the device, traversed edges = (sum of reached
degrees) / 2, and the step records summing to the reached set."""
import os

import numpy as np
import pytest

import beedrill_amd as bd

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "beedrill_amd", "models", "direction_mlp_32x32.f32")
LEVEL_KEYS = ("level", "dir", "frontier", "frontier_deg", "unvisited", "unvisited_deg",
              "edges_to_check", "result", "next", "next_deg", "max_frontier_deg",
              "edges_examined")


def offsets(ctx):
    """The device CSR's offsets alone (the column ids stay on the device)."""
    nv, _, _ = ctx.info()
    off = np.empty(nv + 1, np.int64)
    ctx._ok(ctx._L.bd_graph_download(ctx._h, off.ctypes.data, None))
    return off


def assert_records(gpu, ref):
    assert len(gpu) == len(ref)
    for g, o in zip(gpu, ref):
        for k in LEVEL_KEYS:
            assert g[k] == o[k], (k, g, o)


def size_independent(ctx, off, root, st):
    """hybrid_bfs::check on the device; traversed edges from the reached
    vertices' degrees; the records account for every reached vertex."""
    assert ctx.check(root)
    depth, _ = ctx.fetch(parent=False)
    deg = np.diff(off)
    reached = depth >= 0
    assert st["reached"] == int(reached.sum())
    assert st["traversed_edges"] == int(deg[reached].sum()) // 2
    assert ctx.count_traversed() == st["traversed_edges"]
    lv = ctx.levels()
    assert 1 + sum(r["next"] for r in lv) == st["reached"]
    assert lv[-1]["next"] == 0
    return depth, lv


@pytest.fixture(scope="module")
def s24(ctx, orc):
    """RMAT scale 24, edge factor 16, A/B/C/D .57/.19/.19/.05, seed 0 (the
    metric's graph) and the head index the bottom-up step reads."""
    m = ctx.generate_rmat(24, 16, seed=0)
    off, adj = ctx.csr()
    assert len(off) - 1 == 1 << 24 and 0.9 * (16 << 24) < m < 16 << 24
    return off, adj, m, orc.head_index(off, adj)


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_s24_beamer_bit_exact(ctx, orc, s24, k):
    """Roots 1-4 of the metric's 64: depth bit-exact, every step record
    (direction, frontier, Beamer counters, edges examined) equal to the
    oracle's run_beamer, parents accepted by the oracle's check()."""
    off, adj, m, head = s24
    ctx.generate_rmat(24, 16, seed=0)  # resident again if another test replaced it
    roots = ctx.pick_roots(4)
    np.testing.assert_array_equal(roots, orc.pick_roots(off, 4))
    root = int(roots[k])
    st = ctx.bfs(root, 15, 18, "beamer")
    depth, parent = ctx.fetch()
    ref_parent, ref_depth, ref_lv = orc.beamer(off, adj, m, root, head=head)
    np.testing.assert_array_equal(depth, ref_depth)
    assert orc.check(off, adj, root, parent)
    assert ctx.check(root)
    assert st["traversed_edges"] == orc.count_traversed(off, ref_parent)
    assert_records(ctx.levels(), ref_lv)


def test_s24_learned_policy_bit_exact(ctx, orc, s24):
    """The shipped network on the metric's graph: directions may differ from
    Beamer's, depths may not."""
    off, adj, m, _ = s24
    ctx.generate_rmat(24, 16, seed=0)
    ctx.mlp_set(np.fromfile(MODEL, np.float32), 32, 32)
    for root in ctx.pick_roots(4):
        root = int(root)
        st = ctx.bfs(root, 15, 18, "learned")
        depth, parent = ctx.fetch()
        np.testing.assert_array_equal(depth, orc.bfs_depth(off, adj, root))
        assert orc.check(off, adj, root, parent)
        assert ctx.check(root)
        assert 1 + sum(r["next"] for r in ctx.levels()) == st["reached"]


def test_s24_check_rejects_wrong_trees(ctx, orc, s24):
    """The FAIL half of hybrid_bfs::check (hybrid_bfs.cc:329-376): a tree
    checked from another source, and planted mistakes in the parent array,
    are rejected exactly when the oracle's check rejects them."""
    off, adj, m, _ = s24
    ctx.generate_rmat(24, 16, seed=0)
    roots = [int(r) for r in ctx.pick_roots(2)]
    ctx.bfs(roots[0])
    assert ctx.check(roots[0])
    assert not ctx.check(roots[1])
    depth, parent = ctx.fetch()
    assert ctx.check_parents(roots[0], parent)
    assert not ctx.check_parents(roots[1], parent)
    rng = np.random.default_rng(24)
    deg = np.diff(off)
    reached = np.flatnonzero((depth > 1) & (deg > 1))
    for mut in range(6):
        p = parent.copy()
        v = int(rng.choice(reached))
        nb = adj[off[v]:off[v + 1]]
        if mut == 0:    # a neighbour at the same or a deeper level
            same = nb[depth[nb] >= depth[v]]
            p[v] = int(same[0]) if len(same) else roots[0]
        elif mut == 1:  # not a neighbour at all
            p[v] = int(np.flatnonzero(depth == depth[v] - 1)[-1])
            if p[v] in set(nb.tolist()):
                p[v] = roots[0] if v != roots[0] else 0
        elif mut == 2:  # a reached vertex marked unreached
            p[v] = -1
        elif mut == 3:  # an unreached vertex given a parent
            u = int(np.flatnonzero(depth < 0)[0])
            p[u] = roots[0]
        elif mut == 4:  # the source not its own parent
            p[roots[0]] = int(adj[off[roots[0]]])
        else:           # a second root
            p[v] = v
        expect = orc.check(off, adj, roots[0], p)
        assert not expect, mut
        assert ctx.check_parents(roots[0], p) == expect, mut


def test_s22_records_and_depths(ctx, orc):
    """RMAT scale 22 (4,194,304 vertices, 67,108,864 generated edges): 8 of
    the lcg(0) roots, depth bit-exact and records equal to the oracle's."""
    m = ctx.generate_rmat(22, 16)
    off, adj = ctx.csr()
    head = orc.head_index(off, adj)
    for root in ctx.pick_roots(8):
        root = int(root)
        st = ctx.bfs(root)
        depth, parent = ctx.fetch()
        _, ref_depth, ref_lv = orc.beamer(off, adj, m, root, head=head)
        np.testing.assert_array_equal(depth, ref_depth)
        assert orc.check(off, adj, root, parent)
        assert_records(ctx.levels(), ref_lv)
        assert st["reached"] == int((depth >= 0).sum())


def test_s23_multi_root_64(ctx, orc):
    """The 64-root bit-parallel search at RMAT scale 23: 8 of the 64 depth
    rows against the oracle's textbook BFS, every row's reached count against
    a single-root search from the same root."""
    ctx.generate_rmat(23, 16)
    off, adj = ctx.csr()
    roots = ctx.pick_roots(64)
    dp, ms, levels = ctx.multi_root(roots)
    assert dp.shape == (64, len(off) - 1) and ms > 0
    for r in (0, 1, 9, 22, 31, 40, 57, 63):
        np.testing.assert_array_equal(dp[r], orc.bfs_depth(off, adj, int(roots[r])), err_msg=str(r))
    for r in range(64):
        st = ctx.bfs(int(roots[r]))
        assert int((dp[r] >= 0).sum()) == st["reached"], r
    assert levels == int(dp.max()) + 1


def test_256_uneven_graphs(ctx, orc):
    """The many-graph batch as configured: 256 seeded RMAT graphs, scale
    uniform in 10-16, edge factor uniform in 8-32, one root each, searched in
    one go; every graph's depths equal the oracle's BFS of that graph alone
    and its parents pass the oracle's check()."""
    rng = np.random.default_rng(7)
    graphs, roots = [], []
    for g in range(256):
        scale, ef = int(rng.integers(10, 17)), int(rng.integers(8, 33))
        ctx.generate_rmat(scale, ef, seed=1000 + g)
        off, adj = ctx.csr()
        graphs.append((off, adj, ctx.info()[2]))
        roots.append(int(ctx.pick_roots(1)[0]))
    vbase = ctx.upload_union(graphs)
    dp, ms = ctx.multi_graph(vbase, roots)
    assert ms > 0
    parent = ctx.fetch()[1]
    for g, (off, adj, _) in enumerate(graphs):
        np.testing.assert_array_equal(dp[vbase[g]:vbase[g + 1]],
                                      orc.bfs_depth(off, adj, roots[g]), err_msg=str(g))
        p = parent[vbase[g]:vbase[g + 1]]
        assert orc.check(off, adj, roots[g], np.where(p >= 0, p - vbase[g], -1)), g


@pytest.mark.parametrize("name", ["rmat26", "er26", "lattice4096"])
def test_scale26_stress_graphs(ctx, name):
    """RMAT scale 26 (1.07 G generated edges), Erdos-Renyi with 2^26 vertices
    and 2^30 edges, and the 4096 x 4096 lattice: two lcg(0) roots each with
    Beamer and with the learned policy, checked by the size-independent
    properties; both policies reach the same vertices."""
    if name == "rmat26":
        ctx.generate_rmat(26, 16, seed=0)
    elif name == "er26":
        ctx.generate_er(1 << 26, 16 << 26, seed=0)
    else:
        ctx.generate_lattice(4096, 4096)
    off = offsets(ctx)
    ctx.mlp_set(np.fromfile(MODEL, np.float32), 32, 32)
    for root in ctx.pick_roots(2):
        root = int(root)
        st = ctx.bfs(root, 15, 18, "beamer")
        depth, lv = size_independent(ctx, off, root, st)
        st2 = ctx.bfs(root, 15, 18, "learned")
        depth2, _ = size_independent(ctx, off, root, st2)
        np.testing.assert_array_equal(depth, depth2)
        if name == "lattice4096":
            # a lattice's depths are Manhattan distances
            r, c = divmod(root, 4096)
            rr, cc = np.divmod(np.arange(1 << 24), 4096)
            np.testing.assert_array_equal(depth, np.abs(rr - r) + np.abs(cc - c))
