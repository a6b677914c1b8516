"""GPU parity: the HIP path, called through the C ABI, against the oracle on
the same seeded inputs.  Integer work, so every comparison is bit-exact."""
import os
import tempfile

import numpy as np
import pytest

import beedrill_amd as bd
from conftest import complete_graph, path_graph, star_graph, two_components

pytestmark = pytest.mark.gpu

LEVEL_KEYS = ("level", "dir", "frontier", "frontier_deg", "unvisited", "unvisited_deg",
              "edges_to_check", "result", "next", "next_deg", "max_frontier_deg",
              "edges_examined")


def oracle_csr(orc, nv, s, d):
    off, adj = orc.build_csr(nv, s, d)
    return off, adj, len(s)


def assert_same_levels(gpu_levels, orc_levels, keys=LEVEL_KEYS):
    assert len(gpu_levels) == len(orc_levels)
    for g, o in zip(gpu_levels, orc_levels):
        for k in keys:
            assert g[k] == o[k], (k, g, o)


def run_and_compare(ctx, orc, off, adj, m, root, policy="beamer", alpha=15, beta=18,
                    levels=True):
    st = ctx.bfs(root, alpha, beta, policy)
    depth, parent = ctx.fetch()
    # the bottom-up kernel reads each vertex's head index first unless told not to
    head = None if ctx.get_tuning()["bu_head"] == 2 else orc.head_index(off, adj)
    ref_parent, ref_depth, ref_lv = orc.beamer(off, adj, m, root, alpha, beta, head=head)
    np.testing.assert_array_equal(depth, ref_depth)
    assert orc.check(off, adj, root, parent)
    assert ((parent < 0) == (depth < 0)).all()
    assert st["traversed_edges"] == orc.count_traversed(off, ref_parent)
    assert ctx.count_traversed() == st["traversed_edges"]
    if levels and policy == "beamer":
        assert_same_levels(ctx.levels(), ref_lv)
    return st


# ------------------------------------------------------------- generator

@pytest.mark.parametrize("scale,ef,seed,abcd", [(10, 16, 0, (0.57, 0.19, 0.19, 0.05)),
                                                (12, 8, 3, (0.45, 0.25, 0.25, 0.05)),
                                                (4, 16, 1, (0.57, 0.19, 0.19, 0.05))])
def test_device_rmat_matches_oracle(ctx, orc, scale, ef, seed, abcd):
    m = ctx.generate_rmat(scale, ef, *abcd, seed=seed)
    s, d = ctx.edges()
    nv, os_, od = orc.make_rmat(scale, ef, *abcd, seed=seed)
    assert m == len(os_)
    np.testing.assert_array_equal(s, os_)
    np.testing.assert_array_equal(d, od)
    off, cols = ctx.csr()
    ooff, oadj = orc.build_csr(nv, os_, od)
    np.testing.assert_array_equal(off, ooff)
    np.testing.assert_array_equal(cols, oadj)


def test_device_er_and_lattice_match_oracle(ctx, orc):
    ctx.generate_er(5000, 40000, seed=1)
    s, d = ctx.edges()
    _, os_, od = orc.make_er(5000, 40000, seed=1)
    np.testing.assert_array_equal(s, os_)
    np.testing.assert_array_equal(d, od)
    ctx.generate_lattice(20, 30, seed=2)
    s, d = ctx.edges()
    _, os_, od = orc.make_lattice(20, 30, seed=2)
    np.testing.assert_array_equal(s, os_)
    np.testing.assert_array_equal(d, od)


# ------------------------------------------------------ hand-built graphs

@pytest.mark.parametrize("g,root", [(path_graph(100), 0), (star_graph(100), 1),
                                    (star_graph(100), 0), (complete_graph(20), 0),
                                    (two_components(), 0), (two_components(), 3)])
def test_small_graphs_match_oracle(ctx, orc, g, root):
    nv, s, d = g
    ctx.from_edges(nv, s, d)
    off, adj, m = oracle_csr(orc, nv, s, d)
    goff, gcols = ctx.csr()
    np.testing.assert_array_equal(goff, off)
    np.testing.assert_array_equal(gcols, adj)
    run_and_compare(ctx, orc, off, adj, m, root)
    assert ctx.check(root)


def test_depths_from_deep_parent_chains(ctx, orc):
    # The steps write parents only and depths come from the parent tree on
    # readout: a 20000-vertex path under a random relabelling has parent
    # chains of up to 19999 links running both ways through the id space.
    n = 20000
    perm = np.random.default_rng(7).permutation(n).astype(np.int64)
    s, d = perm[:-1], perm[1:]
    ctx.from_edges(n, s, d)
    off, adj, m = oracle_csr(orc, n, s, d)
    for root in (int(perm[0]), int(perm[n // 3])):
        run_and_compare(ctx, orc, off, adj, m, root)
        depth, _ = ctx.fetch()
        assert depth.max() == n - 1 - (0 if root == perm[0] else n // 3)


def test_isolated_source_and_self_loops(ctx, orc):
    # vertex 5 isolated; a self-loop on 2 and a duplicate edge stay as given
    nv = 6
    s = np.array([0, 1, 2, 0], np.int64)
    d = np.array([1, 2, 2, 1], np.int64)
    ctx.from_edges(nv, s, d)
    off, adj, m = oracle_csr(orc, nv, s, d)
    np.testing.assert_array_equal(ctx.csr()[1], adj)
    run_and_compare(ctx, orc, off, adj, m, 0)
    st = ctx.bfs(5)
    depth, parent = ctx.fetch()
    assert depth.tolist() == [-1, -1, -1, -1, -1, 0] and parent[5] == 5
    assert st["steps"] == 1 and st["traversed_edges"] == 0


def test_graph_without_edges(ctx, orc):
    # The reference's loaders refuse such a graph (edge_list.cc:104), and its
    # root picking would never end on one; an edge list with no edges still
    # makes a graph of isolated vertices here, searched like any other.
    nv = 70  # one visited word and a ragged second
    empty = np.zeros(0, np.int64)
    ctx.from_edges(nv, empty, empty)
    off, adj, m = oracle_csr(orc, nv, empty, empty)
    goff, gcols = ctx.csr()
    np.testing.assert_array_equal(goff, off)
    assert gcols.size == 0 and m == 0
    for root in (0, 63, 69):
        st = run_and_compare(ctx, orc, off, adj, m, root)
        assert st["steps"] == 1 and st["traversed_edges"] == 0
        assert ctx.check(root) and not ctx.check((root + 1) % nv)
        run_and_compare(ctx, orc, off, adj, m, root, policy="td_only")
    with pytest.raises(bd.BeedrillError) as e:
        ctx.pick_roots(1)
    assert e.value.code == -4


# Degrees on either side of every cut the kernels make by list length: task
# slices from 16 entries, one lane per list below 64, hub lists from 256, the
# batch's pull slices from 512 (1024 a slice), task slices of 1536, hub
# slices of 4096 from 8192.
THRESHOLD_DEGREES = (15, 16, 17, 63, 64, 65, 255, 256, 257, 511, 512, 513, 1023, 1024, 1025,
                     1535, 1536, 1537, 4095, 4096, 4097, 8191, 8192, 8193, 12289)


def threshold_graph(seed=11, leaves=16384, extra=20000, isolated=37):
    """Hubs 0..24 of exactly THRESHOLD_DEGREES neighbours, drawn without
    repeats from a shared pool of leaves (so the hubs meet only through
    leaves), random leaf-leaf edges, isolated vertices at the end; the edges
    in shuffled order and direction."""
    rng = np.random.default_rng(seed)
    nh = len(THRESHOLD_DEGREES)
    s, d = [], []
    for h, deg in enumerate(THRESHOLD_DEGREES):
        s.append(np.full(deg, h, np.int64))
        d.append(nh + rng.choice(leaves, deg, replace=False).astype(np.int64))
    s.append(nh + rng.integers(0, leaves, extra))
    d.append(nh + rng.integers(0, leaves, extra))
    s, d = np.concatenate(s), np.concatenate(d)
    flip = rng.random(s.size) < 0.5
    s, d = np.where(flip, d, s), np.where(flip, s, d)
    order = rng.permutation(s.size)
    return nh + leaves + isolated, s[order], d[order]


@pytest.mark.parametrize("vertex_order", [1, 2, 3, 4])
def test_lists_at_every_kernel_threshold(ctx, orc, vertex_order):
    """Searches over lists at and beside each length threshold, on every
    search layout and through several launch shapes: depths, parents and
    Beamer's step records are the oracle's (a list cut one entry short or
    long loses or repeats neighbours)."""
    nv, s, d = threshold_graph()
    t0 = ctx.get_tuning()
    try:
        ctx.set_tuning(vertex_order=vertex_order)
        ctx.from_edges(nv, s, d)
        off, adj, m = oracle_csr(orc, nv, s, d)
        deg = np.diff(off)
        nh = len(THRESHOLD_DEGREES)
        for h, want in enumerate(THRESHOLD_DEGREES):
            assert deg[h] == want, (h, deg[h], want)
        hub_big, hub_small, leaf = nh - 1, 0, int(np.argmax(deg[nh:] > 3)) + nh
        for root in (hub_big, hub_small, leaf):
            run_and_compare(ctx, orc, off, adj, m, root)
            for policy in ("td_only", "bu_only"):
                run_and_compare(ctx, orc, off, adj, m, root, policy=policy)
            for alpha, beta in ((2, 2), (1000, 1)):
                run_and_compare(ctx, orc, off, adj, m, root, alpha=alpha, beta=beta)
        shapes = ({"bu_cap": 8, "bu_batch": 1}, {"bu_batch": 4, "task_edges": 128},
                  {"td_split_edges": 1}, {"bu_head": 2}, {"wave_threshold": 16})
        for shape in shapes:
            ctx.set_tuning(**shape)
            for root in (hub_big, leaf):
                run_and_compare(ctx, orc, off, adj, m, root)
            ctx.set_tuning(**{k: t0[k] for k in shape})
        roots = ctx.pick_roots(64)
        dp, _, _ = ctx.multi_root(roots)
        for r in range(64):
            np.testing.assert_array_equal(dp[r], orc.bfs_depth(off, adj, int(roots[r])),
                                          err_msg=str(r))
    finally:
        ctx.set_tuning(**t0)


# ------------------------------------------------------------ RMAT scale 16

@pytest.fixture(scope="module")
def s16(ctx, orc):
    """Config: RMAT scale 16, edge factor 16, A/B/C/D .57/.19/.19/.05, seed 0."""
    ctx.generate_rmat(16, 16)
    off, cols = ctx.csr()
    _, m_, _ = ctx.info()
    return off, cols, ctx.info()[2]


def test_s16_beamer_parity(ctx, orc, s16):
    ctx.generate_rmat(16, 16)
    off, adj, m = s16
    roots = ctx.pick_roots(8)
    np.testing.assert_array_equal(roots, orc.pick_roots(off, 8))
    for root in roots:
        run_and_compare(ctx, orc, off, adj, m, int(root))


@pytest.mark.parametrize("policy", ["td_only", "bu_only"])
def test_s16_other_policies_same_depth(ctx, orc, s16, policy):
    ctx.generate_rmat(16, 16)
    off, adj, m = s16
    root = int(ctx.pick_roots(1)[0])
    st = run_and_compare(ctx, orc, off, adj, m, root, policy=policy, levels=False)
    lv = ctx.levels()
    want = bd.DIR_TD if policy == "td_only" else bd.DIR_BU
    assert all(r["dir"] == want for r in lv)
    if policy == "td_only":
        _, ref_depth = orc.td_only(off, adj, m, root)
        np.testing.assert_array_equal(ctx.fetch()[0], ref_depth)
    assert st["steps"] == len(lv)


def test_s16_scripted_directions_same_depth(ctx, orc, s16):
    ctx.generate_rmat(16, 16)
    off, adj, m = s16
    rng = np.random.default_rng(0)
    for root in ctx.pick_roots(3):
        ctx.set_script(rng.integers(0, 2, size=32).astype(np.int32))
        run_and_compare(ctx, orc, off, adj, m, int(root), policy="script", levels=False)


def test_bitmap_top_down_over_hubs_every_layout(ctx, orc):
    """A top-down step right after a bottom-up one reads its frontier from
    the bitmap; when that frontier holds hubs (lists of 256 or more) and has
    2^20 edges or more, the hubs are taken off the workgroups' ranges and
    spread over the grid.  Scripted directions force such steps at levels 2
    and 3, with claims and with remote writes, on every search layout (RMAT
    scale 17): depths are the oracle's and the parent trees pass check()."""
    ctx.generate_rmat(17, 16)
    off, adj = ctx.csr()
    m = ctx.info()[2]
    roots = [int(r) for r in ctx.pick_roots(2)]
    t0 = ctx.get_tuning()
    hub_steps = 0
    try:
        for order in (1, 2, 3, 4):
            for split in (2147483647, 1):
                ctx.set_tuning(vertex_order=order, td_split_edges=split)
                for script in ([0, 1, 0, 0], [0, 0, 1, 0]):
                    ctx.set_script(np.array(script + [0] * 28, np.int32))
                    for root in roots:
                        run_and_compare(ctx, orc, off, adj, m, root, policy="script",
                                        levels=False)
                        lv = ctx.levels()
                        k = script.index(1) + 1
                        if len(lv) > k and lv[k]["dir"] == bd.DIR_TD and \
                                lv[k]["max_frontier_deg"] >= 256 and \
                                lv[k]["frontier_deg"] >= 1 << 20:
                            hub_steps += 1
    finally:
        ctx.set_tuning(**t0)
    assert hub_steps >= 8


def test_s16_host_driven_levels_identical(ctx, orc, s16):
    ctx.generate_rmat(16, 16)
    off, adj, m = s16
    root = int(ctx.pick_roots(2)[1])
    t0 = ctx.get_tuning()
    try:
        ctx.set_tuning(level_driver=2)
        run_and_compare(ctx, orc, off, adj, m, root)
        ctx.set_tuning(level_driver=1, block=512, wave_threshold=8, task_threshold=256,
                       task_edges=128, bu_words=8, bu_cap=4)
        run_and_compare(ctx, orc, off, adj, m, root)
        ctx.set_tuning(block=128, wave_threshold=1000, task_threshold=64, task_edges=64,
                       bu_words=64, bu_cap=1, grid_per_cu=1)
        run_and_compare(ctx, orc, off, adj, m, root)
        # lists only, in CSR order: edges_examined is the in-order early-exit count
        ctx.set_tuning(bu_head=2, bu_words=32, bu_cap=32, block=256, grid_per_cu=8)
        run_and_compare(ctx, orc, off, adj, m, root)
        # one kernel carrying both directions; odd bottom-up grid
        ctx.set_tuning(bu_head=1, step_kernels=2)
        run_and_compare(ctx, orc, off, adj, m, root)
        # every bottom-up batching mode, with and without the head index
        for batch in (1, 4, 8):
            for head in (1, 2):
                ctx.set_tuning(bu_batch=batch, bu_head=head)
                run_and_compare(ctx, orc, off, adj, m, root)
        ctx.set_tuning(bu_head=1, bu_batch=8)
        ctx.set_tuning(step_kernels=1, bu_grid_per_cu=3, bu_words=5)
        run_and_compare(ctx, orc, off, adj, m, root)
        ctx.set_tuning(level_driver=2)
        run_and_compare(ctx, orc, off, adj, m, root)
    finally:
        ctx.set_tuning(**t0)


def _launches_follow(lv):
    # a level's launches start where the previous level's end
    starts = np.cumsum([0] + [r["launches"] for r in lv])[:-1]
    return [r["launch"] for r in lv] == starts.tolist()


def test_s16_remote_write_topdown(ctx, orc, s16):
    # Top-down steps with remote writes take two launches (candidate writes,
    # then the claim sweep) and give the oracle's depths and records: on the
    # fused kernel, one kernel per direction, host-driven and per-launch
    # timed, for the top-down-only policy, and with a threshold that splits
    # only the heavy steps.
    ctx.generate_rmat(16, 16)
    off, adj, m = s16
    roots = [int(r) for r in ctx.pick_roots(3)]
    t0 = ctx.get_tuning()
    try:
        ctx.set_tuning(td_split_edges=1)
        for root in roots:
            run_and_compare(ctx, orc, off, adj, m, root)
            lv = ctx.levels()
            assert all(r["launches"] == (2 if r["dir"] == bd.DIR_TD else 1) for r in lv)
            assert _launches_follow(lv)
        ctx.set_tuning(step_kernels=1)
        run_and_compare(ctx, orc, off, adj, m, roots[0])
        ctx.set_tuning(level_driver=2)
        run_and_compare(ctx, orc, off, adj, m, roots[0])
        ctx.set_tuning(level_driver=1, step_kernels=2, time_steps=3)
        run_and_compare(ctx, orc, off, adj, m, roots[1])
        lv = ctx.levels()
        lt = ctx.launch_times()
        for r in lv:
            assert r["ms"] == pytest.approx(lt[r["launch"]:r["launch"] + r["launches"]].sum())
        ctx.set_tuning(time_steps=2)
        run_and_compare(ctx, orc, off, adj, m, roots[1], policy="td_only", levels=False)
        _, ref_depth = orc.td_only(off, adj, m, roots[1])
        np.testing.assert_array_equal(ctx.fetch()[0], ref_depth)
        split = 20000
        ctx.set_tuning(td_split_edges=split)
        for root in roots:
            run_and_compare(ctx, orc, off, adj, m, root)
            lv = ctx.levels()
            for r in lv:
                heavy = r["dir"] == bd.DIR_TD and r["frontier_deg"] >= split
                assert r["launches"] == (2 if heavy else 1)
            assert _launches_follow(lv)
    finally:
        ctx.set_tuning(**t0)


def test_s16_remote_write_policies(ctx, orc, s16):
    # run_with_remote_writes: every step top-down with remote writes, the
    # levels of the top-down-only search.  run_with_remote_writes_hybrid:
    # Beamer's state machine with remote-write steps exactly where Beamer
    # goes bottom-up, so its records carry Beamer's level quantities.
    ctx.generate_rmat(16, 16)
    off, adj, m = s16
    head = orc.head_index(off, adj)
    same = ("level", "frontier", "frontier_deg", "unvisited", "unvisited_deg",
            "edges_to_check", "next", "next_deg", "max_frontier_deg")
    for root in [int(r) for r in ctx.pick_roots(3)]:
        run_and_compare(ctx, orc, off, adj, m, root, policy="remote_writes", levels=False)
        _, ref_depth = orc.td_only(off, adj, m, root)
        np.testing.assert_array_equal(ctx.fetch()[0], ref_depth)
        lv = ctx.levels()
        assert all(r["dir"] == bd.DIR_TD and r["launches"] == 2 for r in lv)
        assert _launches_follow(lv)
        run_and_compare(ctx, orc, off, adj, m, root, policy="remote_writes_hybrid", levels=False)
        lv = ctx.levels()
        _, _, ref_lv = orc.beamer(off, adj, m, root, 15, 18, head=head)
        assert_same_levels(lv, ref_lv, keys=same)
        assert [r["launches"] for r in lv] == [2 if o["dir"] == bd.DIR_BU else 1 for o in ref_lv]
        assert all(r["dir"] == bd.DIR_TD for r in lv)
    with pytest.raises(bd.BeedrillError):
        ctx.profile(int(ctx.pick_roots(1)[0]), policy="remote_writes")


@pytest.mark.parametrize("alg", ["beamer_hybrid", "remote_writes", "remote_writes_hybrid",
                                 "migrating_threads"])
def test_cli_algorithms_check(alg):
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "beedrill_amd", "bin", "hybrid_bfs")
    import subprocess
    r = subprocess.run([exe, "--generate", "graph500-scale12", "--num_trials", "3",
                        "--algorithm", alg, "--check_results"],
                       capture_output=True, text=True, timeout=120)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert out.count("PASS") >= 3 and "FAIL" not in out, out


@pytest.mark.parametrize("g,root", [(path_graph(100), 0), (star_graph(100), 1),
                                    (complete_graph(20), 0), (two_components(), 3)])
def test_small_graphs_remote_write_topdown(ctx, orc, g, root):
    nv, s, d = g
    ctx.from_edges(nv, s, d)
    off, adj, m = oracle_csr(orc, nv, s, d)
    t0 = ctx.get_tuning()
    try:
        ctx.set_tuning(td_split_edges=1)
        run_and_compare(ctx, orc, off, adj, m, root)
        run_and_compare(ctx, orc, off, adj, m, root, policy="td_only", levels=False)
        assert ctx.check(root)
    finally:
        ctx.set_tuning(**t0)


def test_s16_alpha_beta_sweep(ctx, orc, s16):
    ctx.generate_rmat(16, 16)
    off, adj, m = s16
    root = int(ctx.pick_roots(1)[0])
    for alpha, beta in [(1, 1), (2, 100), (15, 18), (100, 2), (10 ** 9, 10 ** 9)]:
        run_and_compare(ctx, orc, off, adj, m, root, alpha=alpha, beta=beta)


def test_profile_times_both_directions(ctx, orc, s16):
    ctx.generate_rmat(16, 16)
    off, adj, m = s16
    root = int(ctx.pick_roots(1)[0])
    ctx.profile(root)
    lv = ctx.levels()
    depth, parent = ctx.fetch()
    _, ref_depth, ref_lv = orc.beamer(off, adj, m, root, head=orc.head_index(off, adj))
    np.testing.assert_array_equal(depth, ref_depth)
    assert orc.check(off, adj, root, parent)
    assert_same_levels(lv, ref_lv)
    assert all(r["ms"] > 0 and r["ms_other"] > 0 for r in lv)


# ------------------------------------------------------------ el64 files

def test_el64_roundtrip_and_oracle_file(ctx, orc):
    nv, s, d = orc.make_rmat(10, 16, seed=9)
    off, adj, m = oracle_csr(orc, nv, s, d)
    with tempfile.TemporaryDirectory() as t:
        p = os.path.join(t, "g.el64")
        orc.write_el64(p, nv, s, d)
        ctx.load_el64(p)
        goff, gcols = ctx.csr()
        np.testing.assert_array_equal(goff, off)
        np.testing.assert_array_equal(gcols, adj)
        p2 = os.path.join(t, "g2.el64")
        ctx.save_el64(p2)
        with open(p, "rb") as f1, open(p2, "rb") as f2:
            assert f1.read() == f2.read()
        run_and_compare(ctx, orc, off, adj, m, int(orc.pick_roots(off, 1)[0]))


def test_el64_errors(ctx):
    with tempfile.TemporaryDirectory() as t:
        def write(name, header, payload=b""):
            p = os.path.join(t, name)
            with open(p, "wb") as f:
                f.write(header + payload)
            return p
        pairs = np.array([[0, 1], [1, 2]], np.int64).tobytes()
        good = b" --format el64 --num_edges 2 --num_vertices 3 --is_deduped\n"
        ctx.load_el64(write("ok", good, pairs))
        assert ctx.info() == (3, 4, 2)
        prefix = b" --form el64 --num_e 2 --num_v=3 --is_dedup\n"
        ctx.load_el64(write("prefix", prefix, pairs))
        for name, hdr, body, code in [
                ("short", good, pairs[:16], -5),
                ("notdedup", b" --format el64 --num_edges 2 --num_vertices 3\n", pairs, -5),
                ("fmt", b" --format el32 --num_edges 2 --num_vertices 3 --is_deduped\n", pairs, -5),
                ("size", b" --format el64 --num_edges 0 --num_vertices 3 --is_deduped\n", b"", -5),
                ("unknown", b" --format el64 --bogus --num_edges 2 --num_vertices 3 --is_deduped\n",
                 pairs, -5),
                ("range", b" --format el64 --num_edges 2 --num_vertices 2 --is_deduped\n", pairs, -6)]:
            with pytest.raises(bd.BeedrillError) as e:
                ctx.load_el64(write(name, hdr, body))
            assert e.value.code == code, name
        with pytest.raises(bd.BeedrillError) as e:
            ctx.load_el64(os.path.join(t, "missing"))
        assert e.value.code == -5


def test_argument_errors(ctx, orc):
    nv, s, d = path_graph(10)
    ctx.from_edges(nv, s, d)
    for kw, code in [(dict(source=10), -6), (dict(source=-1), -6), (dict(source=0, alpha=0), -1),
                     (dict(source=0, beta=0), -1)]:
        with pytest.raises(bd.BeedrillError) as e:
            ctx.bfs(**kw)
        assert e.value.code == code
    with pytest.raises(bd.BeedrillError) as e:
        ctx.from_edges(4, np.array([0]), np.array([7]))
    assert e.value.code == -6


# ------------------------------------------------------------ batched searches

def test_multi_root_matches_oracle(ctx, orc):
    """64 roots in one bit-parallel search: every root's depth array is the
    oracle's textbook BFS from that root."""
    ctx.generate_rmat(12, 16, seed=5)
    off, adj = ctx.csr()
    roots = ctx.pick_roots(64)
    dp, ms, levels = ctx.multi_root(roots)
    assert dp.shape == (64, len(off) - 1) and ms > 0
    for r, root in enumerate(roots):
        np.testing.assert_array_equal(dp[r], orc.bfs_depth(off, adj, int(root)), err_msg=str(r))
    assert levels == int(dp.max()) + 1
    # fewer roots, a repeated root, an isolated root, then the 64 again
    deg = np.diff(off)
    iso = int(np.flatnonzero(deg == 0)[0])
    few = np.array([roots[3], iso, roots[3], roots[7], 0], np.int64)
    dp2, _, _ = ctx.multi_root(few)
    for r, root in enumerate(few):
        np.testing.assert_array_equal(dp2[r], orc.bfs_depth(off, adj, int(root)))
    dp3, _, _ = ctx.multi_root(roots)
    np.testing.assert_array_equal(dp3, dp)


def test_multi_root_s16_and_lattice(ctx, orc):
    ctx.generate_rmat(16, 16)
    off, adj = ctx.csr()
    roots = ctx.pick_roots(64)
    dp, _, _ = ctx.multi_root(roots)
    for r in (0, 17, 63):
        np.testing.assert_array_equal(dp[r], orc.bfs_depth(off, adj, int(roots[r])))
    _, ms, _ = ctx.multi_root(roots, want_depths=False)
    assert ms > 0
    ctx.generate_lattice(40, 50, seed=1)  # long, thin levels: every level pushes or pulls little
    off, adj = ctx.csr()
    roots = np.array([0, 1999, 1000, 77], np.int64)
    dp, _, levels = ctx.multi_root(roots)
    for r, root in enumerate(roots):
        np.testing.assert_array_equal(dp[r], orc.bfs_depth(off, adj, int(root)))


def test_multi_root_every_layout_with_hubs(ctx, orc):
    """The batch on every search layout, on a graph with hubs (lists of 512
    entries or more, which a pull cuts into slices scanned by the whole grid
    and finishes in its second launch): every depth row is the oracle's, and
    the level records agree across layouts."""
    ctx.generate_rmat(16, 16, seed=3)
    off, adj = ctx.csr()
    assert int(np.diff(off).max()) >= 4 * 512
    roots = ctx.pick_roots(64)
    ref = [orc.bfs_depth(off, adj, int(r)) for r in roots]
    t0 = ctx.get_tuning()
    recs = None
    try:
        for order in (1, 2, 3, 4):
            ctx.set_tuning(vertex_order=order)
            dp, _, levels = ctx.multi_root(roots)
            for r in range(64):
                np.testing.assert_array_equal(dp[r], ref[r], err_msg=f"order {order} row {r}")
            rec = [(r["dir"], r["frontier"], r["frontier_deg"], r["next"]) for r in ctx.levels()]
            assert any(d == 1 for d, *_ in rec), "no pull level"
            if recs is None:
                recs = rec
            assert rec == recs, order
    finally:
        ctx.set_tuning(**t0)


def test_multi_root_graph_replays_match_host_driven(ctx, orc):
    """The level launches captured once in a HIP graph and replayed, against
    the same launches enqueued level by level from the host: the same depth
    rows and level records on every replay, with ordinary searches (eager
    work and synchronisation) between the replays."""
    ctx.generate_rmat(14, 16, seed=9)
    off, adj = ctx.csr()
    roots = ctx.pick_roots(64)
    t0 = ctx.get_tuning()
    try:
        ctx.set_tuning(level_driver=2)
        ref, _, ref_levels = ctx.multi_root(roots)
        ref_rec = [(r["dir"], r["frontier"], r["frontier_deg"], r["next"]) for r in ctx.levels()]
        ctx.set_tuning(level_driver=1)
        for rep in range(4):
            dp, ms, levels = ctx.multi_root(roots)
            assert levels == ref_levels and ms > 0
            np.testing.assert_array_equal(dp, ref)
            assert [(r["dir"], r["frontier"], r["frontier_deg"], r["next"])
                    for r in ctx.levels()] == ref_rec
            ctx.bfs(int(roots[rep]))
    finally:
        ctx.set_tuning(**t0)
    assert len(ref_rec) == ref_levels
    for r in (0, 31, 63):
        np.testing.assert_array_equal(ref[r], orc.bfs_depth(off, adj, int(roots[r])))


def test_multi_root_errors(ctx):
    nv, s, d = path_graph(10)
    ctx.from_edges(nv, s, d)
    for roots, code in [(np.zeros(0, np.int64), -1), (np.zeros(65, np.int64), -1),
                        (np.array([3, 10], np.int64), -6), (np.array([-1], np.int64), -6)]:
        with pytest.raises(bd.BeedrillError) as e:
            ctx.multi_root(roots)
        assert e.value.code == code


def test_multi_graph_matches_oracle(ctx, orc):
    """Many disjoint graphs of uneven size searched from their roots at once:
    each graph's depths are the oracle's BFS of that graph alone."""
    rng = np.random.default_rng(3)
    graphs, roots = [], []
    for g in range(24):
        scale, ef = int(rng.integers(4, 11)), int(rng.integers(8, 33))
        nv, s, d = orc.make_rmat(scale, ef, seed=100 + g)
        off, adj = orc.build_csr(nv, s, d)
        graphs.append((off, adj, len(s)))
        roots.append(int(orc.pick_roots(off, 1)[0]) if g != 5 else
                     int(np.flatnonzero(np.diff(off) == 0)[0]) if (np.diff(off) == 0).any() else 0)
    vbase = ctx.upload_union(graphs)
    dp, ms = ctx.multi_graph(vbase, roots)
    assert ms > 0
    for g, (off, adj, _) in enumerate(graphs):
        np.testing.assert_array_equal(dp[vbase[g]:vbase[g + 1]],
                                      orc.bfs_depth(off, adj, roots[g]), err_msg=str(g))
    parent = ctx.fetch()[1]
    for g, (off, adj, _) in enumerate(graphs):
        p = parent[vbase[g]:vbase[g + 1]]
        assert orc.check(off, adj, roots[g], np.where(p >= 0, p - vbase[g], -1)), g
    lv = ctx.levels()
    assert lv[0]["frontier"] == len(set(int(vbase[g]) + r for g, r in enumerate(roots)))
    # errors: an edge between graphs, bad boundaries, a root outside its graph
    with pytest.raises(bd.BeedrillError) as e:
        ctx.multi_graph(np.array([0, 3, vbase[-1]]), [0, 0])
    assert e.value.code == -1
    with pytest.raises(bd.BeedrillError) as e:
        ctx.multi_graph(vbase[:-1], roots[:-1])
    assert e.value.code == -1
    bad = list(roots)
    bad[2] = int(vbase[3] - vbase[2])
    with pytest.raises(bd.BeedrillError) as e:
        ctx.multi_graph(vbase, bad)
    assert e.value.code == -6


def test_learned_policy_any_network_same_depth(ctx, orc, s16):
    """Whatever the network predicts, the learned policy only changes
    directions: depths stay bit-exact and parents valid.  Random networks at
    several override margins, and the trained network."""
    ctx.generate_rmat(16, 16)
    off, adj, m = s16
    roots = [int(r) for r in ctx.pick_roots(3)]
    rng = np.random.default_rng(11)
    t0 = ctx.get_tuning()
    dirs_seen = set()
    try:
        for trial in range(3):
            ctx.mlp_set(rng.normal(0, 1.0, 32 * 8 + 32 + 32 * 32 + 32 + 2 * 32 + 2), 32, 32)
            for margin in (1, 500, 4000):
                ctx.set_tuning(learned_margin_milli=margin)
                for root in roots:
                    run_and_compare(ctx, orc, off, adj, m, root, policy="learned", levels=False)
                    dirs_seen.add(tuple(r["dir"] for r in ctx.levels()))
    finally:
        ctx.set_tuning(**t0)
    assert len(dirs_seen) > 1  # the networks did change directions


def test_learned_decisions_follow_the_network(ctx, orc, s16):
    """The network is consulted where Beamer's rules switch direction and
    decides whether to switch or stay; elsewhere Beamer's choice stands.
    (1) The trained network with the smallest margin: every step's direction
    is either the previous step's or the network's preference, evaluated in
    the kernel (bf16 MFMA) from features it computes itself, as the same
    network in fp32 on the host prefers it from the step records' features
    (steps where the two predictions are within bf16 rounding are skipped).
    (2) A network that always prefers bottom-up: Beamer's directions up to
    its first bottom-up step, bottom-up from there on; with learned_near 4 up
    to the first top-down step whose frontier edges * alpha * 4 exceed the
    unvisited edges, if that comes earlier.  (3) One that always prefers
    top-down: top-down throughout."""
    import os
    ctx.generate_rmat(16, 16)
    off, adj, m = s16
    nv = len(off) - 1
    path = os.path.join(os.path.dirname(os.path.dirname(__file__)), "beedrill_amd", "models",
                        "direction_mlp_32x32.f32")
    params = np.fromfile(path, np.float32)
    ctx.mlp_set(params, 32, 32)
    t0 = ctx.get_tuning()
    checked = 0
    roots = [int(r) for r in ctx.pick_roots(6)]
    try:
        ctx.set_tuning(learned_margin_milli=1)
        for root in roots:
            ctx.bfs(root, 15, 18, "learned")
            lv = ctx.levels()
            for i, rec in enumerate(lv):
                if i == 0:
                    continue
                prev = lv[i - 1]["dir"]
                x = bd.features(rec, nv, m, prev)
                y = orc.mlp_forward(params, np.asarray([x], np.float32))[0]
                if abs(y[1] - y[0]) < 0.1:
                    continue
                pref = bd.DIR_BU if y[1] < y[0] else bd.DIR_TD
                assert rec["dir"] in (prev, pref), (root, i, y)
                checked += 1
        n_out = 2
        nparams = params.size
        for want in (bd.DIR_BU, bd.DIR_TD):
            fixed = np.zeros(nparams, np.float32)
            # output biases: log2 cost (top-down, bottom-up)
            fixed[nparams - n_out:] = [10.0, -10.0] if want == bd.DIR_BU else [-10.0, 10.0]
            ctx.mlp_set(fixed, 32, 32)
            for near in (1, 4):
                ctx.set_tuning(learned_near=near)
                for root in roots:
                    ctx.bfs(root, 15, 18, "beamer")
                    blv = ctx.levels()
                    beamer = [r["dir"] for r in blv]
                    st = ctx.bfs(root, 15, 18, "learned")
                    dirs = [r["dir"] for r in ctx.levels()]
                    assert len(dirs) == len(beamer)
                    if want == bd.DIR_BU:
                        k = beamer.index(bd.DIR_BU) if bd.DIR_BU in beamer else len(beamer)
                        for i in range(1, k):
                            # Beamer's counters before step i's charge
                            scout = blv[i]["frontier_deg"]
                            if scout * 15 * near > blv[i]["edges_to_check"] + scout:
                                k = i
                                break
                        assert dirs == beamer[:k] + [bd.DIR_BU] * (len(beamer) - k), \
                            (root, near, dirs)
                    else:
                        assert dirs == [bd.DIR_TD] * len(beamer), (root, dirs)
                    assert ctx.check(root)
                    assert st["reached"] == 1 + sum(r["next"] for r in ctx.levels())
    finally:
        ctx.set_tuning(**t0)
        ctx.mlp_set(params, 32, 32)
    assert checked >= 10


def test_learned_policy_is_beamer_on_graphs_without_hubs(ctx, orc):
    """The network only overrides Beamer on graphs whose largest degree is at
    least learned_min_skew times the mean (it was trained on graphs with
    hubs).  A network that always prefers bottom-up changes an RMAT search
    (bottom-up from Beamer's first bottom-up step on) but leaves a lattice
    search exactly Beamer's."""
    params = np.zeros(32 * 8 + 32 + 32 * 32 + 32 + 2 * 32 + 2, np.float32)
    params[-2:] = [10.0, -10.0]  # predicted log2 cost: top-down 2^10, bottom-up 2^-10
    ctx.mlp_set(params, 32, 32)
    t0 = ctx.get_tuning()
    try:
        # switch points only, so the expected directions follow from Beamer's
        ctx.set_tuning(learned_margin_milli=1, learned_near=1)
        ctx.generate_lattice(30, 40, seed=3)
        off, adj = ctx.csr()
        for root in (0, 611):
            ctx.bfs(root, 15, 18, "beamer")
            beamer = [(r["dir"], r["next"]) for r in ctx.levels()]
            ctx.bfs(root, 15, 18, "learned")
            assert [(r["dir"], r["next"]) for r in ctx.levels()] == beamer
        ctx.generate_rmat(12, 16, seed=5)
        off, adj = ctx.csr()
        m = ctx.info()[2]
        root = int(ctx.pick_roots(1)[0])
        ctx.bfs(root, 15, 18, "beamer")
        beamer = [r["dir"] for r in ctx.levels()]
        ctx.bfs(root, 15, 18, "learned")
        dirs = [r["dir"] for r in ctx.levels()]
        k = beamer.index(bd.DIR_BU)
        assert dirs == beamer[:k] + [bd.DIR_BU] * (len(beamer) - k) and dirs != beamer
        np.testing.assert_array_equal(ctx.fetch()[0], orc.bfs_depth(off, adj, root))
    finally:
        ctx.set_tuning(**t0)
