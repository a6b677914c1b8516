"""GPU parity for connected components (components.cc): the device labels,
called through the C ABI, equal the oracle's (each component labelled by its
smallest vertex id, whatever the pass order), with the same component count,
and the check passes."""
import os
import subprocess

import numpy as np
import pytest

from conftest import complete_graph, path_graph, star_graph, two_components

pytestmark = pytest.mark.gpu


def compare(ctx, orc, off, adj):
    st = ctx.components()
    lab = ctx.component_labels()
    ref, n, _ = orc.components(off, adj)
    np.testing.assert_array_equal(lab, ref)
    assert st["num_components"] == n
    assert st["num_iters"] >= 1 and st["ms"] > 0
    assert 0 < st["pass_ms"] <= st["ms"]
    # the first pass takes neighbour ids, later pulls read every label and
    # pushes the changed vertices' labels only
    nnz, nv = int(off[-1]), len(off) - 1
    assert 0 <= st["labels_read"] <= (st["num_iters"] - 1) * nnz
    assert st["edge_bytes"] >= 4 * nnz + 12 * nv + 8 * st["labels_read"]
    assert st["num_iters"] <= st["pass_launches"] <= 9 * st["num_iters"]
    assert ctx.components_check()
    return st


@pytest.mark.parametrize("g", [path_graph(300), star_graph(100), complete_graph(30),
                               two_components()])
def test_small_graphs(ctx, orc, g):
    nv, s, d = g
    ctx.from_edges(nv, s, d)
    off, adj = orc.build_csr(nv, s, d)
    compare(ctx, orc, off, adj)


def test_isolated_vertices_and_self_loops(ctx, orc):
    nv = 10
    s = np.array([2, 3, 7], np.int64)
    d = np.array([3, 4, 7], np.int64)
    ctx.from_edges(nv, s, d)
    off, adj = orc.build_csr(nv, s, d)
    compare(ctx, orc, off, adj)
    assert ctx.component_labels().tolist() == [0, 1, 2, 2, 2, 5, 6, 7, 8, 9]


def test_long_path_under_relabelling(ctx, orc):
    # one component whose smallest id sits in the middle of a 20000-vertex
    # path: labels travel both ways through the id space
    n = 20000
    perm = np.random.default_rng(3).permutation(n).astype(np.int64)
    ctx.from_edges(n, perm[:-1], perm[1:])
    off, adj = orc.build_csr(n, perm[:-1], perm[1:])
    st = compare(ctx, orc, off, adj)
    assert st["num_components"] == 1


@pytest.mark.parametrize("scale,ef,seed", [(16, 16, 0), (14, 2, 5), (12, 1, 9)])
def test_rmat(ctx, orc, scale, ef, seed):
    ctx.generate_rmat(scale, ef, seed=seed)
    off, adj = ctx.csr()
    compare(ctx, orc, off, adj)


def test_long_lists_and_slices(ctx, orc):
    # lists of 255 and 256 edges (either side of the long-list cut), hubs of
    # several 2048-edge slices, long and short lists in one 64-vertex group,
    # and a group whose vertices all have long lists
    rng = np.random.default_rng(11)
    nv = 40000
    src, dst = [], []
    for hub, deg in [(5, 255), (6, 256), (7, 257), (70, 5000), (71, 2048), (72, 2049),
                     (20000, 9000)] + [(128 + i, 300 + i) for i in range(64)]:
        nb = rng.choice(np.arange(30000, nv), deg, replace=False)
        src.append(np.full(deg, hub)); dst.append(nb)
    a = rng.integers(1000, 30000, 6000)
    src.append(a); dst.append(a + 1)  # short lists elsewhere
    s = np.concatenate(src).astype(np.int64)
    d = np.concatenate(dst).astype(np.int64)
    ctx.from_edges(nv, s, d)
    off, adj = orc.build_csr(nv, s, d)
    compare(ctx, orc, off, adj)
    compare(ctx, orc, off, adj)  # second run reuses the slices


def test_er_and_lattice(ctx, orc):
    ctx.generate_er(20000, 12000, seed=4)
    off, adj = ctx.csr()
    st = compare(ctx, orc, off, adj)
    assert st["num_components"] > 100
    ctx.generate_lattice(60, 70, seed=1)
    off, adj = ctx.csr()
    compare(ctx, orc, off, adj)


def test_cli_components_check():
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "beedrill_amd", "bin", "components")
    r = subprocess.run([exe, "--generate", "graph500-scale12", "--num_trials", "2",
                        "--check_results"], capture_output=True, text=True, timeout=120)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert out.count("Found ") == 2 and "PASS" in out and "FAIL" not in out, out
