"""The components oracle (components.cc) pinned against scipy's connected
components and closed forms; planted wrong labels must fail its check."""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components

from conftest import complete_graph, path_graph, star_graph, two_components


def min_id_labels(nv, off, adj):
    """scipy's partition, each part named by its smallest vertex id."""
    m = csr_matrix((np.ones(len(adj)), adj, off), shape=(nv, nv))
    _, part = connected_components(m, directed=False)
    first = np.full(part.max() + 1 if nv else 0, nv, np.int64)
    np.minimum.at(first, part, np.arange(nv))
    return first[part]


@pytest.mark.parametrize("g,labels", [
    (path_graph(50), [0] * 50),
    (star_graph(20), [0] * 20),
    (complete_graph(7), [0] * 7),
    (two_components(), [0, 0, 0, 3, 3]),
])
def test_closed_forms(orc, g, labels):
    nv, s, d = g
    off, adj = orc.build_csr(nv, s, d)
    lab, n, it = orc.components(off, adj)
    assert lab.tolist() == labels
    assert n == len(set(labels))
    assert it >= 1
    assert orc.components_check(off, adj, lab)


def test_isolated_vertices_are_their_own_components(orc):
    nv = 10
    s = np.array([2, 3, 7], np.int64)
    d = np.array([3, 4, 7], np.int64)  # a self-loop on 7
    off, adj = orc.build_csr(nv, s, d)
    lab, n, _ = orc.components(off, adj)
    assert lab.tolist() == [0, 1, 2, 2, 2, 5, 6, 7, 8, 9]
    assert n == 8


@pytest.mark.parametrize("kind,args", [("er", (3000, 2200, 5)), ("er", (2000, 6000, 9)),
                                       ("rmat", (11, 4, 0.57, 0.19, 0.19, 0.05, 3)),
                                       ("rmat", (12, 2, 0.45, 0.25, 0.25, 0.05, 8))])
def test_matches_scipy(orc, kind, args):
    if kind == "er":
        nv, s, d = orc.make_er(*args[:2], seed=args[2])
    else:
        nv, s, d = orc.make_rmat(*args[:6], seed=args[6])
    off, adj = orc.build_csr(nv, s, d)
    lab, n, _ = orc.components(off, adj)
    np.testing.assert_array_equal(lab, min_id_labels(nv, off, adj))
    assert n == len(np.unique(lab))
    assert n > 1  # sparse enough to have several components
    assert orc.components_check(off, adj, lab)


def test_check_catches_wrong_labels(orc):
    nv, s, d = orc.make_er(500, 300, seed=2)
    off, adj = orc.build_csr(nv, s, d)
    lab, _, _ = orc.components(off, adj)
    assert orc.components_check(off, adj, lab)
    # two components merged under one label
    roots = np.unique(lab)
    merged = lab.copy()
    merged[lab == roots[1]] = roots[0]
    assert not orc.components_check(off, adj, merged)
    # one vertex of a component of size > 1 given another label
    sizes = np.bincount(lab, minlength=nv)
    big = int(np.argmax(sizes))
    v = int(np.flatnonzero(lab == big)[-1])
    split = lab.copy()
    split[v] = v if v != big else big + 1
    assert not orc.components_check(off, adj, split)
