"""Host side of the many-graph batch: the disjoint union of CSR graphs."""
import numpy as np
import scipy.sparse as sp

import beedrill_amd as bd


def test_union_csr_is_block_diagonal(orc):
    graphs, mats = [], []
    for g, (scale, ef) in enumerate([(5, 8), (7, 16), (4, 4), (6, 32)]):
        nv, s, d = orc.make_rmat(scale, ef, seed=g)
        off, adj = orc.build_csr(nv, s, d)
        graphs.append((off, adj, len(s)))
        mats.append(sp.csr_matrix((np.ones(len(adj)), adj, off), shape=(nv, nv)))
    off, cols, m, vbase = bd.union_csr(graphs)
    assert m == sum(x[2] for x in graphs)
    assert vbase.tolist() == [0] + np.cumsum([len(o) - 1 for o, _, _ in graphs]).tolist()
    got = sp.csr_matrix((np.ones(len(cols)), cols, off), shape=(vbase[-1], vbase[-1]))
    want = sp.block_diag(mats, format="csr")
    assert (got != want).nnz == 0
    # each list keeps its order
    for g, (o, c, _) in enumerate(graphs):
        v = len(o) // 2
        np.testing.assert_array_equal(cols[off[vbase[g] + v]:off[vbase[g] + v + 1]],
                                      c[o[v]:o[v + 1]] + vbase[g])
