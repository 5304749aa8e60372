"""Project graph partitioner (doug_partition_graph) and the subdomain source
it feeds (doug_set_num_subdomains).  METIS is absent, so partitions are not
compared with the reference's: parity unpinned.  These tests pin what the
partitioner promises (cover, balance, determinism, a cut near the ideal on
grids) and what the subdomain rule of Partitioning_metis.F90:34-86 fixes
(whole aggregates, nodes in node order)."""
import numpy as np
import pytest

import doug_amd as D
from doug_amd import problems as P


def grid_graph(nx, ny):
    idx = np.arange(nx * ny).reshape(nx, ny)
    pairs = [(idx[:-1, :].ravel(), idx[1:, :].ravel()), (idx[:, :-1].ravel(), idx[:, 1:].ravel())]
    a = np.concatenate([p[0] for p in pairs] + [p[1] for p in pairs])
    b = np.concatenate([p[1] for p in pairs] + [p[0] for p in pairs])
    order = np.lexsort((b, a))
    a, b = a[order], b[order]
    xadj = np.concatenate([[0], np.cumsum(np.bincount(a, minlength=nx * ny))])
    return xadj.astype(np.int32), b.astype(np.int32)


def edge_cut(xadj, adj, part):
    rows = np.repeat(np.arange(len(xadj) - 1), np.diff(xadj))
    return int(np.sum(part[rows] != part[adj])) // 2


@pytest.mark.parametrize("k", [1, 2, 3, 7, 16])
def test_grid_partition_cover_balance_cut(k):
    nx = ny = 64
    xadj, adj = grid_graph(nx, ny)
    part = D.partition_graph(xadj, adj, k)
    assert part.min() == 0 and part.max() == k - 1
    sizes = np.bincount(part, minlength=k)
    assert sizes.min() > 0
    assert sizes.max() <= (nx * ny) / k * 1.05 + 1
    if k == 16:  # 4 x 4 boxes of 16 x 16 cut 2 * 3 * 64 edges
        assert edge_cut(xadj, adj, part) <= 1.6 * 384
    np.testing.assert_array_equal(part, D.partition_graph(xadj, adj, k))


def test_vertex_weights_balance():
    xadj, adj = grid_graph(40, 40)
    w = np.ones(1600, np.int32)
    w[:800] = 3  # the first half weighs three times as much
    part = D.partition_graph(xadj, adj, 4, vwgt=w)
    loads = np.bincount(part, weights=w, minlength=4)
    assert loads.max() <= w.sum() / 4 * 1.05 + 3


def test_bad_arguments_are_refused():
    xadj, adj = grid_graph(4, 4)
    bad = adj.copy()
    bad[0] = 99
    with pytest.raises(D.DougError):
        D.partition_graph(xadj, bad, 2)


@pytest.mark.parametrize("nsub", [3, 8])
def test_subdomains_are_whole_aggregates_in_node_order(nsub):
    n, i, j, v = P.diffusion3d(16, 1.0, 5)
    with D.Solver(-1) as s:
        s.set_params(radius1=2, strong1=0.08, smoothers=1, overlap=0, symmstruct=1, symmnumeric=1)
        s.set_matrix(n, i, j, v, 1)
        s.set_num_subdomains(nsub)
        info = s.setup_host()
        starts, inds = s.subdomains()
        num, _, _ = s.fine_aggregates()
    assert 1 < info["nsub"] <= nsub
    inds = inds - 1
    assert np.array_equal(np.sort(inds), np.arange(n))  # overlap 0: a partition
    owner = np.repeat(np.arange(len(starts) - 1), np.diff(starts))
    node_part = np.empty(n, np.int64)
    node_part[inds] = owner
    for a in np.unique(num[num > 0]):
        assert len(np.unique(node_part[num == a])) == 1
    for q in range(len(starts) - 1):
        seg = inds[starts[q] - 1:starts[q + 1] - 1]
        assert np.all(np.diff(seg) > 0)
