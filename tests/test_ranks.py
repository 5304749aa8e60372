"""Virtual ranks: DOUG's decomposition over several MPI processes
(SpMtx_DistributeAssembled / SpMtx_localize / SpMtx_build_ghost /
SpMtx_Build_lggl, per-rank aggregation of Partitioning_aggr.F90:81-92 and the
per-rank subdomains of aggr.F90:200-205) restated on one host.

The library formulates the rank's node labels as graph distances; the
oracle follows the reference's own loops (front arrays, per-neighbour
breadth-first passes with layer offsets, the four matrix blocks).  Both are
checked against each other bit for bit, and each against what the
definitions fix independently: labels and local numbering from algebraic
reachability (boolean matrix powers), subdomains as graph balls, and the
one-rank case against the one-process aggregation."""
import numpy as np
import pytest
import scipy.sparse as sp

import doug_amd as D
import oracle as O
from doug_amd import problems as P

BASE = dict(strong1=0.08, smoothers=1, overlap=1, symmstruct=1, symmnumeric=1)


def _host(n, i, j, v, nranks, rank=None, nsub=0, **ctl):
    s = D.Solver(-1)
    s.set_params(**ctl)
    s.set_matrix(n, i, j, v, 1)
    if nsub:
        s.set_num_subdomains(nsub)
    s.set_ranks(nranks, rank)
    s.setup_host()
    return s


def _ranks_of(s, nranks, n):
    rank = np.zeros(n, np.int32)
    for r in range(nranks):
        ni, lg = s.rank_map(r)
        rank[lg[:ni] - 1] = r + 1
    return rank


def _dist(G, sources, maxd):
    """Hop distance from a node set by boolean matrix powers (inf beyond maxd)."""
    n = G.shape[0]
    d = np.full(n, np.inf)
    cur = np.zeros(n, bool)
    cur[sources] = True
    d[cur] = 0
    for level in range(1, maxd + 1):
        reach = (G @ cur.astype(np.int8)) > 0
        new = reach & np.isinf(d)
        d[new] = level
        cur = cur | reach
    return d


def _graph(n, i, j):
    G = sp.csr_matrix((np.ones(len(i), np.int8), (i - 1, j - 1)), shape=(n, n))
    return ((G + G.T) > 0).astype(np.int8)


CASES = {
    "2d-quadrants": lambda: (P.laplace2d(16), None, 4, dict(BASE, radius1=-2)),
    "2d-field-overlap2": lambda: (P.diffusion2d(40, 1.0, 3), None, 3, dict(BASE, radius1=-3, overlap=2)),
    "3d-stripes-smoothers2": lambda: (P.diffusion3d(10, 1.0, 2), "stripes", 4,
                                      dict(BASE, radius1=-1, strong1=0.1, smoothers=2, overlap=-1,
                                           symmstruct=0, symmnumeric=0)),
    "3d-scattered": lambda: (P.diffusion3d(10, 1.0, 2), "random", 6, dict(BASE, radius1=2)),
    # entries in random order: the rows reach the ranks sorted by column
    # (SpMtx_distributeWithOverlap arranges with sort)
    "2d-shuffled-input": lambda: (_shuffled(P.diffusion2d(24, 1.0, 9)), "stripes", 3,
                                  dict(BASE, radius1=-2)),
}


def _shuffled(m):
    n, i, j, v = m
    q = np.random.default_rng(3).permutation(len(v))
    return n, i[q], j[q], v[q]


def _case(name):
    (n, i, j, v), kind, nr, ctl = CASES[name]()
    if kind == "stripes":
        rank = ((np.arange(n) // 7) % nr + 1).astype(np.int32)
    elif kind == "random":
        rank = np.random.default_rng(5).integers(1, nr + 1, n).astype(np.int32)
    else:
        with _host(n, i, j, v, nr, None, **ctl) as s:
            rank = _ranks_of(s, nr, n)
    return n, i, j, v, rank, nr, ctl


@pytest.mark.parametrize("name", sorted(CASES))
def test_library_matches_oracle_bit_for_bit(name):
    n, i, j, v, rank, nr, ctl = _case(name)
    o = O.Oracle.ranks(n, i, j, v, O.params(**ctl), rank)
    with _host(n, i, j, v, nr, rank, **ctl) as s:
        for r in range(nr):
            ni, lg = s.rank_map(r)
            assert ni == o.rank_maps[r][0]
            np.testing.assert_array_equal(lg, o.rank_maps[r][1])
        for a, b in zip(s.fine_aggregates(), o.fine_aggregates()):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(s.subdomains(), o.subdomains()):
            np.testing.assert_array_equal(a, b)
        assert s.info()["nsub"] == nr


@pytest.mark.parametrize("name", sorted(CASES))
def test_local_numbering_from_distances(name):
    # Build_lggl's order: [sent | inner overlap | inner | outer overlap |
    # received], each ascending; sent = exactly L steps from some other
    # rank's nodes, inner overlap = within L steps, received = outside nodes
    # exactly L steps from the rank (L = max(overlap, smoothers))
    n, i, j, v, rank, nr, ctl = _case(name)
    L = max(ctl["overlap"], ctl["smoothers"])
    G = _graph(n, i, j)
    o = O.Oracle.ranks(n, i, j, v, O.params(**ctl), rank)
    for r in range(1, nr + 1):
        own = np.flatnonzero(rank == r)
        dk = [_dist(G, np.flatnonzero(rank == k), L) for k in range(1, nr + 1) if k != r]
        d_own = _dist(G, own, L)
        sent = np.zeros(n, bool)
        near = np.zeros(n, bool)
        for d in dk:
            sent |= d == L
            near |= d <= L
        mine = rank == r
        groups = [mine & sent, mine & near & ~sent, mine & ~near,
                  ~mine & (d_own >= 1) & (d_own < L), ~mine & (d_own == L)]
        want = np.concatenate([np.flatnonzero(g) for g in groups]) + 1
        ninner, lg = o.rank_maps[r - 1]
        assert ninner == len(own)
        np.testing.assert_array_equal(lg, want)


@pytest.mark.parametrize("name", sorted(CASES))
def test_subdomains_are_balls_around_the_ranks(name):
    # one subdomain per rank: its nodes plus every node within `overlap`
    # steps (the local matrix holds all rows nearer than L), rank by rank
    n, i, j, v, rank, nr, ctl = _case(name)
    ol = ctl["overlap"] if ctl["overlap"] >= 0 else min(ctl["smoothers"], abs(ctl["radius1"]))
    G = _graph(n, i, j)
    o = O.Oracle.ranks(n, i, j, v, O.params(**ctl), rank)
    st, inds = o.subdomains()
    assert len(st) == nr + 1
    for r in range(1, nr + 1):
        sub = inds[st[r - 1] - 1:st[r] - 1]
        assert len(np.unique(sub)) == len(sub)
        ball = np.flatnonzero(_dist(G, np.flatnonzero(rank == r), ol) <= ol) + 1
        np.testing.assert_array_equal(np.sort(sub), ball)
        # the rank's own nodes come first, in its local order
        ninner, lg = o.rank_maps[r - 1]
        np.testing.assert_array_equal(sub[:ninner], lg[:ninner])


@pytest.mark.parametrize("name", sorted(CASES))
def test_aggregates_stay_inside_their_rank(name):
    n, i, j, v, rank, nr, ctl = _case(name)
    o = O.Oracle.ranks(n, i, j, v, O.params(**ctl), rank)
    num, st, nodes = o.fine_aggregates()
    assert (num > 0).all()
    first = np.zeros(nr + 1, int)
    for a in range(len(st) - 1):
        members = nodes[st[a] - 1:st[a + 1] - 1]
        owners = np.unique(rank[members - 1])
        assert len(owners) == 1
        # numbered rank by rank (CoarseAllgathers.F90:531-540)
        assert owners[0] >= first.max()
        first[owners[0]] = owners[0]


def test_one_rank_is_the_one_process_aggregation():
    # with one rank every node is inner and unsent: the local matrix is the
    # assembled one with sorted rows, and getLocal's copy carries no
    # symmetry flags, so the aggregates are those of one process with row
    # scaling on the sorted input
    n, i, j, v = P.diffusion2d(30, 1.0, 7)
    order = np.lexsort((j, i))
    i, j, v = i[order], j[order], v[order]
    ctl = dict(BASE, radius1=2, symmstruct=0, symmnumeric=0)
    one = np.ones(n, np.int32)
    o = O.Oracle.ranks(n, i, j, v, O.params(**ctl), one)
    # (one injected part: the one-process coarse aggregation is not compared)
    ref = O.Oracle(n, i, j, v, O.params(**ctl),
                   partition=(np.array([1, n + 1]), np.arange(1, n + 1)))
    np.testing.assert_array_equal(o.rank_maps[0][1], np.arange(1, n + 1))
    for a, b in zip(o.fine_aggregates(), ref.fine_aggregates()):
        np.testing.assert_array_equal(a, b)
    with _host(n, i, j, v, 1, one, **ctl) as s:
        for a, b in zip(s.fine_aggregates(), ref.fine_aggregates()):
            np.testing.assert_array_equal(a, b)


def test_planted_numbering_mistake_is_caught():
    # the distance check above discriminates: swapping the inner-overlap and
    # sent groups (a plausible misreading of Build_lggl's marks) fails it
    n, i, j, v, rank, nr, ctl = _case("2d-field-overlap2")
    L = 2
    G = _graph(n, i, j)
    o = O.Oracle.ranks(n, i, j, v, O.params(**ctl), rank)
    r = 1
    dk = [_dist(G, np.flatnonzero(rank == k), L) for k in range(1, nr + 1) if k != r]
    sent = np.zeros(n, bool)
    near = np.zeros(n, bool)
    for d in dk:
        sent |= d == L
        near |= d <= L
    mine = rank == r
    wrong = np.concatenate([np.flatnonzero(mine & near & ~sent), np.flatnonzero(mine & sent)]) + 1
    assert not np.array_equal(o.rank_maps[0][1][:len(wrong)], wrong)


def test_several_subdomains_per_rank_cover_each_rank():
    n, i, j, v = P.diffusion3d(12, 1.0, 4)
    with _host(n, i, j, v, 4, None, nsub=3, **dict(BASE, radius1=-2)) as s:
        rank = _ranks_of(s, 4, n)
        st, inds = s.subdomains()
        assert len(st) - 1 == 12
        seen = np.zeros(n, int)
        for q in range(12):
            sub = inds[st[q] - 1:st[q + 1] - 1]
            r = q // 3 + 1
            core = sub[rank[sub - 1] == r]
            seen[core - 1] += 1
        assert (seen >= 1).all()


def test_rank_errors():
    n, i, j, v = P.laplace2d(8)
    with D.Solver(-1) as s:
        s.set_matrix(n, i, j, v, 1)
        with pytest.raises(RuntimeError, match="out of range"):
            s.set_ranks(2, np.full(n, 3, np.int32))
        with pytest.raises(RuntimeError, match="owns no unknown"):
            s.set_ranks(3, np.r_[np.ones(n - 1), 2].astype(np.int32))
        s.set_params(**dict(BASE, radius1=-2, overlap=0, smoothers=0))
        s.set_ranks(2)
        with pytest.raises(RuntimeError, match="max\\(overlap, smoothers\\)"):
            s.setup_host()
        s.set_params(**dict(BASE, radius1=-2))
        st = np.array([1, n + 1], np.int32)
        s.set_partition(st, np.arange(1, n + 1, dtype=np.int32))
        with pytest.raises(RuntimeError, match="clear the partition"):
            s.setup_host()


def test_repeated_setups_reuse_scratch_cleanly():
    # the rank workers' scratch is kept by the context between set-ups:
    # later set-ups (other options, other rank maps) equal fresh ones
    n, i, j, v, rank, nr, ctl = _case("2d-field-overlap2")
    runs = [(dict(ctl, overlap=1), rank, nr), (dict(ctl, overlap=2), rank, nr),
            (dict(ctl, overlap=1), ((np.arange(n) // 11) % 5 + 1).astype(np.int32), 5)]
    with D.Solver(-1) as s:
        s.set_matrix(n, i, j, v, 1)
        for c, rk, k in runs:
            s.set_params(**c)
            s.set_ranks(k, rk)
            s.setup_host()
            with _host(n, i, j, v, k, rk, **c) as fresh:
                for a, b in zip(s.fine_aggregates(), fresh.fine_aggregates()):
                    np.testing.assert_array_equal(a, b)
                for a, b in zip(s.subdomains(), fresh.subdomains()):
                    np.testing.assert_array_equal(a, b)


def test_methods_2_to_4_need_several_processes():
    # with one process methods 2-4 restrict a residual nothing has written
    # (pcg.F90:146-147,253-255): refused by the library and the oracle
    n, i, j, v = P.laplace2d(8)
    for m in (2, 3, 4):
        ctl = dict(BASE, radius1=2, method=m)
        with pytest.raises(RuntimeError, match="undefined with one process"):
            O.Oracle(n, i, j, v, O.params(**ctl))
        with D.Solver(-1) as s:
            s.set_params(**ctl)
            s.set_matrix(n, i, j, v, 1)
            with pytest.raises(RuntimeError, match="need virtual ranks"):
                s.setup_host()
            s.set_ranks(2)
            s.setup_host()  # accepted with several processes


def test_methods_2_to_4_with_ranks_follow_prec2level():
    # prec2Level_solve with several processes: 3 adds the coarse correction
    # like 1, 4 keeps the fine sum alone, 2 (overlap > 0) the coarse
    # correction alone; checked on the oracle against its method 1 and
    # one-level applies
    n, i, j, v = P.diffusion2d(24, 1.0, 2)
    rank = ((np.arange(n) // 96) + 1).astype(np.int32)
    r = np.random.default_rng(3).uniform(-1, 1, n)
    ctl = dict(BASE, radius1=2)
    z = {m: O.Oracle.ranks(n, i, j, v, O.params(**dict(ctl, method=m)), rank).prec(r) for m in (1, 2, 3, 4)}
    fine = O.Oracle.ranks(n, i, j, v, O.params(**dict(ctl, levels=1)), rank).prec(r)
    np.testing.assert_array_equal(z[3], z[1])
    np.testing.assert_array_equal(z[4], fine)
    coarse = z[1] - fine
    np.testing.assert_allclose(z[2], coarse, rtol=0, atol=1e-12 * np.abs(z[1]).max())
    assert np.abs(z[2]).max() > 0


def test_stale_rank_map_or_partition_is_refused():
    n, i, j, v = P.laplace2d(8)
    n2, i2, j2, v2 = P.laplace2d(10)
    with D.Solver(-1) as s:
        s.set_params(**dict(BASE, radius1=2))
        s.set_matrix(n, i, j, v, 1)
        s.set_ranks(2)
        s.set_matrix(n2, i2, j2, v2, 1)
        with pytest.raises(RuntimeError, match="rank map was set for another matrix"):
            s.setup_host()
        s.set_ranks(0)
        s.set_matrix(n, i, j, v, 1)
        s.set_partition(np.array([1, n + 1], np.int32), np.arange(1, n + 1, dtype=np.int32))
        s.set_matrix(n2, i2, j2, v2, 1)
        with pytest.raises(RuntimeError, match="partition was set for another matrix"):
            s.setup_host()


@pytest.mark.parametrize("case", ["node-per-rank", "node-per-rank-overlap2", "checkerboard",
                                  "disconnected-matrix"])
def test_edge_rank_maps_match_oracle(case):
    ctl = dict(BASE, radius1=2)
    n, i, j, v = P.laplace2d(6)
    g = np.arange(n)
    if case == "node-per-rank":
        rank = g + 1
    elif case == "node-per-rank-overlap2":
        rank, ctl = g + 1, dict(ctl, overlap=2)
    elif case == "checkerboard":  # every rank in many disconnected pieces
        rank = (g // 6 + g % 6) % 2 + 1
    else:  # two independent grids, ranks crossing both
        n1, i1, j1, v1 = P.laplace2d(5)
        n, i, j, v = 2 * n1, np.r_[i1, i1 + n1], np.r_[j1, j1 + n1], np.r_[v1, v1]
        rank = (np.arange(n) // 7) % 3 + 1
    rank = rank.astype(np.int32)
    nr = int(rank.max())
    o = O.Oracle.ranks(n, i, j, v, O.params(**ctl), rank)
    with _host(n, i, j, v, nr, rank, **ctl) as s:
        for a, b in zip(s.fine_aggregates(), o.fine_aggregates()):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(s.subdomains(), o.subdomains()):
            np.testing.assert_array_equal(a, b)
