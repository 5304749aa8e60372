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
    # iid lognormal coefficients (sigma 2): weak connections at the rank
    # boundaries, where A_ghost's rows smooth differently from own rows
    "3d-contrast-boxes": lambda: (_contrast(12, 7), "boxes", 8, dict(BASE, radius1=2, strong1=0.25)),
}


def _contrast(N, seed):
    kappa = np.exp(2.0 * np.random.default_rng(seed).standard_normal((N, N, N)))
    return P.diffusion3d(N, kappa=kappa)


def _shuffled(m):
    n, i, j, v = m
    q = np.random.default_rng(3).permutation(len(v))
    return n, i[q], j[q], v[q]


def _case(name):
    (n, i, j, v), kind, nr, ctl = CASES[name]()
    if kind == "stripes":
        rank = ((np.arange(n) // 7) % nr + 1).astype(np.int32)
    elif kind == "boxes":
        N = round(n ** (1 / 3))
        st, nd = P.block_partition((N,) * 3, N // 2)
        rank = np.zeros(n, np.int32)
        for r in range(len(st) - 1):
            rank[nd[st[r] - 1:st[r + 1] - 1] - 1] = r + 1
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
        # the coarse space of the several-process run: entries and bits
        assert _same_entries(s.restriction(), o.R(), True)
        assert _same_entries(s.coarse_matrix(), o.AC(), False)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


def _same_entries(csr, coo, sort_coo):
    """The library's 0-based CSR against the oracle's 1-based COO (R: sorted
    first, the oracle keeps it arranged by columns)."""
    rp, c, v = csr
    oi, oj, ov = coo
    if sort_coo:
        order = np.lexsort((oj, oi))
        oi, oj, ov = oi[order], oj[order], ov[order]
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return np.array_equal(rows, oi - 1) and np.array_equal(c, oj - 1) and _bits_equal(v, ov)


def _owner_view_R(n, i, j, v, rank, num, ctl):
    """R = T S^T of DOUG's several-process run, from its definition and not
    from either restatement (smoothers 1): R(a, j) sums S(j, k) over the
    nodes k of aggregate a, with row j of S as a's owner sees it.  A row of
    a node the owner holds (its own rows, and with overlap >= 2 the first
    outside layer) is smoothed with the filtered diagonal a_jj + the weak
    a_jk; a row of the outside layer the owner receives (overlap 1) is
    A_ghost's: its diagonal entry alone.  Strength: |a_ij| > alpha max|a_ik|
    on the scaled matrix (symmetric scaling when the matrix is flagged
    symmetric, else by rows, which leaves the test unscaled)."""
    om = 0.33333333333333
    A = sp.coo_matrix((v, (i - 1, j - 1)), shape=(n, n)).tocsr()
    A.sum_duplicates()
    C = A.tocoo()
    d = A.diagonal()
    if ctl.get("symmstruct") and ctl.get("symmnumeric"):
        sc = C.data / np.sqrt(abs(d[C.row])) / np.sqrt(abs(d[C.col]))
    else:
        sc = C.data
    off = C.row != C.col
    mx = np.zeros(n)
    np.maximum.at(mx, C.row[off], abs(sc[off]))
    strong = np.zeros(len(sc), bool)
    strong[off] = abs(sc[off]) > ctl["strong1"] * mx[C.row[off]]
    weak = off & ~strong
    dfilt = d.copy()
    np.add.at(dfilt, C.row[weak], C.data[weak])
    L = max(ctl.get("overlap", 1), ctl.get("smoothers", 1), 1)
    owner = np.zeros(num.max() + 1, int)
    owner[num[num > 0]] = rank[num > 0]
    keep = num[C.col] > 0
    a = num[C.col][keep] - 1
    jr, kk, val, st = C.row[keep], C.col[keep], C.data[keep], strong[keep]
    ghost = (L == 1) & (rank[jr] != owner[a + 1])
    dj = np.where(ghost, d[jr], dfilt[jr])
    s_ = np.where(jr == kk, np.where(ghost, 1 - om / dj * val, 1 - om),
                  np.where(st, -om / dj * val, 0.0))
    return sp.csr_matrix((s_, (a, jr)), shape=(num.max(), n))


@pytest.mark.parametrize("name", sorted(CASES))
def test_coarse_space_of_several_processes(name):
    # the oracle's restatement of the several-process coarse space against
    # what its definition fixes: A_c = R A R^T over the ranks' parts, and
    # (smoothers 1) R from the owner-view formula
    n, i, j, v, rank, nr, ctl = _case(name)
    o = O.Oracle.ranks(n, i, j, v, O.params(**ctl), rank)
    _check_coarse_space(o, n, i, j, v, rank, ctl)


def _check_coarse_space(o, n, i, j, v, rank, ctl):
    ri, rj, rv = o.R()
    nc = int(ri.max())
    R = sp.csr_matrix((rv, (ri - 1, rj - 1)), shape=(nc, n))
    ai, aj, av = o.A()
    A = sp.csr_matrix((av, (ai - 1, aj - 1)), shape=(n, n))
    ci, cj, cv = o.AC()
    AC = sp.csr_matrix((cv, (ci - 1, cj - 1)), shape=(nc, nc))
    G = (R @ A @ R.T).tocsr()
    assert abs(AC - G).max() <= 1e-13 * abs(G).max()
    if ctl.get("smoothers", 1) == 1:
        num, _, _ = o.fine_aggregates()
        Ri = _owner_view_R(n, i, j, v, rank, num, ctl)
        assert abs(R - Ri).max() <= 1e-14 * abs(Ri).max()


def test_one_rank_coarse_space_is_the_one_process_one():
    n, i, j, v = P.diffusion3d(12, 1.0, 4)
    ctl = dict(BASE, radius1=2)
    o1 = O.Oracle.ranks(n, i, j, v, O.params(**ctl), np.ones(n, np.int32))
    num, _, _ = o1.fine_aggregates()
    # the same aggregates given to one process: the same R and A_c bits
    with D.Solver(-1) as s:
        s.set_params(**ctl)
        s.set_matrix(n, i, j, v, 1)
        s.set_aggregates(num)
        s.set_partition(*P.block_partition((12,) * 3, 6))
        s.setup_host()
        assert _same_entries(s.restriction(), o1.R(), True)
        cp, cc, cv = s.coarse_matrix()
    ci, cj, cv1 = o1.AC()  # sorted by row and column
    rows = np.repeat(np.arange(len(cp) - 1), np.diff(cp))
    order = np.lexsort((cc, rows))
    np.testing.assert_array_equal(rows[order], ci - 1)
    np.testing.assert_array_equal(cc[order], cj - 1)
    assert _bits_equal(cv[order], cv1)


def test_planted_coarse_space_mistakes_are_caught(tmp_path):
    import ctypes as C
    import subprocess
    import os
    src = os.path.join(os.path.dirname(O.__file__), "doug_oracle.c")
    n, i, j, v, rank, nr, ctl = _case("3d-contrast-boxes")
    for k in (9, 10):
        out = str(tmp_path / f"liboracle_m{k}.so")
        subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off",
                        f"-DOR_MUTANT={k}", "-shared", "-o", out, src, "-lm"], check=True)
        L = O.bind(C.CDLL(out))
        # ghost rows with the filtered diagonal (9) fail the owner-view
        # formula; a coarse matrix missing the other ranks' parts (10) fails
        # R A R^T or is not even positive definite
        with pytest.raises((AssertionError, RuntimeError)):
            om = O.Oracle.ranks(n, i, j, v, O.params(**ctl), rank, L=L)
            _check_coarse_space(om, n, i, j, v, rank, ctl)


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
