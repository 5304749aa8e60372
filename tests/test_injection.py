"""Caller-given aggregates (doug_set_aggregates) and subdomain index sets
(doug_set_subdomains), the boundary's replacements for
Partitionings_aggr_InitFine and FinePreconditioner_Init: the coarse space
built from given aggregates against R = T S^T and A_c = R A R^T computed with
scipy from the formulas (CoarseMtx.F90:50-516), and the oracle's own
aggregates and index sets fed back through the boundary."""
import numpy as np
import pytest
import scipy.sparse as sp

import doug_amd as D
import oracle as O
from doug_amd import problems as P

CTL = dict(radius1=2, strong1=0.08, smoothers=1, overlap=1, symmstruct=1, symmnumeric=1)


def _csr(getter_rowptr, n_rows, n_cols):
    rp, c, v = getter_rowptr
    return sp.csr_matrix((v, c, rp), shape=(n_rows, n_cols))


def _library(n, i, j, v, num=None, subd=None, part=None, **ctl):
    s = D.Solver(-1)
    s.set_params(**ctl)
    s.set_matrix(n, i, j, v, 1)
    if part is not None:
        s.set_partition(*part)
    if num is not None:
        s.set_aggregates(num)
    if subd is not None:
        s.set_subdomains(*subd)
    s.setup_host()
    return s


def test_given_aggregates_give_the_formula_coarse_space():
    # 2x2 blocks of a 2-D field as aggregates (one corner node left out)
    m = 20
    n, i, j, v = P.diffusion2d(m, 1.0, 8)
    g = np.arange(n)
    num = ((g // m) // 2) * (m // 2) + (g % m) // 2 + 1
    num[0] = 0  # an unaggregated unknown: its own full aggregate, no row of R
    num = np.where(num > 0, num, 0).astype(np.int32)
    # renumber densely (aggregate 1 lost its corner but keeps three nodes)
    o = O.Oracle(n, i, j, v, O.params(**CTL), partition=P.block_partition((m, m), 10))
    Ai, Aj, Av = o.A()  # the values after DOUG's scale/unscale round trip
    A = sp.csr_matrix((Av, (Ai - 1, Aj - 1)), shape=(n, n))
    d = A.diagonal()
    As = sp.diags(1 / np.sqrt(d)) @ A @ sp.diags(1 / np.sqrt(d))
    Ac = As.tocoo()
    off = Ac.row != Ac.col
    mx = np.zeros(n)
    np.maximum.at(mx, Ac.row[off], np.abs(Ac.data[off]))
    strong = off & (np.abs(Ac.data) > 0.08 * mx[Ac.row])
    Acoo = A.tocoo()
    dfilt = d + np.bincount(Acoo.row[off & ~strong], Acoo.data[off & ~strong], minlength=n)
    w = 0.33333333333333
    sval = np.where(~off, 1 - w, np.where(strong, -w * Acoo.data / dfilt[Acoo.row], 0.0))
    S = sp.csr_matrix((sval, (Acoo.row, Acoo.col)), shape=(n, n))
    nagr = int(num.max())
    keep = num > 0
    T = sp.csr_matrix((np.ones(keep.sum()), (num[keep] - 1, np.flatnonzero(keep))), shape=(nagr, n))
    Rexp = (T @ S.T).toarray()
    with _library(n, i, j, v, num=num, part=P.block_partition((m, m), 10), **CTL) as s:
        R = _csr(s.restriction(), nagr, n).toarray()
        np.testing.assert_allclose(R, Rexp, rtol=1e-14, atol=1e-15)
        Acx = _csr(s.coarse_matrix(), nagr, nagr).toarray()
        np.testing.assert_allclose(Acx, Rexp @ A.toarray() @ Rexp.T, rtol=1e-12,
                                   atol=1e-12 * np.abs(Acx).max())
        gn, gst, gnodes = s.fine_aggregates()
        np.testing.assert_array_equal(gn, num)


def test_oracle_aggregates_and_subdomains_fed_back_give_the_same_bits():
    n, i, j, v = P.diffusion3d(10, 1.0, 4)
    o = O.Oracle(n, i, j, v, O.params(**CTL))  # one process: coarse-aggregate subdomains
    num, _, _ = o.fine_aggregates()
    st, inds = o.subdomains()
    with _library(n, i, j, v, num=num, subd=(st, inds), **CTL) as s:
        for a, b in zip(s.subdomains(), (st, inds)):
            np.testing.assert_array_equal(a, b)
        ri, rj, rv = o.R()
        R = _csr(s.restriction(), len(np.unique(num[num > 0])), n).tocoo()
        Ro = sp.coo_matrix((rv, (ri - 1, rj - 1)), shape=R.shape).tocsr()
        assert (abs(R.tocsr() - Ro) > 0).nnz == 0
    # with the library's own aggregation the same sets come out
    with _library(n, i, j, v, **CTL) as s2:
        for a, b in zip(s2.subdomains(), (st, inds)):
            np.testing.assert_array_equal(a, b)


def test_injection_errors():
    n, i, j, v = P.laplace2d(6)
    with D.Solver(-1) as s:
        s.set_params(**CTL)
        s.set_matrix(n, i, j, v, 1)
        with pytest.raises(RuntimeError, match="out of range"):
            s.set_aggregates(np.r_[np.ones(n - 1), -1].astype(np.int32))
        bad = np.ones(n, np.int32)
        bad[0] = 3  # aggregate 2 is empty
        with pytest.raises(RuntimeError, match="holds no unknown"):
            s.set_aggregates(bad)
        with pytest.raises(RuntimeError, match="is empty"):
            s.set_subdomains(np.array([1, 1, n + 1]), np.arange(1, n + 1))
        with pytest.raises(RuntimeError, match="twice in one subdomain"):
            s.set_subdomains(np.array([1, 3]), np.array([2, 2]))
        s.set_aggregates(np.ones(n, np.int32))
        s.set_ranks(2)
        with pytest.raises(RuntimeError, match="make their own aggregates"):
            s.setup_host()


# ---- strength flags (doug_set_strong_mask) and R (doug_set_restriction_coo)

def _bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def _check_against_oracle(s, o, n):
    """Aggregates, subdomains, R and A_c of the host setup against the
    oracle, entries and bits."""
    for a, b in zip(o.fine_aggregates(), s.fine_aggregates()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(o.coarse_aggregates(), s.coarse_aggregates()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(o.subdomains(), s.subdomains()):
        np.testing.assert_array_equal(a, b)
    ri, rj, rv = o.R()
    rp, rc, rvals = s.restriction()
    order = np.lexsort((rj, ri))
    np.testing.assert_array_equal(ri[order] - 1, np.repeat(np.arange(len(rp) - 1), np.diff(rp)))
    np.testing.assert_array_equal(rj[order] - 1, rc)
    assert _bits(rv[order], rvals)
    ci, cj, cv = o.AC()
    cp, cc, cvals = s.coarse_matrix()
    np.testing.assert_array_equal(ci - 1, np.repeat(np.arange(len(cp) - 1), np.diff(cp)))
    np.testing.assert_array_equal(cj - 1, cc)
    assert _bits(cv, cvals)


def test_oracle_strength_flags_fed_back_give_the_same_bits():
    # the oracle's own flags (SpMtx_find_strong) through the boundary: every
    # product of the set-up is unchanged; shuffled input order included
    n, i, j, v = P.diffusion3d(10, 2.0, 6)
    perm = np.random.default_rng(7).permutation(len(v))
    i, j, v = i[perm], j[perm], v[perm]
    ctl = dict(CTL, strong1=0.3)  # weak couplings present
    o = O.Oracle(n, i, j, v, O.params(**ctl))
    flags = o.strong()
    assert len(flags) == len(v) and 0 < flags.sum() < len(v)
    with D.Solver(-1) as s:
        s.set_params(**ctl)
        s.set_matrix(n, i, j, v, 1)
        s.set_strong_mask(flags)
        s.setup_host()
        _check_against_oracle(s, o, n)


def test_given_strength_flags_drive_aggregation_and_smoothing():
    # flags unlike the computed ones (a third of the strong couplings
    # dropped at random, the diagonal kept) in the oracle and the library:
    # the same aggregates, R and A_c, and not those of the computed flags
    n, i, j, v = P.diffusion3d(10, 1.0, 9)
    o0 = O.Oracle(n, i, j, v, O.params(**CTL))
    flags = o0.strong().copy()
    rng = np.random.default_rng(11)
    drop = (i != j) & (flags == 1) & (rng.uniform(size=len(v)) < 0.33)
    flags[drop] = 0
    o = O.Oracle(n, i, j, v, O.params(**CTL), strong=flags)
    np.testing.assert_array_equal(o.strong(), flags)
    assert not np.array_equal(o.fine_aggregates()[0], o0.fine_aggregates()[0])
    with D.Solver(-1) as s:
        s.set_params(**CTL)
        s.set_matrix(n, i, j, v, 1)
        s.set_strong_mask(flags)
        s.setup_host()
        _check_against_oracle(s, o, n)
        # cleared: back to the computed flags
        s.set_strong_mask(None)
        s.setup_host()
        _check_against_oracle(s, o0, n)


def test_oracle_restriction_fed_back_gives_the_same_bits():
    # the oracle's R of another smoother count (unsmoothed aggregation) in a
    # smoothers-1 set-up: the library and the oracle, both given that R,
    # form the A_c of the unsmoothed run bit for bit
    n, i, j, v = P.diffusion2d(24, 1.0, 5)
    ctl0 = dict(CTL, smoothers=0)
    o0 = O.Oracle(n, i, j, v, O.params(**ctl0))
    ri, rj, rv = o0.R()
    nagr = o0.L.or_nagr_fine(o0.h)
    o = O.Oracle(n, i, j, v, O.params(**CTL), restriction=(nagr, ri, rj, rv))
    for a, b in zip(o.AC(), o0.AC()):
        assert _bits(a, b) if a.dtype == np.float64 else np.array_equal(a, b)
    with D.Solver(-1) as s:
        s.set_params(**CTL)
        s.set_matrix(n, i, j, v, 1)
        # reversed order: the boundary sorts the triples itself
        s.set_restriction(nagr, ri[::-1], rj[::-1], rv[::-1])
        s.setup_host()
        _check_against_oracle(s, o, n)
        info = s.info()
        assert info["ncoarse"] == nagr


def test_given_restriction_with_repeated_pairs_and_a_partition():
    # an R of other values and sizes (random weights on 2x2 blocks, some
    # pairs given twice) with a partition: oracle and library agree bit for
    # bit on R (sorted, repeated pairs summed in input order) and A_c
    m = 16
    n, i, j, v = P.diffusion2d(m, 1.0, 2)
    g = np.arange(n)
    agg = (g // m) // 2 * (m // 2) + (g % m) // 2  # 0-based, 64 aggregates of 4
    rng = np.random.default_rng(3)
    ri = agg + 1
    rj = g + 1
    rv = rng.uniform(0.2, 1.0, n)
    rep = rng.choice(n, 20, replace=False)
    ri = np.r_[ri, ri[rep]]
    rj = np.r_[rj, rj[rep]]
    rv = np.r_[rv, rng.uniform(0.1, 0.3, 20)]
    perm = rng.permutation(len(rv))
    ri, rj, rv = ri[perm], rj[perm], rv[perm]
    part = P.block_partition((m, m), 4)
    nagr = int(agg.max()) + 1
    o = O.Oracle(n, i, j, v, O.params(**CTL), partition=part, restriction=(nagr, ri, rj, rv))
    with D.Solver(-1) as s:
        s.set_params(**CTL)
        s.set_matrix(n, i, j, v, 1)
        s.set_partition(*part)
        s.set_restriction(nagr, ri, rj, rv)
        s.setup_host()
        oi, oj, ov = o.R()
        rp, rc, rvals = s.restriction()
        order = np.lexsort((oj, oi))
        np.testing.assert_array_equal(oj[order] - 1, rc)
        assert _bits(ov[order], rvals)
        assert len(rc) == n  # every repeated pair summed into one entry
        ci, cj, cv = o.AC()
        cp, cc, cvals = s.coarse_matrix()
        np.testing.assert_array_equal(cj - 1, cc)
        assert _bits(cv, cvals)
        # R A R^T from scipy within rounding
        R = sp.coo_matrix((rv, (ri - 1, rj - 1)), shape=(nagr, n)).tocsr()
        Ai, Aj, Av = o.A()
        A = sp.csr_matrix((Av, (Ai - 1, Aj - 1)), shape=(n, n))
        Ac = _csr((cp, cc, cvals), nagr, nagr).toarray()
        np.testing.assert_allclose(Ac, (R @ A @ R.T).toarray(), rtol=1e-13, atol=1e-13)


def test_strength_and_restriction_errors():
    n, i, j, v = P.laplace2d(8)
    with D.Solver(-1) as s:
        s.set_params(**CTL)
        s.set_matrix(n, i, j, v, 1)
        with pytest.raises(RuntimeError, match="one flag per matrix entry"):
            s.set_strong_mask(np.ones(len(v) - 1, np.uint8))
        with pytest.raises(RuntimeError, match="out of range"):
            s.set_restriction(2, [1, 3], [1, 2], [1.0, 1.0])
        # coarse-aggregate subdomains need one row of R per fine aggregate
        s.set_restriction(3, [1, 2, 3], [1, 2, 3], [1.0, 1.0, 1.0])
        with pytest.raises(RuntimeError, match="rows and the fine aggregation"):
            s.setup_host()
        s.set_restriction(0, None, None, None)
        s.setup_host()
        s.set_strong_mask(np.ones(len(v), np.uint8))
        s.set_ranks(2)
        with pytest.raises(RuntimeError, match="own strength flags"):
            s.setup_host()
        s.set_ranks(0)
        n2, i2, j2, v2 = P.laplace2d(6)
        s.set_matrix(n2, i2, j2, v2, 1)
        with pytest.raises(RuntimeError, match="set for another matrix"):
            s.setup_host()
