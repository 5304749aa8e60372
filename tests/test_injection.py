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
