"""The CPU oracle pinned against the reference's own fixtures and against
closed forms / brute force.  No GPU."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from conftest import LAP16_CTL, LAP4_CTL
from doug_amd import problems as P


def _l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)))


@pytest.mark.parametrize("smoothers", [0, 1])
@pytest.mark.parametrize("levels", [1, 2])
def test_lap16x16_known_answer(lap16, smoothers, levels):
    # scripts/doug/test.py:159-172 accepts ||x - x_ref||_2 <= 1e-6; the ctl
    # asks for solve_tolerance 1e-12 (tests/assembled/lap16x16/DOUG.ctl)
    o = O.Oracle(lap16["n"], lap16["i"], lap16["j"], lap16["v"],
                 O.params(smoothers=smoothers, levels=levels, **LAP16_CTL))
    x, it, cond, hist = o.pcg(lap16["b"], 1e-12, 300)
    assert _l2(x, lap16["x"]) <= 1e-6
    assert it < 300 and hist[-1] <= 1e-12 * np.linalg.norm(lap16["b"])


def test_lap4x4_known_answer(lap4):
    o = O.Oracle(lap4["n"], lap4["i"], lap4["j"], lap4["v"], O.params(**LAP4_CTL))
    x, it, cond, _ = o.pcg(lap4["b"], 1e-12, 300)
    assert _l2(x, lap4["x"]) <= 1e-6


def test_tql1_closed_form():
    # symmetric tridiagonal (2, -1): eigenvalues 2 - 2 cos(k pi / (n+1))
    n = 40
    d, ierr = O.tql1(np.full(n, 2.0), np.concatenate([[0.0], np.full(n - 1, -1.0)]))
    k = np.arange(1, n + 1)
    assert ierr == 0
    np.testing.assert_allclose(d, np.sort(2 - 2 * np.cos(k * np.pi / (n + 1))), rtol=0, atol=1e-13)


def test_tql1_random_against_lapack():
    r = np.random.default_rng(3)
    for n in (1, 2, 7, 33):
        dd, ee = r.standard_normal(n), r.standard_normal(n)
        T = np.diag(dd) + np.diag(ee[1:], 1) + np.diag(ee[1:], -1)
        d, ierr = O.tql1(dd, ee)
        assert ierr == 0
        np.testing.assert_allclose(d, np.linalg.eigvalsh(T), atol=1e-12)


def test_lanczos_condition_of_laplacian():
    # unpreconditioned CG (method 0) on the 5-point Laplacian: the Lanczos
    # estimate converges to lambda_max / lambda_min of the operator
    m = 12
    n, i, j, v = P.laplace2d(m)
    o = O.Oracle(n, i, j, v, O.params(method=0, radius1=2, strong1=0.67, symmstruct=1,
                                      symmnumeric=1, smoothers=0))
    b = np.random.default_rng(7).uniform(-1, 1, n)  # excites every eigenvector
    x, it, cond, _ = o.pcg(b, 1e-14, 500)
    k = np.arange(1, m + 1)
    lam = 4 - 2 * np.cos(k[:, None] * np.pi / (m + 1)) - 2 * np.cos(k[None, :] * np.pi / (m + 1))
    assert abs(cond / (lam.max() / lam.min()) - 1) < 1e-6


def test_spmv_csr_matches_dense():
    r = np.random.default_rng(0)
    A = sp.random(50, 50, density=0.1, random_state=1, format="csr")
    x = r.standard_normal(50)
    y = O.spmv_csr(A.indptr, A.indices, A.data, x)
    np.testing.assert_allclose(y, A @ x, rtol=1e-14, atol=1e-14)


def _bfs_layers(indptr, indices, start_nodes, nlayers):
    out = list(start_nodes)
    seen = set(out)
    front = list(start_nodes)
    for _ in range(nlayers):
        nxt = []
        for u in front:
            for w in indices[indptr[u]:indptr[u + 1]]:
                if w not in seen:
                    seen.add(w)
                    nxt.append(w)
        out += nxt
        front = nxt
    return out


@pytest.mark.parametrize("nlayers", [0, 1, 2, 3])
def test_add_layers_is_bfs(nlayers):
    r = np.random.default_rng(nlayers)
    A = sp.random(200, 200, density=0.02, random_state=nlayers, format="csr")
    A = (A + A.T + sp.eye(200)).tocsr()
    A.sort_indices()
    nodes = r.choice(200, 15, replace=False)
    got = O.add_layers(A.indptr + 1, A.indices + 1, 200, nodes + 1, nlayers) - 1
    assert list(got) == _bfs_layers(A.indptr, A.indices, list(nodes), nlayers)


def _path(n):
    i = np.concatenate([np.arange(n), np.arange(n - 1), np.arange(1, n)])
    j = np.concatenate([np.arange(n), np.arange(1, n), np.arange(n - 1)])
    v = np.concatenate([np.full(n, 2.0), np.full(2 * n - 2, -1.0)])
    o = np.lexsort((j, i))
    return n, i[o] + 1, j[o] + 1, v[o]


def test_aggregation_hand_traced_chain():
    # 7-node path, radius 1, maxasize 3, traced by hand through lets_colour /
    # can_add_roundlayer: seeds 1, 3, 5, 7 -> {1,2}, {3,4}, {5,6}, {7}
    n, i, j, v = _path(7)
    base = dict(radius1=1, maxasize1=3, strong1=0.67, symmstruct=1, symmnumeric=1,
                smoothers=0, radius2=1)
    o = O.Oracle(n, i, j, v, O.params(minasize1=1, **base), partition=([1, 8], np.arange(1, 8)))
    assert list(o.fine_aggregates()[0]) == [1, 1, 2, 2, 3, 3, 4]
    # with minasize 2 the singleton {7} is too small; its only neighbour
    # aggregate cannot eat it (weight 0.5 < alpha 0.67) so node 7 is given
    # away to aggregate 3 (|a~| = 0.5 >= beta = alpha/4)
    o = O.Oracle(n, i, j, v, O.params(minasize1=2, **base), partition=([1, 8], np.arange(1, 8)))
    assert list(o.fine_aggregates()[0]) == [1, 1, 2, 2, 3, 3, 3]


def _dense_preconditioner(o, n):
    st, inds = o.subdomains()
    Ai, Aj, Av = o.A()
    A = sp.coo_matrix((Av, (Ai - 1, Aj - 1)), shape=(n, n)).toarray()
    M = np.zeros((n, n))
    for s in range(len(st) - 1):
        idx = inds[st[s] - 1:st[s + 1] - 1] - 1
        Rs = np.zeros((len(idx), n))
        Rs[np.arange(len(idx)), idx] = 1
        M += Rs.T @ np.linalg.solve(A[np.ix_(idx, idx)], Rs)
    ri, rj, rv = o.R()
    nc = int(ri.max())
    R = sp.coo_matrix((rv, (ri - 1, rj - 1)), shape=(nc, n)).toarray()
    ci, cj, cv = o.AC()
    Ac = sp.coo_matrix((cv, (ci - 1, cj - 1)), shape=(nc, nc)).toarray()
    np.testing.assert_allclose(Ac, R @ A @ R.T, rtol=1e-12, atol=1e-12)
    return M + R.T @ np.linalg.solve(Ac, R)


@pytest.mark.parametrize("smoothers", [0, 1])
def test_two_level_schwarz_matches_dense_formula(smoothers):
    n, i, j, v = P.diffusion3d(8, 1.0, seed=5)
    o = O.Oracle(n, i, j, v, O.params(radius1=1, strong1=0.08, smoothers=smoothers, overlap=1,
                                      symmstruct=1, symmnumeric=1),
                 partition=P.block_partition((8, 8, 8), 4))
    M = _dense_preconditioner(o, n)
    r = np.random.default_rng(1).standard_normal(n)
    np.testing.assert_allclose(o.prec(r), M @ r, rtol=1e-10, atol=1e-12)


def test_restriction_matches_formula():
    # R = T S^T with S = I - omega D_f^{-1} A_strong, D_f the filtered diagonal
    n, i, j, v = P.diffusion3d(6, 1.0, seed=2)
    alpha = 0.08
    o = O.Oracle(n, i, j, v, O.params(radius1=2, strong1=alpha, smoothers=1, overlap=1,
                                      symmstruct=1, symmnumeric=1),
                 partition=P.block_partition((6, 6, 6), 3))
    Ai, Aj, Av = o.A()
    A = sp.csr_matrix((Av, (Ai - 1, Aj - 1)), shape=(n, n))
    d = A.diagonal()
    As = sp.diags(1 / np.sqrt(d)) @ A @ sp.diags(1 / np.sqrt(d))
    Ac = As.tocoo()
    off = Ac.row != Ac.col
    mx = np.zeros(n)
    np.maximum.at(mx, Ac.row[off], np.abs(Ac.data[off]))
    strong = off & (np.abs(Ac.data) > alpha * mx[Ac.row])
    Acoo = A.tocoo()
    dfilt = d + np.bincount(Acoo.row[off & ~strong], Acoo.data[off & ~strong], minlength=n)
    sval = np.where(~off, 1 - 0.33333333333333,
                    np.where(strong, -0.33333333333333 * Acoo.data / dfilt[Acoo.row], 0.0))
    S = sp.csr_matrix((sval, (Acoo.row, Acoo.col)), shape=(n, n))
    num, st, nodes = o.fine_aggregates()
    T = sp.csr_matrix((np.ones(n), (num - 1, np.arange(n))), shape=(st.size - 1, n))
    Rexp = (T @ S.T).toarray()
    ri, rj, rv = o.R()
    Rgot = sp.coo_matrix((rv, (ri - 1, rj - 1)), shape=Rexp.shape).toarray()
    np.testing.assert_allclose(Rgot, Rexp, rtol=1e-14, atol=1e-15)
    # explicit zeros of weak couplings are part of R's pattern
    pattern = (T @ (abs(A) + sp.eye(n)).T).toarray() != 0
    got_pat = np.zeros_like(pattern)
    got_pat[ri - 1, rj - 1] = True
    assert (got_pat == pattern).all()


def _smoothers_reference(o, n, alpha, smoothers, omega=0.33333333333333):
    """R for smoothers >= 2 from the formulas of CoarseMtx.F90:152-417 with
    scipy products: S = I(1-w) - w D1^-1 A_strong (D1 = a_ii + weak row
    sum), S2 the same with D2 = a_ii + the weak entries stored after the
    diagonal, SS = S2 S, and the chain each value of smoothers selects."""
    Ai, Aj, Av = o.A()  # storage order: rows arranged, entries as stored
    r, c = Ai - 1, Aj - 1
    d = np.zeros(n)
    d[r[r == c]] = Av[r == c]
    sc = Av / np.sqrt(np.abs(d[r])) / np.sqrt(np.abs(d[c]))
    off = r != c
    mx = np.zeros(n)
    np.maximum.at(mx, r[off], np.abs(sc[off]))
    strong = off & (np.abs(sc) > alpha * mx[r])
    d1 = d + np.bincount(r[off & ~strong], Av[off & ~strong], minlength=n)
    d2 = np.zeros(n)
    weak_before_diag = 0
    seen_diag = np.zeros(n, bool)
    for k in range(len(Av)):
        if r[k] == c[k]:
            d2[r[k]] = Av[k]
            seen_diag[r[k]] = True
        elif not strong[k]:
            d2[r[k]] += Av[k]
            weak_before_diag += not seen_diag[r[k]]

    def smoother(dd):
        val = np.where(~off, 1 - omega, np.where(strong, -omega * Av / dd[r], 0.0))
        return sp.csr_matrix((val, (r, c)), shape=(n, n))

    S, S2 = smoother(d1), smoother(d2)
    SS = S2 @ S
    X = {3: lambda: SS @ S2, 4: lambda: SS @ SS, 5: lambda: SS @ SS @ S2,
         6: lambda: SS @ SS @ SS, 7: lambda: SS @ SS @ SS @ S2,
         8: lambda: (SS @ SS) @ (SS @ SS)}.get(smoothers, lambda: SS.T)()
    num, st, _ = o.fine_aggregates()
    T = sp.csr_matrix((np.ones(n), (num - 1, np.arange(n))), shape=(st.size - 1, n))
    return (T @ X).toarray(), weak_before_diag


@pytest.mark.parametrize("smoothers", [2, 3, 4, 5, 6, 7, 8, 9])
def test_restriction_powers_match_formula(smoothers):
    # above 8 the reference warns and builds the smoothers=2 product
    n, i, j, v = P.diffusion2d(40, 2.0, seed=3)
    alpha = 0.6
    o = O.Oracle(n, i, j, v, O.params(radius1=9, strong1=alpha, smoothers=smoothers, overlap=1,
                                      symmstruct=1, symmnumeric=1),
                 partition=P.block_partition((40, 40), 20))
    assert o.smoothers == smoothers
    Rexp, weak_before = _smoothers_reference(o, n, alpha, smoothers)
    assert weak_before > 0  # the second smoother's diagonal differs from the first's
    ri, rj, rv = o.R()
    Rgot = sp.coo_matrix((rv, (ri - 1, rj - 1)), shape=Rexp.shape).toarray()
    np.testing.assert_allclose(Rgot, Rexp, rtol=1e-13, atol=1e-15)
    # the coarse matrix is still the Galerkin product R A R^T
    Ai, Aj, Av = o.A()
    A = sp.csr_matrix((Av, (Ai - 1, Aj - 1)), shape=(n, n)).toarray()
    ci, cj, cv = o.AC()
    Ac = sp.coo_matrix((cv, (ci - 1, cj - 1)), shape=(Rexp.shape[0],) * 2).toarray()
    np.testing.assert_allclose(Ac, Rexp @ A @ Rexp.T, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("iters", [1, 3])
def test_sgs_fine_preconditioner_matches_formula(iters):
    # fine_method 2 (FinePreconditioner_sgs.F90, stationary.F90:33-88): per
    # subdomain I, `iters` symmetric Gauss-Seidel sweeps on A_II from x = 0,
    # forward x <- (D+L)^-1 (b - U x), backward x <- (D+U)^-1 (b - L x),
    # summed over the overlapping subdomains (one level: no coarse term)
    import scipy.linalg as sla
    n, i, j, v = P.diffusion3d(8, 1.0, seed=6)
    o = O.Oracle(n, i, j, v, O.params(radius1=2, strong1=0.08, smoothers=1, overlap=1, levels=1,
                                      fine_method=2, num_iters=iters, symmstruct=1,
                                      symmnumeric=1),
                 partition=P.block_partition((8, 8, 8), 4))
    Ai, Aj, Av = o.A()
    A = sp.csr_matrix((Av, (Ai - 1, Aj - 1)), shape=(n, n)).toarray()
    r = np.random.default_rng(7).uniform(-1, 1, n)
    starts, inds = o.subdomains()
    z = np.zeros(n)
    for a in range(len(starts) - 1):
        I = np.sort(inds[starts[a] - 1:starts[a + 1] - 1] - 1)
        M = A[np.ix_(I, I)]
        Lo, Up, Dg = np.tril(M, -1), np.triu(M, 1), np.diag(np.diag(M))
        x = np.zeros(len(I))
        for _ in range(iters):
            x = sla.solve_triangular(Dg + Lo, r[I] - Up @ x, lower=True)
            x = sla.solve_triangular(Dg + Up, r[I] - Lo @ x, lower=False)
        z[I] += x
    np.testing.assert_allclose(o.prec(r), z, rtol=1e-12, atol=1e-14)
    # num_iters unset means three sweeps (aggr.F90:220-222)
    o3 = O.Oracle(n, i, j, v, O.params(radius1=2, strong1=0.08, smoothers=1, overlap=1, levels=1,
                                       fine_method=2, symmstruct=1, symmnumeric=1),
                  partition=P.block_partition((8, 8, 8), 4))
    if iters == 3:
        np.testing.assert_array_equal(o3.prec(r), o.prec(r))
