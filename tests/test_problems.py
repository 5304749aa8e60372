"""Synthetic workload generators: closed forms that pin the P1 assembly and
the properties the FEM workload relies on (symmetric, positive definite,
assembled without duplicates)."""
import numpy as np
import scipy.sparse as sp

from doug_amd import problems as P


def _structured_triangles(m):
    """(m+1)^2 grid points of spacing h = 1/m, each square split along one diagonal."""
    x = np.linspace(0.0, 1.0, m + 1)
    pts = np.array([(a, b) for a in x for b in x])
    idx = np.arange((m + 1) ** 2).reshape(m + 1, m + 1)
    tri = []
    for a in range(m):
        for b in range(m):
            p, q, r, s = idx[a, b], idx[a + 1, b], idx[a + 1, b + 1], idx[a, b + 1]
            tri += [(p, q, r), (p, r, s)]
    return pts, np.array(tri)


def test_p1_on_right_triangles_is_the_five_point_laplacian():
    # textbook closed form: on a uniform right-triangle mesh the P1 stiffness
    # matrix is the 5-point stencil (4 on the diagonal, -1 to the axis
    # neighbours, 0 across the split diagonals), independent of h
    m = 6
    pts, tri = _structured_triangles(m)
    K = P.p1_stiffness(pts, tri).toarray()
    interior = [(a, b) for a in range(1, m) for b in range(1, m)]
    for a, b in interior:
        i = a * (m + 1) + b
        expect = np.zeros(len(pts))
        expect[i] = 4.0
        for da, db in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            expect[(a + da) * (m + 1) + b + db] = -1.0
        np.testing.assert_allclose(K[i], expect, atol=1e-12)
    # constants are in the kernel of the Neumann matrix
    np.testing.assert_allclose(K @ np.ones(len(pts)), 0.0, atol=1e-12)


def test_p1_linear_energy_is_exact():
    # P1 reproduces linear functions, so for u = 2x + 3y on any triangulation
    # of the unit square the discrete energy u^T K u equals the integral of
    # |grad u|^2 = 13
    pts, tri = _structured_triangles(5)
    rng = np.random.default_rng(0)
    pts = pts + np.where((pts > 0) & (pts < 1), 0.03 * rng.standard_normal(pts.shape), 0.0)
    K = P.p1_stiffness(pts, tri)
    u = 2 * pts[:, 0] + 3 * pts[:, 1]
    assert abs(u @ (K @ u) - 13.0) < 1e-12


def test_fem_workload_properties():
    n, i, j, v, xy = P.fem_p1_square(3000, 7)
    A = sp.coo_matrix((v, (i - 1, j - 1)), shape=(n, n)).tocsr()
    assert len(xy) == n and ((xy > 0) & (xy < 1)).all()
    assert A.nnz == len(v)  # assembled: no duplicate entries
    assert abs(A - A.T).max() == 0.0
    assert (A.diagonal() > 0).all()
    order = np.lexsort((j, i))
    assert (order == np.arange(len(i))).all()  # rows ascending, columns ascending
    w = np.linalg.eigvalsh(A.toarray())
    assert w.min() > 0
