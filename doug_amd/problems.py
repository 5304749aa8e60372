"""Synthetic assembled problems in DOUG's input form (1-based COO triples).

All seeds go through numpy's Generator(PCG64(seed)).  Rows are emitted in
ascending order and, inside a row, columns ascending -- the order DOUG's own
generators produce for a stencil matrix (SpMtx_class.F90:709-787 writes the
diagonal block first; the matrices here are written sorted instead, which
DOUG's stable arrangement preserves).
"""
from __future__ import annotations

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _coo_from_stencil(n, rows, cols, vals):
    order = np.lexsort((cols, rows))
    return (rows[order] + 1).astype(np.int32), (cols[order] + 1).astype(np.int32), vals[order]


def laplace2d(m: int):
    """5-point Laplacian on the m x m interior grid (Dirichlet eliminated):
    diagonal 4, neighbours -1, lexicographic numbering."""
    idx = np.arange(m * m).reshape(m, m)
    r, c, v = [idx.ravel()], [idx.ravel()], [np.full(m * m, 4.0)]
    for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
        a = idx[max(0, -di):m - max(0, di), max(0, -dj):m - max(0, dj)]
        b = idx[max(0, di):m - max(0, -di) or None, max(0, dj):m - max(0, -dj) or None]
        r.append(a.ravel())
        c.append(b.ravel())
        v.append(np.full(a.size, -1.0))
    rows, cols, vals = np.concatenate(r), np.concatenate(c), np.concatenate(v)
    n = m * m
    return (n,) + _coo_from_stencil(n, rows, cols, vals)


def smoothed_gaussian(N: int, seed: int, width: int = 5, dims: int = 3) -> np.ndarray:
    """iid N(0,1) per cell, box-filtered width^dims with reflecting ends,
    normalised to zero mean and unit variance."""
    g = rng(seed).standard_normal((N,) * dims)
    h = width // 2
    for ax in range(dims):
        pad = [(0, 0)] * dims
        pad[ax] = (h, h)
        gp = np.pad(g, pad, mode="reflect")
        cs = np.cumsum(gp, axis=ax)
        zero_shape = list(cs.shape)
        zero_shape[ax] = 1
        cs = np.concatenate([np.zeros(zero_shape), cs], axis=ax)
        hi = np.take(cs, np.arange(width, width + N), axis=ax)
        lo = np.take(cs, np.arange(0, N), axis=ax)
        g = (hi - lo) / width
    g = g - g.mean()
    return g / g.std()


def diffusion3d(N: int, sigma: float = 1.0, seed: int = 0, kappa: np.ndarray | None = None):
    """Cell-centred 7-point diffusion on an N^3 grid.

    Cell coefficients kappa = exp(sigma * g) with g = smoothed_gaussian(N, seed);
    an interior face carries the harmonic mean of its two cells, a boundary face
    (Dirichlet, eliminated) the cell value.  A_ii = sum of the cell's face
    coefficients, A_ij = -face coefficient.  kappa=1 gives the 7-point
    Laplacian (diagonal 6, neighbours -1).
    """
    if kappa is None:
        kappa = np.exp(sigma * smoothed_gaussian(N, seed))
    return _diffusion(N, 3, kappa)


def diffusion2d(N: int, sigma: float = 1.0, seed: int = 0):
    """The 5-point analogue of diffusion3d on an N^2 grid (kappa from a
    smoothed Gaussian field on N^2 cells)."""
    return _diffusion(N, 2, np.exp(sigma * smoothed_gaussian(N, seed, dims=2)))


def _diffusion(N: int, dims: int, kappa: np.ndarray):
    kappa = np.asarray(kappa, np.float64).reshape((N,) * dims)
    n = N ** dims
    idx = np.arange(n).reshape((N,) * dims)
    diag = np.zeros((N,) * dims)
    rows, cols, vals = [], [], []
    for ax in range(dims):
        sl_a = [slice(None)] * dims
        sl_b = [slice(None)] * dims
        sl_a[ax] = slice(0, N - 1)
        sl_b[ax] = slice(1, N)
        ka, kb = kappa[tuple(sl_a)], kappa[tuple(sl_b)]
        kf = 2.0 * ka * kb / (ka + kb)
        diag[tuple(sl_a)] += kf
        diag[tuple(sl_b)] += kf
        ia, ib = idx[tuple(sl_a)].ravel(), idx[tuple(sl_b)].ravel()
        rows += [ia, ib]
        cols += [ib, ia]
        vals += [-kf.ravel(), -kf.ravel()]
        lo = [slice(None)] * dims
        hi = [slice(None)] * dims
        lo[ax] = 0
        hi[ax] = N - 1
        diag[tuple(lo)] += kappa[tuple(lo)]
        diag[tuple(hi)] += kappa[tuple(hi)]
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    vals.append(diag.ravel())
    return (n,) + _coo_from_stencil(n, np.concatenate(rows), np.concatenate(cols),
                                    np.concatenate(vals))


def block_partition(shape, block):
    """Non-overlapping partition of a lexicographically numbered grid into
    boxes of edge `block` (the last box along an axis may be smaller).  Parts
    are numbered box-lexicographically; nodes inside a part ascend.  Returns
    1-based (starts, nodes) for doug_set_partition."""
    shape = tuple(shape)
    coords = np.indices(shape).reshape(len(shape), -1)
    nb = [-(-s // block) for s in shape]
    part = np.zeros(coords.shape[1], np.int64)
    for ax in range(len(shape)):
        part = part * nb[ax] + coords[ax] // block
    order = np.argsort(part, kind="stable")
    counts = np.bincount(part, minlength=int(np.prod(nb)))
    starts = np.concatenate([[0], np.cumsum(counts)]) + 1
    return starts.astype(np.int32), (order + 1).astype(np.int32)


def graded_partition(pts: np.ndarray, k: int = 16, ratio: float = 1.05):
    """Non-overlapping partition of points in the unit square into k x k
    cells whose widths grow in geometric `ratio` along each axis, numbered
    row by row; on a locally refined mesh the parts are uneven (for the
    10^6-point mesh of fem_p1_square: 1.7k-7.7k points).  Nodes ascend
    inside a part.  Returns 1-based (starts, nodes) for doug_set_partition."""
    w = ratio ** np.arange(k)
    e = np.concatenate([[0.0], np.cumsum(w / w.sum())])
    cx = np.clip(np.searchsorted(e, pts[:, 0], side="right") - 1, 0, k - 1)
    cy = np.clip(np.searchsorted(e, pts[:, 1], side="right") - 1, 0, k - 1)
    part = cy * k + cx
    order = np.argsort(part, kind="stable")
    counts = np.bincount(part, minlength=k * k)
    starts = np.concatenate([[0], np.cumsum(counts)]) + 1
    return starts.astype(np.int32), (order + 1).astype(np.int32)


def uniform_rhs(n: int, seed: int) -> np.ndarray:
    return rng(seed).uniform(-1.0, 1.0, n)


def p1_stiffness(pts: np.ndarray, tri: np.ndarray):
    """Assembled P1 stiffness matrix (scipy CSR, duplicates summed) of a
    triangulation: K_e = area * grad(phi_i) . grad(phi_j), with grad(phi_i) =
    perp(p_k - p_j) / det for the cyclic vertex order (i, j, k).  Triangles of
    area below 1e-14 are skipped."""
    from scipy.sparse import coo_matrix

    p0, p1, p2 = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    det = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1])
    area = 0.5 * np.abs(det)
    keep = area > 1e-14
    tri, area, det = tri[keep], area[keep], det[keep]
    P = pts[tri]  # (ne, 3, 2)
    grads = np.empty((len(tri), 3, 2))
    for i in range(3):
        j, k = (i + 1) % 3, (i + 2) % 3
        grads[:, i, 0] = (P[:, j, 1] - P[:, k, 1]) / det
        grads[:, i, 1] = (P[:, k, 0] - P[:, j, 0]) / det
    Ke = area[:, None, None] * np.einsum("eid,ejd->eij", grads, grads)
    rows = np.repeat(tri, 3, axis=1).ravel()
    cols = np.tile(tri, (1, 3)).ravel()
    return coo_matrix((Ke.ravel(), (rows, cols)), shape=(len(pts), len(pts))).tocsr()


def fem_p1_square(npoints: int, seed: int, circle_fraction: float = 0.02):
    """P1 finite-element Poisson (-Laplace u = f, u = 0 on the boundary) on the
    Delaunay triangulation of seeded points in the unit square: uniform points,
    plus `circle_fraction` of them clustered within 0.002 of a random circle
    (local refinement), plus evenly spaced points on the four sides.  Element
    matrices K_e = area * grad(phi_i) . grad(phi_j) are summed into the assembled
    matrix (the elemental-to-assembled step of DOUG's input_type 1), boundary
    unknowns eliminated.  Returns (n, indi, indj, val) 1-based, rows ascending,
    columns ascending within a row, no duplicates, and the interior points."""
    from scipy.spatial import Delaunay

    g = rng(seed)
    m = max(2, int(round(np.sqrt(npoints))))  # points per side
    n_circle = int(npoints * circle_fraction)
    n_uniform = max(0, npoints - n_circle - 4 * (m - 1))
    inner = g.uniform(0.0, 1.0, (n_uniform, 2))
    center = g.uniform(0.3, 0.7, 2)
    radius = g.uniform(0.1, 0.25)
    theta = g.uniform(0.0, 2.0 * np.pi, n_circle)
    rr = radius + g.normal(0.0, 0.002, n_circle)
    ring = center + rr[:, None] * np.stack([np.cos(theta), np.sin(theta)], axis=1)
    ring = ring[(ring > 1e-6).all(axis=1) & (ring < 1.0 - 1e-6).all(axis=1)]
    t = np.linspace(0.0, 1.0, m, endpoint=False)
    side = np.concatenate([np.stack([t, np.zeros(m)], 1), np.stack([np.ones(m), t], 1),
                           np.stack([1.0 - t, np.ones(m)], 1), np.stack([np.zeros(m), 1.0 - t], 1)])
    pts = np.concatenate([side, inner, ring])
    tri = Delaunay(pts).simplices
    K = p1_stiffness(pts, tri)
    boundary = (pts <= 1e-12).any(axis=1) | (pts >= 1.0 - 1e-12).any(axis=1)
    interior = np.flatnonzero(~boundary)
    K = K[interior][:, interior].tocsr()
    K.sort_indices()
    K.sum_duplicates()
    coo = K.tocoo()
    return (len(interior), (coo.row + 1).astype(np.int32), (coo.col + 1).astype(np.int32),
            coo.data.astype(np.float64), pts[interior])
