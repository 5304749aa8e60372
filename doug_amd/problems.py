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
    kappa = np.asarray(kappa, np.float64).reshape(N, N, N)
    n = N ** 3
    idx = np.arange(n).reshape(N, N, N)
    diag = np.zeros((N, N, N))
    rows, cols, vals = [], [], []
    for ax in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
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
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
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


def uniform_rhs(n: int, seed: int) -> np.ndarray:
    return rng(seed).uniform(-1.0, 1.0, n)
