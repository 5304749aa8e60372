"""Differentiable solve x = A(v)^{-1} b through the Schwarz-PCG path.

`DougSolve.apply(values, b, solver)` returns x for the matrix whose COO
pattern the solver holds and whose entries are `values` (a float64 CUDA
tensor in the pattern's input order).  The backward pass is the adjoint:
a second preconditioned solve with A^T (= A for the symmetric matrices of
this path) gives lambda = A^{-1} dL/dx, then one GPU kernel scatters
-lambda_i x_j onto the pattern (doug_value_grad_device), so
dL/dv_e = -lambda_{i_e} x_{j_e} and dL/db = lambda.

`diffusion3d_values(theta, N)` assembles the 7-point cell-centred diffusion
matrix of `problems.diffusion3d` from per-cell log-coefficients theta
(kappa = exp(theta)) with torch operations, in the same entry order, so the
chain rule to theta is ordinary autograd.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import problems
from .solver import Solver


class PatternSolver:
    """A Solver bound to one COO pattern; the values change per forward.

    `version` counts the value sets taken: a backward pass whose forward ran
    on other values than the solver now holds sets its own values again
    before the adjoint solve."""

    def __init__(self, n, indi, indj, params: dict, partition=None, precision=0, tol=1e-10,
                 maxit=1000, device=0):
        self.n = int(n)
        self.indi = np.ascontiguousarray(indi, np.int32)
        self.indj = np.ascontiguousarray(indj, np.int32)
        self.solver = Solver(device)
        self.solver.set_params(**params)
        self.partition = partition
        self.precision = precision
        self.tol = tol
        self.maxit = maxit
        self.iterations = []
        self.solve_times = []  # seconds per PCG (forward or adjoint)
        self.setup_times = []  # seconds per set_values (host copy + doug_setup)
        self.version = 0

    def set_values(self, values: torch.Tensor) -> int:
        t0 = time.perf_counter()
        v = values.detach().to("cpu", torch.float64).numpy()
        self.solver.set_matrix(self.n, self.indi, self.indj, v, 1)
        if self.partition is not None:
            self.solver.set_partition(*self.partition)
        self.solver.setup(self.precision)
        self.setup_times.append(time.perf_counter() - t0)
        self.version += 1
        return self.version

    def solve(self, b: torch.Tensor) -> torch.Tensor:
        b = b.detach().contiguous()
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, _, _ = self.solver.pcg_device(b.data_ptr(), x.data_ptr(), self.tol, self.maxit)
        self.solve_times.append(time.perf_counter() - t0)
        self.iterations.append(it)
        return x

    def close(self):
        self.solver.close()


class DougSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values: torch.Tensor, b: torch.Tensor, ps: PatternSolver):
        ctx.version = ps.set_values(values)
        x = ps.solve(b)
        ctx.ps = ps
        ctx.save_for_backward(x, values.detach())
        return x

    @staticmethod
    def backward(ctx, gx: torch.Tensor):
        ps: PatternSolver = ctx.ps
        x, values = ctx.saved_tensors
        if ps.version != ctx.version:
            # another forward has set other values since: the adjoint needs
            # this forward's matrix
            ctx.version = ps.set_values(values)
        lam = ps.solve(gx.contiguous())  # adjoint solve, A^T = A
        gv = torch.empty(len(ps.indi), dtype=torch.float64, device=x.device)
        torch.cuda.synchronize()
        ps.solver.value_grad_device(lam.data_ptr(), x.data_ptr(), gv.data_ptr())
        return gv, lam, None


def _diffusion_pattern(N: int):
    """Construction order of problems.diffusion3d and its sort permutation."""
    n = N ** 3
    idx = np.arange(n).reshape(N, N, N)
    rows, cols = [], []
    for ax in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[ax] = slice(0, N - 1)
        sl_b[ax] = slice(1, N)
        ia, ib = idx[tuple(sl_a)].ravel(), idx[tuple(sl_b)].ravel()
        rows += [ia, ib]
        cols += [ib, ia]
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    order = np.lexsort((cols, rows))
    return rows, cols, order


def diffusion3d_values(theta: torch.Tensor, N: int) -> torch.Tensor:
    """Matrix values of problems.diffusion3d for kappa = exp(theta), as a
    differentiable function of theta (shape N^3), in the generator's order."""
    kappa = torch.exp(theta.reshape(N, N, N))
    diag = torch.zeros_like(kappa)
    offs = []
    for ax in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[ax] = slice(0, N - 1)
        sl_b[ax] = slice(1, N)
        ka, kb = kappa[tuple(sl_a)], kappa[tuple(sl_b)]
        kf = 2.0 * ka * kb / (ka + kb)
        # accumulate face coefficients into both cells
        d_a = torch.zeros_like(kappa)
        d_a[tuple(sl_a)] = kf
        d_b = torch.zeros_like(kappa)
        d_b[tuple(sl_b)] = kf
        diag = diag + d_a + d_b
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, 1)
        hi[ax] = slice(N - 1, N)
        bnd = torch.zeros_like(kappa)
        bnd[tuple(lo)] = kappa[tuple(lo)]
        bnd2 = torch.zeros_like(kappa)
        bnd2[tuple(hi)] = kappa[tuple(hi)]
        diag = diag + bnd + bnd2
        offs += [-kf.reshape(-1), -kf.reshape(-1)]
    vals = torch.cat(offs + [diag.reshape(-1)])
    _, _, order = _diffusion_pattern(N)
    return vals[torch.as_tensor(order, device=vals.device)]


def diffusion3d_pattern(N: int):
    """(n, indi, indj) 1-based, identical to problems.diffusion3d's."""
    rows, cols, order = _diffusion_pattern(N)
    return N ** 3, (rows[order] + 1).astype(np.int32), (cols[order] + 1).astype(np.int32)


__all__ = ["DougSolve", "PatternSolver", "diffusion3d_values", "diffusion3d_pattern", "problems"]
