"""Writes x.npz: the FP64 sparse-direct solution of the 64^3 instance of the
3-D exp(sigma g) diffusion workload (sigma 1, field seed 1, right-hand side
uniform[-1,1] from seed 2), the accuracy reference of the precision form.

Solver: scipy.sparse.linalg.spsolve (SuperLU) with the MMD_AT_PLUS_A column
order, FP64.  Run once in the build container (a few minutes); the result is
committed so that GPU tests compare against it without a direct solve.

    python tests/golden/sparse_direct_64/make_fixture.py
"""
import os
import sys
import time

import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", ".."))
from doug_amd import problems as P  # noqa: E402

N, SIGMA, SEED = 64, 1.0, 1


def main():
    n, i, j, v = P.diffusion3d(N, SIGMA, SEED)
    b = P.uniform_rhs(n, SEED + 1)
    A = sps.coo_matrix((v, (i - 1, j - 1)), shape=(n, n)).tocsc()
    t0 = time.perf_counter()
    x = spla.spsolve(A, b, permc_spec="MMD_AT_PLUS_A")
    dt = time.perf_counter() - t0
    res = float(np.linalg.norm(A @ x - b) / np.linalg.norm(b))
    print(f"n={n} direct solve {dt:.1f}s relative residual {res:.3e}")
    np.savez_compressed(os.path.join(HERE, "x.npz"), x=x, n=n, sigma=SIGMA, seed=SEED,
                        rel_residual=res)


if __name__ == "__main__":
    main()
