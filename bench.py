"""Benchmark: time-to-solve of DOUG's two-level Schwarz PCG at rtol 1e-8 on one
MI355X, plus the SpMV's HBM bandwidth.

Workload (default): 3-D 7-point variable-coefficient diffusion on 128^3 cells
(2,097,152 unknowns, 14,581,760 nonzeros), kappa = exp(sigma g) with g a seeded
Gaussian field box-filtered 5^3 and normalised, sigma = 1; 512 subdomains
(16^3 boxes) with overlap 1; smoothed-aggregation coarse grid (strength 0.08,
one Jacobi smoothing step); PCG to rtol 1e-8 from a zero start; FP64.

One step = doug_setup (aggregation, coarse space, subdomain index sets, the
batched dense Cholesky factorisations) + doug_pcg_device, with the matrix
already handed to the context and the right-hand side resident in HBM.  Both
forms are timed: plain FP64 and FP32 subdomain factors inside the FP64 PCG;
`value` is the faster one's mean time to solve.

Prints one JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3

# HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE + WRITE_SIZE (each
# scaled by the same counters on a copy of known size) for the 128^3
# workload; the run and its summary are in profiles/hbm_traffic_pmc_128cube.md
MEASURED_TRAFFIC_128 = {
    # HBM bytes per launch (read + written) from the FETCH_SIZE / WRITE_SIZE runs of
    # profiles/hbm_traffic_pmc_128cube.md (FETCH_SIZE scaled by the copy-kernel factor);
    # the subdomain solve per decomposition of the workload
    ("k_trsv", 1, "ranks"): 6_062_536_148 + 22_282_240,
    ("k_trsv", 1, "boxes"): 6_062_693_646 + 22_282_240,
    ("k_spmv", 0): 244_787_743 + 16_777_216,
    ("k_spmv", 1): 213_542_137 + 16_777_216,
    ("k_spmv", 2): 222_963_734 + 16_777_216,
    ("k_spmv", 3): 208_341_549 + 16_777_216,
    ("k_spmv", 4): 210_923_588 + 16_777_216,
    ("k_spmv", 5): 207_855_730 + 16_777_216,
}

CTL = dict(radius1=4, minasize1=100, maxasize1=343, strong1=0.08, smoothers=1, overlap=1,
           symmstruct=1, symmnumeric=1)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def spmv_bytes(n, nnz):
    return 12 * nnz + 4 * (n + 1) + 8 * n + 8 * n


def cpu_baseline(size, block, sigma, seed, decomp):
    """The oracle (a C restatement of the reference's algorithm) on the full
    workload: setup + PCG to 1e-8 from zero, wall time on this host's cores.
    Its per-rank setup, subdomain factorisations and subdomain solves run on
    OpenMP threads (OMP_NUM_THREADS), like the reference's processes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from doug_amd import problems as P

    n, i, j, v = P.diffusion3d(size, sigma, seed)
    b = P.uniform_rhs(n, seed + 1)
    t0 = time.perf_counter()
    if decomp == "ranks":
        rank, nr = box_ranks(size, block)
        o = O.Oracle.ranks(n, i, j, v, O.params(**CTL), rank)
        how = f"{nr} processes owning one {block}^3 box each"
    else:
        o = O.Oracle(n, i, j, v, O.params(**CTL),
                     partition=P.block_partition((size,) * 3, block))
        how = f"one process, {(size // block) ** 3} subdomains of {block}^3"
    t1 = time.perf_counter()
    x, it, cond, _ = o.pcg(b, 1e-8, 1000)
    t2 = time.perf_counter()
    threads = O.num_threads()
    return {"value": round(t2 - t0, 3), "unit": "s", "cores": threads, "kind": "port",
            "setup_s": round(t1 - t0, 3), "pcg_s": round(t2 - t1, 3), "iterations": it,
            "sample": f"the whole workload: oracle setup + PCG(1e-8) at {size}^3 ({how}, "
                      f"{it} iterations) on {threads} threads, not extrapolated"}


FEM_CTL = dict(radius1=5, minasize1=60, maxasize1=121, strong1=0.08, smoothers=1, overlap=2,
               symmstruct=1, symmnumeric=1)


def fem_batched_leg(npoints, nsub, nrhs, steps, warmup, seed):
    """Batching form on the unstructured workload: P1 FEM Poisson on the
    Delaunay mesh of `npoints` seeded points (2 % clustered near a circle),
    `nsub` uneven subdomains (cells of growing width), overlap 2, FP32
    subdomain factors, `nrhs` right-hand sides in one batched PCG to 1e-8 with
    the graph-captured iteration; a single-RHS FP64 solve alongside."""
    import torch
    import doug_amd as D
    from doug_amd import problems as P

    t0 = time.perf_counter()
    n, i, j, v, pts = P.fem_p1_square(npoints, seed)
    log(f"[bench] fem mesh: n={n} nnz={len(v)} in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda:0")
    s = D.Solver(0)
    s.set_params(**FEM_CTL)
    s.set_matrix(n, i, j, v, 1)
    # uneven subdomains: 16 x 16 cells of geometrically growing width over
    # the locally refined mesh (nsub must be a square)
    k = int(round(np.sqrt(nsub)))
    s.set_partition(*P.graded_partition(pts, k))
    B = torch.tensor(np.stack([P.uniform_rhs(n, seed + 100 + k) for k in range(nrhs)], axis=1),
                     dtype=torch.float64, device=dev).contiguous()
    X = torch.zeros_like(B)
    s.setup(1)
    s.prepare_batch()
    tune = s.autotune(20)  # includes the batched solve's tiles per step
    times = []
    for k in range(warmup + steps):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        s.setup(1)
        s.prepare_batch()
        tb = time.perf_counter()
        X.zero_()
        its = s.pcg_batch_device(nrhs, B.data_ptr(), X.data_ptr(), 1e-8, 2000)
        torch.cuda.synchronize()
        tc = time.perf_counter()
        if k >= warmup:
            times.append((tb - ta, tc - tb))
        log(f"[bench] fem batched {nrhs} rhs: setup {tb - ta:.3f}s solve {tc - tb:.3f}s "
            f"iterations {int(its.min())}-{int(its.max())}")
    info = s.info()
    sizes = np.diff(s.subdomains()[0])
    # relative residual of every column
    res = []
    for k in range(nrhs):
        xk = X[:, k].contiguous()
        r = torch.empty_like(xk)
        s.spmv_device(xk.data_ptr(), r.data_ptr())
        res.append(float(torch.linalg.norm(B[:, k] - r) / torch.linalg.norm(B[:, k])))
    prec_ms = s.time_operator(4, 3)
    # single right-hand side, plain FP64
    b0 = B[:, 0].contiguous()
    x0 = torch.zeros_like(b0)
    torch.cuda.synchronize()
    ta = time.perf_counter()
    s.setup(0)
    tb = time.perf_counter()
    it64, _, _ = s.pcg_device(b0.data_ptr(), x0.data_ptr(), 1e-8, 2000)
    torch.cuda.synchronize()
    tc = time.perf_counter()
    s.close()
    setup_s = float(np.mean([a for a, _ in times]))
    solve_s = float(np.mean([c for _, c in times]))
    per_rhs = (setup_s + solve_s) / nrhs
    return {"workload": f"P1 FEM Poisson, Delaunay mesh of {npoints} seeded points in the unit square "
                        f"(2% near a circle), {info['nsub']} uneven subdomains (16 x 16 cells of "
                        f"geometrically growing width), overlap 2, FP32 factors, {nrhs} RHS, "
                        f"PCG rtol 1e-8",
            "n": n, "nnz": int(len(v)), "ncoarse": info["ncoarse"],
            "subdomain_sizes_min_median_max": [int(sizes.min()), int(np.median(sizes)), int(sizes.max())],
            "setup_s": round(setup_s, 4), "solve_s": round(solve_s, 4),
            "time_to_solve_all_s": round(setup_s + solve_s, 4), "time_per_rhs_s": round(per_rhs, 5),
            "iterations_min_max": [int(its.min()), int(its.max())],
            "rel_residual_max": max(res), "prec_apply_ms": round(prec_ms, 3),
            "fp64_single": {"setup_s": round(tb - ta, 4), "solve_s": round(tc - tb, 4),
                            "time_to_solve_s": round(tc - ta, 4), "iterations": int(it64)},
            "speedup_per_rhs_vs_fp64_single": round((tc - ta) / per_rhs, 2),
            "autotune": {k: tune[k] for k in ("spmv_rows", "spmv_xcd_order", "trsv_order",
                                              "trsm_tiles_per_step", "trsm_ms")}}


def sgs_leg(n, i, j, v, part, b_h, num_iters=3):
    """The symmetric Gauss-Seidel fine level (fine_method 2) on the same
    workload: time per preconditioner apply (k_sgs and the overlap sum) and
    an FP64 solve to 1e-8.  Not the benchmarked form: a row of subdomain
    Gauss-Seidel is a chain of dependent rows, so each subdomain runs on one
    wavefront and the apply is bound by that chain's latency."""
    import torch
    import doug_amd as D
    dev = torch.device("cuda:0")
    b = torch.tensor(b_h, dtype=torch.float64, device=dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    with D.Solver(0) as s:
        s.set_params(**dict(CTL, fine_method=2, num_iters=num_iters))
        s.set_matrix(n, i, j, v, 1)
        s.set_partition(*part)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = s.setup(0)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        it, cond, hist = s.pcg_device(b.data_ptr(), x.data_ptr(), 1e-8, 100000)
        torch.cuda.synchronize()
        t_iter = time.perf_counter() - t0
        apply_ms = s.time_operator(1, 10)
    import scipy.sparse as sp
    A = sp.csr_matrix((v, (i - 1, j - 1)), shape=(n, n))
    rel = float(np.linalg.norm(b_h - A @ x.cpu().numpy()) / np.linalg.norm(b_h))
    out = {"num_iters": num_iters, "setup_s": round(t_setup, 4), "iterations": int(it),
           "iterations_s": round(t_iter, 4), "time_to_solve_s": round(t_setup + t_iter, 4),
           "sgs_apply_ms": round(apply_ms, 4), "rel_residual": rel,
           "sub_total": int(info["sub_total"]), "nsub": int(info["nsub"])}
    log(f"[bench] sgs fine level: {out}")
    return out


def box_ranks(N, edge):
    """Rank (1-based) of each unknown of the N^3 grid when processes own
    boxes of edge `edge`, numbered box-lexicographically."""
    from doug_amd import problems as P
    st, nd = P.block_partition((N, N, N), edge)
    rank = np.zeros(N ** 3, np.int32)
    for r in range(len(st) - 1):
        rank[nd[st[r] - 1:st[r + 1] - 1] - 1] = r + 1
    return rank, len(st) - 1


def form_leg(n, i, j, v, N, b_h, x_ref_h, steps, warmup, nranks, nsub, edge=16):
    """The same workload decomposed another way, timed as the headline is
    (setup + PCG to 1e-8, both precision forms).  nranks > 0: a DOUG run on
    `nranks` processes with `nsub` subdomains each (doug_set_ranks), ranks
    owning boxes of edge N / nranks^(1/3); nranks = 0: one process with the
    `edge`^3 boxes as the subdomain partition."""
    import torch
    import doug_amd as D
    from doug_amd import problems as P
    dev = torch.device("cuda:0")
    if nranks > 0:
        rank, nr = box_ranks(N, int(round(N / round(nranks ** (1.0 / 3.0)))))
        assert nr == nranks
        out = {"ranks": nranks, "rank_box": int(round(N / round(nranks ** (1.0 / 3.0)))),
               "subdomains_per_rank": nsub}
    else:
        out = {"ranks": 1, "partition": f"{edge}^3 boxes"}
    b = torch.tensor(b_h, dtype=torch.float64, device=dev)
    x_ref = torch.tensor(x_ref_h, dtype=torch.float64, device=dev)
    with D.Solver(0) as s:
        s.set_params(**CTL)
        s.set_matrix(n, i, j, v, 1)
        if nranks > 0:
            s.set_ranks(nranks, rank)
            s.set_num_subdomains(nsub)
        else:
            s.set_partition(*P.block_partition((N, N, N), edge))
        for name, prec in (("fp64", 0), ("fp32_factors", 1)):
            times, setups = [], []
            for rep in range(warmup + steps):
                x = torch.zeros(n, dtype=torch.float64, device=dev)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                info = s.setup(prec)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                it, cond, hist = s.pcg_device(b.data_ptr(), x.data_ptr(), 1e-8, 100000)
                torch.cuda.synchronize()
                t2 = time.perf_counter()
                if rep >= warmup:
                    times.append(t2 - t0)
                    setups.append((t1 - t0, t2 - t1))
            rel_err = float(torch.linalg.norm(x - x_ref) / torch.linalg.norm(x_ref))
            out[name] = {"time_to_solve_s": round(float(np.mean(times)), 4),
                         "setup_s": round(float(np.mean([a for a, _ in setups])), 4),
                         "iterations_s": round(float(np.mean([c for _, c in setups])), 4),
                         "iterations": int(it), "rel_err_vs_fp64_1e-12_solve": rel_err,
                         "t_aggregate": round(info["t_aggregate"], 4),
                         "nsub": int(info["nsub"]), "ncoarse": int(info["ncoarse"]),
                         "sub_total": int(info["sub_total"])}
    log(f"[bench] decomposition {out}")
    return out


def direct_accuracy_leg(block=16):
    """Both precision forms against the committed FP64 sparse-direct solution
    of the 64^3 instance of the same field (tests/golden/sparse_direct_64,
    scipy SuperLU): relative error at rtol 1e-8, and the time-to-solve for
    each form to reach a relative error of at most 1e-8 (the loosest of the
    tried tolerances that gets there), DOUG's decomposition over (64/block)^3
    processes."""
    import torch
    import doug_amd as D
    from doug_amd import problems as P

    fx = np.load(os.path.join(ROOT, "tests", "golden", "sparse_direct_64", "x.npz"))
    N = 64
    n, i, j, v = P.diffusion3d(N, float(fx["sigma"]), int(fx["seed"]))
    b_h = P.uniform_rhs(n, int(fx["seed"]) + 1)
    dev = torch.device("cuda:0")
    b = torch.tensor(b_h, dtype=torch.float64, device=dev)
    xd = torch.tensor(fx["x"], dtype=torch.float64, device=dev)
    rank, nr = box_ranks(N, block)
    out = {"instance": f"3d diffusion {N}^3 (n={n}), {nr} processes of {block}^3",
           "direct_solver": "scipy.sparse.linalg.spsolve (SuperLU, MMD_AT_PLUS_A), FP64, "
                            "committed solution",
           "direct_rel_residual": float(fx["rel_residual"])}
    with D.Solver(0) as s:
        s.set_params(**CTL)
        s.set_matrix(n, i, j, v, 1)
        s.set_ranks(nr, rank)
        for precision, name in ((0, "fp64"), (1, "fp32_factors")):
            tried = []
            for tol in (1e-8, 3e-9, 1e-9, 3e-10, 1e-10):
                best = None
                for rep in range(3):  # the first run of a tolerance warms up
                    x = torch.zeros(n, dtype=torch.float64, device=dev)
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    s.setup(precision)
                    it, _, _ = s.pcg_device(b.data_ptr(), x.data_ptr(), tol, 1000)
                    torch.cuda.synchronize()
                    dt = time.perf_counter() - t0
                    best = dt if best is None or (rep > 0 and dt < best) else best
                err = float(torch.linalg.norm(x - xd) / torch.linalg.norm(xd))
                tried.append({"rtol": tol, "iterations": int(it), "rel_err": err,
                              "time_to_solve_s": round(best, 4)})
                if err <= 1e-8:
                    break
            out[name] = {"rel_err_at_rtol_1e-8": tried[0]["rel_err"],
                         "iterations_at_rtol_1e-8": tried[0]["iterations"],
                         "to_rel_err_1e-8": tried[-1] if tried[-1]["rel_err"] <= 1e-8 else None,
                         "tried": tried}
    t64 = out["fp64"]["to_rel_err_1e-8"]
    t32 = out["fp32_factors"]["to_rel_err_1e-8"]
    if t64 and t32:
        out["speedup_fp32_vs_fp64_to_rel_err_1e-8"] = round(t64["time_to_solve_s"] /
                                                            t32["time_to_solve_s"], 3)
    log(f"[bench] accuracy against the sparse-direct solution: {out}")
    return out


def differentiation_leg(N, block, ndir, seed):
    """Differentiation form: loss 0.5 |x(theta) - x_target|^2 for 3-D diffusion
    on N^3 with per-cell log-coefficients theta, forward and adjoint PCG to
    1e-10; the adjoint gradient against central differences (h = 1e-6) along
    `ndir` random directions."""
    import torch
    from doug_amd import problems as P
    from doug_amd.autograd import DougSolve, PatternSolver, diffusion3d_pattern, diffusion3d_values

    dev = torch.device("cuda:0")
    n, indi, indj = diffusion3d_pattern(N)
    ctl = dict(radius1=2, strong1=0.08, smoothers=1, overlap=1, symmstruct=1, symmnumeric=1)
    ps = PatternSolver(n, indi, indj, ctl, partition=P.block_partition((N,) * 3, block), tol=1e-10)
    theta = torch.tensor(0.5 * P.smoothed_gaussian(N, seed).ravel(), dtype=torch.float64, device=dev)
    theta_t = torch.tensor(0.5 * P.smoothed_gaussian(N, seed + 1).ravel(), dtype=torch.float64, device=dev)
    b = torch.tensor(P.uniform_rhs(n, seed + 2), dtype=torch.float64, device=dev)

    def loss_of(th):
        x = DougSolve.apply(diffusion3d_values(th, N), b, ps)
        return 0.5 * torch.sum((x - x_target) ** 2)

    with torch.no_grad():
        x_target = DougSolve.apply(diffusion3d_values(theta_t, N), b, ps)
    # one untimed forward + backward first: torch's autograd kernels for the
    # assembly load on their first use
    loss_of(theta_t.clone().requires_grad_(True)).backward()
    th = theta.clone().requires_grad_(True)
    torch.cuda.synchronize()
    nsolve, nsetup = len(ps.solve_times), len(ps.setup_times)
    t0 = time.perf_counter()
    loss = loss_of(th)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd_solve, adj_solve = ps.solve_times[nsolve:nsolve + 2]
    fwd_setup = ps.setup_times[nsetup]
    g = th.grad.detach()
    gen = torch.Generator(device="cpu").manual_seed(seed)
    h = 1e-6
    errs = []
    with torch.no_grad():
        for _ in range(ndir):
            d = torch.randn(n, generator=gen, dtype=torch.float64).to(dev)
            d /= torch.linalg.norm(d)
            fd = (loss_of(theta + h * d) - loss_of(theta - h * d)).item() / (2 * h)
            an = float(torch.dot(g, d))
            errs.append(abs(fd - an) / max(abs(an), 1e-300))
    ps.close()
    return {"workload": f"3d diffusion {N}^3, theta = 0.5 g (seeded smoothed Gaussian), "
                        f"{(N // block) ** 3} subdomains, forward + adjoint PCG rtol 1e-10",
            "n": n, "forward_s": round(t1 - t0, 4), "backward_s": round(t2 - t1, 4),
            "forward_setup_s": round(fwd_setup, 4), "forward_pcg_s": round(fwd_solve, 4),
            "adjoint_pcg_s": round(adj_solve, 4),
            "iterations_forward_adjoint": ps.iterations[nsolve:nsolve + 2],
            "fd_directions": ndir, "fd_h": h, "fd_rel_err_max": max(errs) if errs else None,
            "fd_rel_err_median": float(np.median(errs)) if errs else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--block", type=int, default=16)
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mode", choices=["solve", "kernels"], default="solve",
                    help="kernels: one setup, then a few launches of the copy, SpMV and "
                         "subdomain-solve kernels only (for hardware-counter runs)")
    ap.add_argument("--precision", type=int, default=1, help="factor precision for --mode kernels")
    ap.add_argument("--batch", action="store_true",
                    help="--mode kernels: also two iterations of the 32-column batch (eight-wavefront "
                         "envelope solve, one tile per step)")
    ap.add_argument("--fem-points", type=int, default=1_000_000,
                    help="points of the FEM batching workload (0: skip it)")
    ap.add_argument("--diff-size", type=int, default=64,
                    help="grid edge of the differentiation workload (0: skip it)")
    ap.add_argument("--sgs", type=int, default=1,
                    help="1: also time the Gauss-Seidel fine level (fine_method 2) on the workload")
    ap.add_argument("--decomp", choices=["ranks", "boxes"], default="ranks",
                    help="headline decomposition: DOUG on (size/block)^3 processes with one "
                         "block^3 box each (ranks) or one process with the boxes as partition")
    ap.add_argument("--ranks", default="8x64",
                    help="further multi-process legs as RANKSxSUBDOMAINS, comma separated "
                         "('' skips them)")
    ap.add_argument("--direct", type=int, default=1,
                    help="1: accuracy against the committed 64^3 sparse-direct solution")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build runs the whole solve "
              "on one MI355X (its subdomains are co-resident on the GPU).", file=sys.stderr)
        sys.exit(2)

    import torch
    import doug_amd as D
    from doug_amd import problems as P

    N, blk = args.size, args.block
    t0 = time.perf_counter()
    n, i, j, v = P.diffusion3d(N, args.sigma, args.seed)
    b_h = P.uniform_rhs(n, args.seed + 1)
    part = P.block_partition((N, N, N), blk)
    log(f"[bench] generated {N}^3: n={n} nnz={len(v)} in {time.perf_counter() - t0:.1f}s")

    s = D.Solver(0)
    s.set_params(**CTL)
    s.set_matrix(n, i, j, v, 1)
    if args.decomp == "ranks":
        # "METIS into K subdomains" as DOUG runs it: K processes with one
        # subdomain each; process r owns the r-th blk^3 box
        rank, nranks = box_ranks(N, blk)
        s.set_ranks(nranks, rank)
    else:
        s.set_partition(*part)
    dev = torch.device("cuda:0")
    b = torch.tensor(b_h, dtype=torch.float64, device=dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    nnz = len(v)

    if args.mode == "kernels":
        s.setup(args.precision)
        tune = s.autotune(20)
        log(f"[bench] autotune: {tune}")
        s.pcg_device(b.data_ptr(), x.data_ptr(), 1e-8, 2)
        for which, reps in ((3, 5), (0, 5), (1, 3)):
            ms = s.time_operator(which, reps)
            log(f"[bench] operator {which}: {ms:.4f} ms")
        if args.batch:
            s.prepare_batch()
            s.set_option(9, 3)
            s.set_option(6, 1)
            Bm = torch.tensor(np.stack([P.uniform_rhs(n, args.seed + 1 + k) for k in range(32)], axis=1),
                              dtype=torch.float64, device=dev).contiguous()
            Xm = torch.zeros_like(Bm)
            its_b = s.pcg_batch_device(32, Bm.data_ptr(), Xm.data_ptr(), 1e-8, 2)
            torch.cuda.synchronize()
            log(f"[bench] batch: {int(its_b.max())} iterations of 32 columns")
        inf = s.info()
        print(json.dumps({"mode": "kernels", "precision": args.precision, "n": n, "nnz": nnz,
                          "copy_bytes_per_launch": 16 * n, "spmv_bytes": spmv_bytes(n, nnz),
                          "trsv_bytes": 2 * inf["factor_bytes"] + 20 * inf["sub_total"]}))
        s.close()
        return

    def step(precision):
        s.setup(precision)
        x.zero_()
        torch.cuda.synchronize()
        ts = time.perf_counter()
        it, cond, hist = s.pcg_device(b.data_ptr(), x.data_ptr(), 1e-8, 1000)
        return it, cond, time.perf_counter() - ts

    def run_form(precision, name):
        for w in range(args.warmup):
            it, cond, _ = step(precision)
            log(f"[bench] {name} warmup {w}: {it} iterations, setup {s.info()['t_total']:.2f}s")
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        setups, iters_t = [], []
        for k in range(args.steps):
            it, cond, ti = step(precision)
            inf = s.info()
            setups.append(inf["t_total"])
            iters_t.append(ti)
            log(f"[bench] {name} step {k}: {it} it, setup {inf['t_total']:.3f}s "
                f"(aggr {inf['t_aggregate']:.2f} coarse {inf['t_coarse_build']:.2f} "
                f"upload {inf['t_upload']:.2f} factor {inf['t_factor']:.2f} "
                f"cfactor {inf['t_coarse_factor']:.2f}), iterations {ti:.3f}s")
        torch.cuda.synchronize()
        t_total = time.perf_counter() - t_start
        info = s.info()
        # accuracy: true residual and distance to the FP64 reference solve
        xs = x.clone()
        r = torch.empty_like(b)
        s.spmv_device(xs.data_ptr(), r.data_ptr())
        rel_res = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
        rel_err = float(torch.linalg.norm(xs - x_ref) / torch.linalg.norm(x_ref))
        # time-to-solve to a relative error of at most 1e-8 against the
        # FP64 1e-12 solve: the loosest tried stopping tolerance that gets there
        to_err = None
        for tol in (3e-9, 1e-9, 3e-10):
            xe = torch.zeros_like(b)
            torch.cuda.synchronize()
            te0 = time.perf_counter()
            s.setup(precision)
            ite, _, _ = s.pcg_device(b.data_ptr(), xe.data_ptr(), tol, 1000)
            torch.cuda.synchronize()
            te = time.perf_counter() - te0
            erre = float(torch.linalg.norm(xe - x_ref) / torch.linalg.norm(x_ref))
            to_err = {"rtol": tol, "iterations": int(ite), "rel_err": erre,
                      "time_to_solve_s": round(te, 4)}
            if erre <= 1e-8:
                break
        # kernel timings on the library's stream (HIP events)
        s.pcg_device(b.data_ptr(), x.data_ptr(), 1e-8, 2)  # leaves p, r, z populated
        spmv_ms = s.time_operator(0, 50)
        trsv_ms = s.time_operator(1, 5)
        prec_ms = s.time_operator(2, 5)
        spmv_gbs = spmv_bytes(n, nnz) / (spmv_ms * 1e-3) / 1e9
        trsv_bytes = 2 * info["factor_bytes"] + 20 * info["sub_total"]
        trsv_gbs = trsv_bytes / (trsv_ms * 1e-3) / 1e9
        chol_flops = float(info["factor_flops"])  # matrix-core work over the stored envelope
        chol_tflops = chol_flops / max(info["t_factor"], 1e-9) / 1e12
        return {"time_to_solve_s": round(float(np.mean(setups) + np.mean(iters_t)), 4),
                "ms_per_step": round(1000.0 * t_total / args.steps, 2),
                "setup_s": round(float(np.mean(setups)), 4),
                "iterations_s": round(float(np.mean(iters_t)), 4), "iterations": it,
                "rel_residual": rel_res, "rel_err_vs_fp64_1e-12_solve": rel_err,
                "to_rel_err_1e-8": to_err,
                "cond_estimate": cond, "prec_apply_ms": round(prec_ms, 4),
                "trsv_ms": round(trsv_ms, 4), "trsv_bytes": int(trsv_bytes),
                "trsv_GB/s": round(trsv_gbs, 1), "factor_bytes": info["factor_bytes"],
                "factor_tiles": info["factor_tiles"],
                "cholesky_s": round(info["t_factor"], 4), "cholesky_tflops": round(chol_tflops, 3),
                "cholesky_mfma_frac": round(chol_tflops / (FP32_MFMA_PEAK_TFLOPS if precision == 1
                                                           else FP64_MFMA_PEAK_TFLOPS), 4),
                "spmv_ms": spmv_ms, "spmv_GB/s": round(spmv_gbs, 1),
                "setup_breakdown_s": {k: round(info[k], 4) for k in
                                      ("t_aggregate", "t_coarse_build", "t_subdomains",
                                       "t_upload", "t_factor", "t_coarse_factor")},
                "info": {k: info[k] for k in ("nsub", "ncoarse", "sub_total", "overlap")}}

    # FP64 reference solution for the error figures (tighter tolerance)
    s.setup(0)
    x_ref = torch.zeros_like(b)
    it_ref, _, _ = s.pcg_device(b.data_ptr(), x_ref.data_ptr(), 1e-12, 3000)
    log(f"[bench] reference FP64 solve to 1e-12: {it_ref} iterations")
    # launch shapes chosen once per matrix pattern; kept by later setups
    tune = s.autotune(20)
    log(f"[bench] autotune: SpMV {tune['spmv_rows']} rows/workgroup, XCD order "
        f"{tune['spmv_xcd_order']}, subdomain order {tune['trsv_order']}; "
        f"SpMV ms per shape {[round(t, 4) for t in tune['spmv_ms']]}")

    fp64 = run_form(0, "fp64")
    fp32 = run_form(1, "fp32-factors")

    # batched form: 32 right-hand sides through one pass (FP32 factors, explicit
    # inverse subdomain factors), setup and solve timed together
    nb = 32
    Bm = torch.tensor(np.stack([P.uniform_rhs(n, args.seed + 1 + k) for k in range(nb)], axis=1),
                      dtype=torch.float64, device=dev).contiguous()
    Xm = torch.zeros_like(Bm)
    bt = []
    for k in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        tb0 = time.perf_counter()
        s.setup(1)
        s.prepare_batch()
        tb1 = time.perf_counter()
        if k == 0:
            tune_batch = s.autotune(20)  # the batched solve's tiles per step
        Xm.zero_()
        its_b = s.pcg_batch_device(nb, Bm.data_ptr(), Xm.data_ptr(), 1e-8, 1000)
        torch.cuda.synchronize()
        tb2 = time.perf_counter()
        if k >= args.warmup:
            bt.append((tb1 - tb0, tb2 - tb1))
        log(f"[bench] batched {nb} rhs: setup {tb1 - tb0:.3f}s, solve {tb2 - tb1:.3f}s, "
            f"iterations {int(its_b.min())}-{int(its_b.max())}")
    prec32_ms = s.time_operator(4, 3)
    # SpMV on its own: the tuned shape, many back-to-back launches
    spmv_ms = s.time_operator(0, 200)
    spmv_gbs = spmv_bytes(n, nnz) / (spmv_ms * 1e-3) / 1e9
    xr0 = Xm[:, 0].contiguous()
    rr = torch.empty_like(xr0)
    s.spmv_device(xr0.data_ptr(), rr.data_ptr())
    batched = {"nrhs": nb, "setup_s": round(float(np.mean([a for a, _ in bt])), 4),
               "solve_s": round(float(np.mean([c for _, c in bt])), 4),
               "time_to_solve_all_s": round(float(np.mean([a + c for a, c in bt])), 4),
               "time_per_rhs_s": round(float(np.mean([a + c for a, c in bt])) / nb, 5),
               "iterations_min_max": [int(its_b.min()), int(its_b.max())],
               "prec_apply_32rhs_ms": round(prec32_ms, 3),
               "tiles_per_step": tune_batch["trsm_tiles_per_step"],
               "column_layout": tune_batch["trsm_col_groups"],
               "prec_apply_ms_per_tiles_per_step": [round(t, 3) for t in tune_batch["trsm_ms"]],
               "rel_residual_col0": float(torch.linalg.norm(Bm[:, 0] - rr) / torch.linalg.norm(Bm[:, 0])),
               "speedup_per_rhs_vs_fp64_single": round(fp64["time_to_solve_s"] /
                                                       (float(np.mean([a + c for a, c in bt])) / nb), 2)}
    best, best_name = (fp32, "fp32_factors") if fp32["time_to_solve_s"] < fp64["time_to_solve_s"] \
        else (fp64, "fp64")
    info = s.info()
    dominant_trsv = best["iterations"] * best["trsv_ms"] * 1e-3 >= best["cholesky_s"]
    if dominant_trsv:
        traffic = MEASURED_TRAFFIC_128.get(("k_trsv", 1 if best_name == "fp32_factors" else 0,
                                            args.decomp)) if N == 128 and blk == 16 else None
        roof = {"kernel": "k_trsv (batched subdomain forward+backward solves)", "bound": "hbm",
                "achieved": best["trsv_GB/s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(best["trsv_GB/s"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_launch": best["trsv_bytes"], "ms_per_launch": best["trsv_ms"]}
    else:
        peak = FP64_MFMA_PEAK_TFLOPS if best_name == "fp64" else FP32_MFMA_PEAK_TFLOPS
        roof = {"kernel": "k_chol_diag+k_chol_offdiag (batched tile Cholesky)", "bound": "mfma",
                "achieved": best["cholesky_tflops"], "peak": peak, "unit": "TFLOP/s",
                "frac": round(best["cholesky_tflops"] / peak, 4), "traffic": None}

    out = {
        "metric": "time_to_solve_rtol1e-8",
        "value": best["time_to_solve_s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (fp32 subdomain factors)" if best_name == "fp32_factors" else "f64",
        "data": "synthetic",
        "config": {"workload": f"3d 7-point exp(sigma g) diffusion {N}^3, {info['nsub']} subdomains "
                               + (f"(DOUG decomposition over {info['nsub']} processes, one {blk}^3 "
                                  "box and one subdomain each)" if args.decomp == "ranks" else
                                  f"({blk}^3 boxes, one process)")
                               + f" overlap {info['overlap']}, smoothed aggregation, PCG rtol 1e-8",
                   "decomposition": args.decomp,
                   "n": n, "nnz": nnz, "sigma": args.sigma, "seed": args.seed,
                   "ncoarse": info["ncoarse"], "ctl": CTL},
        "form": best_name,
        "speedup_fp32_factors_vs_fp64": round(fp64["time_to_solve_s"] / fp32["time_to_solve_s"], 3),
        "spmv": {"GB/s": round(spmv_gbs, 1), "frac_of_hbm_peak": round(spmv_gbs / HBM_PEAK_GBS, 4),
                 "ms": round(spmv_ms, 5), "bytes": spmv_bytes(n, nnz),
                 "shape": {k: tune[k] for k in ("spmv_rows", "spmv_xcd_order")},
                 "traffic": MEASURED_TRAFFIC_128.get(("k_spmv", tune["spmv_shape"])) if N == 128 else None},
        "fp64": fp64,
        "fp32_factors": fp32,
        "batched_32rhs": batched,
        "autotune": tune,
        "roofline": roof,
    }
    s.close()
    # the other decompositions of the same workload
    out["decompositions"] = {}
    if args.decomp == "ranks":
        out["decompositions"]["one_process"] = form_leg(
            n, i, j, v, N, b_h, x_ref.cpu().numpy(), min(args.steps, 2), min(args.warmup, 1), 0, 1, blk)
    for cfg in [c for c in args.ranks.split(",") if c]:
        nr, ns = (int(t) for t in cfg.split("x"))
        out["decompositions"][f"{nr}_ranks_x_{ns}"] = form_leg(
            n, i, j, v, N, b_h, x_ref.cpu().numpy(), min(args.steps, 2), min(args.warmup, 1), nr, ns)
    if args.sgs:
        out["sgs_fine_level"] = sgs_leg(n, i, j, v, part, b_h)
    if args.fem_points > 0:
        out["fem_batched_32rhs"] = fem_batched_leg(args.fem_points, 256, 32, min(args.steps, 2),
                                                   min(args.warmup, 1), args.seed + 7)
    if args.direct:
        log("[bench] accuracy against the sparse-direct solution ...")
        out["accuracy_vs_sparse_direct"] = direct_accuracy_leg(args.block)
    if args.diff_size > 0:
        log("[bench] differentiation ...")
        out["differentiation"] = differentiation_leg(args.diff_size, 16, 20, args.seed + 11)
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(N, blk, args.sigma, args.seed, args.decomp)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
