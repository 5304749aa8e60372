"""Differentiation form: the adjoint behind torch.autograd (doug_amd.autograd)
checked against an FP64 dense reference gradient and central differences."""
import numpy as np
import pytest

from doug_amd import problems as P

pytestmark = pytest.mark.gpu

CTL = dict(radius1=2, strong1=0.08, smoothers=1, overlap=1, symmstruct=1, symmnumeric=1)


def _setup(N=16, block=8):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    from doug_amd.autograd import PatternSolver, diffusion3d_pattern

    n, indi, indj = diffusion3d_pattern(N)
    ps = PatternSolver(n, indi, indj, CTL, partition=P.block_partition((N,) * 3, block), tol=1e-10)
    dev = torch.device("cuda:0")
    theta = torch.tensor(0.5 * P.smoothed_gaussian(N, 11).ravel(), dtype=torch.float64, device=dev)
    theta_t = torch.tensor(0.5 * P.smoothed_gaussian(N, 12).ravel(), dtype=torch.float64, device=dev)
    b = torch.tensor(P.uniform_rhs(n, 13), dtype=torch.float64, device=dev)
    return torch, ps, theta, theta_t, b, n, indi, indj


def _loss_doug(torch, ps, theta, b, x_target, N):
    from doug_amd.autograd import DougSolve, diffusion3d_values

    x = DougSolve.apply(diffusion3d_values(theta, N), b, ps)
    return 0.5 * torch.sum((x - x_target) ** 2)


def _loss_dense(torch, theta, b, x_target, N, indi, indj):
    from doug_amd.autograd import diffusion3d_values

    n = N ** 3
    v = diffusion3d_values(theta, N)
    A = torch.zeros((n, n), dtype=torch.float64)
    A = A.index_put((torch.as_tensor(indi - 1), torch.as_tensor(indj - 1)), v, accumulate=True)
    x = torch.linalg.solve(A, b)
    return 0.5 * torch.sum((x - x_target) ** 2)


def test_gradient_matches_dense_reference():
    N = 16
    torch, ps, theta, theta_t, b, n, indi, indj = _setup(N)
    print(f"adjoint dense check n={n}", flush=True)
    from doug_amd.autograd import DougSolve, diffusion3d_values

    with torch.no_grad():
        x_target = DougSolve.apply(diffusion3d_values(theta_t, N), b, ps)
    th = theta.clone().requires_grad_(True)
    loss = _loss_doug(torch, ps, th, b, x_target, N)
    loss.backward()
    g = th.grad.cpu()
    th_c = theta.cpu().clone().requires_grad_(True)
    loss_d = _loss_dense(torch, th_c, b.cpu(), x_target.cpu(), N, indi, indj)
    loss_d.backward()
    gd = th_c.grad
    assert abs(loss.item() - loss_d.item()) <= 1e-8 * abs(loss_d.item())
    rel = float(torch.linalg.norm(g - gd) / torch.linalg.norm(gd))
    print(f"relative gradient error vs dense FP64: {rel:.2e}; iterations {ps.iterations[-3:]}")
    assert rel <= 1e-6
    ps.close()


def test_two_solves_in_one_graph_match_dense_reference():
    # two forwards on one PatternSolver before one backward: each adjoint
    # solve runs on its own forward's matrix
    N = 16
    torch, ps, theta, theta_t, b, n, indi, indj = _setup(N)
    print(f"two solves in one graph n={n}", flush=True)
    from doug_amd.autograd import DougSolve, diffusion3d_values

    with torch.no_grad():
        x_target = DougSolve.apply(diffusion3d_values(theta_t, N), b, ps)
    th = theta.clone().requires_grad_(True)
    loss = _loss_doug(torch, ps, th, b, x_target, N) + \
        _loss_doug(torch, ps, 0.5 * th, b, x_target, N)
    loss.backward()
    g = th.grad.cpu()
    th_c = theta.cpu().clone().requires_grad_(True)
    loss_d = _loss_dense(torch, th_c, b.cpu(), x_target.cpu(), N, indi, indj) + \
        _loss_dense(torch, 0.5 * th_c, b.cpu(), x_target.cpu(), N, indi, indj)
    loss_d.backward()
    gd = th_c.grad
    rel = float(torch.linalg.norm(g - gd) / torch.linalg.norm(gd))
    print(f"relative gradient error vs dense FP64: {rel:.2e}")
    assert rel <= 1e-6
    ps.close()


def test_gradient_matches_central_differences():
    N = 16
    torch, ps, theta, theta_t, b, n, indi, indj = _setup(N)
    print(f"adjoint finite differences n={n}", flush=True)
    from doug_amd.autograd import DougSolve, diffusion3d_values

    with torch.no_grad():
        x_target = DougSolve.apply(diffusion3d_values(theta_t, N), b, ps)
    th = theta.clone().requires_grad_(True)
    _loss_doug(torch, ps, th, b, x_target, N).backward()
    g = th.grad
    h = 1e-6
    r = np.random.default_rng(5)
    for k in range(4):
        d = torch.tensor(r.standard_normal(n), dtype=torch.float64, device=theta.device)
        d = d / torch.linalg.norm(d)
        with torch.no_grad():
            lp = _loss_doug(torch, ps, theta + h * d, b, x_target, N).item()
            lm = _loss_doug(torch, ps, theta - h * d, b, x_target, N).item()
        fd = (lp - lm) / (2 * h)
        an = float(torch.dot(g, d))
        print(f"direction {k}: fd {fd:.10e} adjoint {an:.10e}")
        # each solve stops at rtol 1e-10, so the difference quotient carries
        # an error of order 1e-10 / h relative to the loss
        assert abs(fd - an) <= 2e-3 * abs(an) + 1e-8
    ps.close()
