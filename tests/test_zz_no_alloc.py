"""doug_prec_apply_device_noalloc against doug_prec_apply_device, through
the C-ABI: bit-identical results on two set-ups in a row, the second
smaller, each applied twice (the scratch array keeps the previous call's
contents); a set-up needing more work space than the array holds is refused."""
import ctypes as C

import numpy as np
import pytest

import doug_amd as D
from doug_amd import _lib
from doug_amd import problems as P

pytestmark = pytest.mark.gpu

CTL = dict(radius1=2, strong1=0.08, smoothers=1, overlap=1, symmstruct=1, symmnumeric=1)


def _apply_both(torch, s, n, seed):
    dev = torch.device("cuda:0")
    r = torch.tensor(np.random.default_rng(seed).uniform(-1, 1, n), dtype=torch.float64, device=dev)
    z1 = torch.full_like(r, np.nan)
    z2 = torch.full_like(r, np.nan)
    torch.cuda.synchronize()
    L = _lib.load()
    assert L.doug_prec_apply_device(s.h, C.c_void_p(r.data_ptr()), C.c_void_p(z1.data_ptr())) == 0
    rc = L.doug_prec_apply_device_noalloc(s.h, C.c_void_p(r.data_ptr()), C.c_void_p(z2.data_ptr()))
    assert rc == 0, L.doug_last_error(s.h)
    return z1.cpu().numpy(), z2.cpu().numpy()


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_apply_matches_apply(precision):
    torch = pytest.importorskip("torch")
    for N, block in ((24, 12), (16, 8)):  # the second set-up is smaller
        n, i, j, v = P.diffusion3d(N, 1.0, N)
        print(f"no-alloc apply n={n} precision={precision}", flush=True)
        with D.Solver(0) as s:
            s.set_params(**CTL)
            s.set_matrix(n, i, j, v, 1)
            s.set_partition(*P.block_partition((N,) * 3, block))
            s.setup(precision)
            for seed in (1, 2):
                z1, z2 = _apply_both(torch, s, n, seed)
                assert np.array_equal(z1.view(np.int64), z2.view(np.int64))


def test_noalloc_apply_with_virtual_ranks_and_sgs():
    torch = pytest.importorskip("torch")
    n, i, j, v = P.diffusion3d(20, 1.0, 5)
    st, nd = P.block_partition((20,) * 3, 10)
    rank = np.zeros(n, np.int32)
    for r in range(len(st) - 1):
        rank[nd[st[r] - 1:st[r + 1] - 1] - 1] = r + 1
    for ctl in (CTL, dict(CTL, fine_method=2), dict(CTL, levels=1)):
        print(f"no-alloc apply ranks {ctl}", flush=True)
        with D.Solver(0) as s:
            s.set_params(**ctl)
            s.set_matrix(n, i, j, v, 1)
            s.set_ranks(8, rank)
            s.setup(0)
            z1, z2 = _apply_both(torch, s, n, 3)
            assert np.array_equal(z1.view(np.int64), z2.view(np.int64))


def test_noalloc_apply_refuses_oversized_setups():
    torch = pytest.importorskip("torch")
    # 160^3 with 20^3 boxes and overlap 1: about 5.4M subdomain entries
    N = 160
    n, i, j, v = P.diffusion3d(N, 0.0, 0)
    print(f"no-alloc apply oversized n={n}", flush=True)
    with D.Solver(0) as s:
        s.set_params(**dict(CTL, radius1=4, minasize1=100, maxasize1=343))
        s.set_matrix(n, i, j, v, 1)
        s.set_partition(*P.block_partition((N,) * 3, 20))
        s.setup(1)
        assert s.info()["sub_total"] > 4194304
        r = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        z = torch.zeros_like(r)
        torch.cuda.synchronize()
        L = _lib.load()
        rc = L.doug_prec_apply_device_noalloc(s.h, C.c_void_p(r.data_ptr()), C.c_void_p(z.data_ptr()))
        assert rc == 1  # DOUG_E_ARG
        assert b"scratch array holds" in L.doug_last_error(s.h)
