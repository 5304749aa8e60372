"""Launch-shape autotuner: every SpMV shape and both subdomain dispatch
orders give the same bits (SpMV against the oracle, solves against the
default launch), and doug_autotune's choice is one of the measured shapes."""
import numpy as np
import pytest

import doug_amd as D
import oracle as O
from doug_amd import problems as P

pytestmark = pytest.mark.gpu

CTL = dict(radius1=2, strong1=0.08, smoothers=1, overlap=1, symmstruct=1, symmnumeric=1)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


def _slabs(N, widths):
    """Slabs of uneven thickness along the slowest axis (1-based)."""
    assert sum(widths) == N
    starts = np.cumsum([0] + [w * N * N for w in widths]) + 1
    return starts.astype(np.int32), np.arange(1, N ** 3 + 1, dtype=np.int32)


def _solver(n, i, j, v, part):
    s = D.Solver(0)
    s.set_params(**CTL)
    s.set_matrix(n, i, j, v, 1)
    s.set_partition(*part)
    s.setup()
    return s


def test_every_spmv_shape_is_bit_exact():
    # 13,824 rows: 108 blocks of 128 rows, so the XCD renumbering has a tail
    N = 24
    n, i, j, v = P.diffusion3d(N, 1.0, 3)
    part = P.block_partition((N,) * 3, 12)
    print(f"spmv shapes n={n}", flush=True)
    o = O.Oracle(n, i, j, v, O.params(**CTL), partition=part)
    x = P.uniform_rhs(n, 5)
    ref = o.spmv(x)
    with _solver(n, i, j, v, part) as s:
        for shape in range(6):
            s.set_option(2, shape)
            assert s.tuning()["spmv_rows"] == 128 << (shape // 2)
            assert _bits_equal(s.spmv(x), ref), shape


def test_dispatch_order_and_autotune_keep_the_bits():
    N = 20
    n, i, j, v = P.diffusion3d(N, 1.0, 4)
    part = _slabs(N, [3, 7, 2, 8])
    b = P.uniform_rhs(n, 6)
    print(f"dispatch orders n={n} slabs 3/7/2/8", flush=True)
    with _solver(n, i, j, v, part) as s:
        r = P.uniform_rhs(n, 8)
        z0 = s.prec(r)
        x0, it0, _, h0 = s.pcg(b, 1e-8, 500)
        s.set_option(3, 1)
        assert s.tuning()["trsv_order"] == 1
        assert _bits_equal(s.prec(r), z0)
        x1, it1, _, h1 = s.pcg(b, 1e-8, 500)
        assert it1 == it0 and _bits_equal(x1, x0) and _bits_equal(h1, h0)
        tu = s.autotune(5)
        print(tu)
        assert 0 <= tu["spmv_shape"] < 6
        ms = tu["spmv_ms"]
        assert all(t > 0 for t in ms) and ms[tu["spmv_shape"]] == min(ms)
        assert tu["trsv_ms"][0] > 0 and tu["trsv_ms"][1] > 0  # sizes differ: both timed
        x2, it2, _, h2 = s.pcg(b, 1e-8, 500)
        assert it2 == it0 and _bits_equal(x2, x0) and _bits_equal(h2, h0)
        # the choice survives a new setup of the same context
        s.setup()
        assert s.tuning()["spmv_shape"] == tu["spmv_shape"]
        x3, it3, _, _ = s.pcg(b, 1e-8, 500)
        assert it3 == it0 and _bits_equal(x3, x0)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_tiles_per_step_keep_the_bits(precision):
    # one or two off-diagonal tiles staged per step: the same products in the
    # same order, so the batched preconditioner gives the same bits
    N = 16
    n, i, j, v = P.diffusion3d(N, 1.0, 8)
    part = P.block_partition((N,) * 3, 8)
    B = np.stack([P.uniform_rhs(n, 300 + k) for k in range(6)], axis=1)
    print(f"batch tiles per step precision={precision}", flush=True)
    with _solver(n, i, j, v, part) as s:
        s.setup(precision)
        s.prepare_batch()
        out = {}
        for kb in (1, 2):
            s.set_option(6, kb)
            assert s.tuning()["trsm_tiles_per_step"] == kb
            out[kb] = s.pcg_batch(B, 1e-8, 500)
        np.testing.assert_array_equal(out[1][1], out[2][1])
        assert _bits_equal(out[1][0], out[2][0])
        tu = s.autotune(4)
        assert tu["trsm_tiles_per_step"] in (1, 2) and min(tu["trsm_ms"]) > 0
        assert tu["trsm_col_groups"] in (1, 2, 3)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_column_groups_keep_the_bits(precision):
    # 32-column batches (here 20 right-hand sides, padded to 32) solved by one
    # workgroup per subdomain, by two of 16 columns each, or by one of eight
    # wavefronts in two 16-column halves: the same tile products per column
    # in the same order, so the same bits
    N = 16
    n, i, j, v = P.diffusion3d(N, 1.0, 9)
    part = P.block_partition((N,) * 3, 8)
    B = np.stack([P.uniform_rhs(n, 400 + k) for k in range(20)], axis=1)
    print(f"batch column groups precision={precision}", flush=True)
    with _solver(n, i, j, v, part) as s:
        s.setup(precision)
        s.prepare_batch()
        out = {}
        for cg in (1, 2, 3):
            for kb in (1, 2):
                s.set_option(6, kb)
                s.set_option(9, cg)
                assert s.tuning()["trsm_col_groups"] == cg
                out[(cg, kb)] = s.pcg_batch(B, 1e-8, 500)
        ref = out[(1, 2)]
        for key, o in out.items():
            np.testing.assert_array_equal(o[1], ref[1])
            assert _bits_equal(o[0], ref[0]), key
