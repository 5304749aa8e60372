"""The product's host setup (libdoug_hip.so, host-only context) against the
oracle: aggregates, subdomain index sets, restriction and coarse matrix must
agree bit for bit.  No GPU."""
import numpy as np
import pytest

import doug_amd as D
import oracle as O
from conftest import LAP16_CTL
from doug_amd import problems as P
from test_oracle import _path


def _compare(n, i, j, v, part=None, **ctl):
    o = O.Oracle(n, i, j, v, O.params(**ctl), partition=part)
    with D.Solver(-1) as s:
        s.set_params(**ctl)
        s.set_matrix(n, i, j, v, 1)
        if part is not None:
            s.set_partition(*part)
        info = s.setup_host()
        for a, b in zip(o.fine_aggregates(), s.fine_aggregates()):
            np.testing.assert_array_equal(a, b)
        if part is None:
            for a, b in zip(o.coarse_aggregates(), s.coarse_aggregates()):
                np.testing.assert_array_equal(a, b)
        for a, b in zip(o.subdomains(), s.subdomains()):
            np.testing.assert_array_equal(a, b)
        assert info["overlap"] == o.overlap and info["smoothers"] == o.smoothers
        # R: same entries and bits (oracle keeps it column-arranged)
        ri, rj, rv = o.R()
        rp, rc, rvals = s.restriction()
        order = np.lexsort((rj, ri))
        np.testing.assert_array_equal(ri[order] - 1, np.repeat(np.arange(len(rp) - 1), np.diff(rp)))
        np.testing.assert_array_equal(rj[order] - 1, rc)
        np.testing.assert_array_equal(rv[order].view(np.int64), rvals.view(np.int64))
        # A_c: same entry order and bits
        ci, cj, cv = o.AC()
        cp, cc, cvals = s.coarse_matrix()
        np.testing.assert_array_equal(ci - 1, np.repeat(np.arange(len(cp) - 1), np.diff(cp)))
        np.testing.assert_array_equal(cj - 1, cc)
        np.testing.assert_array_equal(cv.view(np.int64), cvals.view(np.int64))
        return info


@pytest.mark.parametrize("smoothers", [0, 1])
def test_lap16x16_setup(lap16, smoothers):
    _compare(lap16["n"], lap16["i"], lap16["j"], lap16["v"], smoothers=smoothers, **LAP16_CTL)


def test_laplace2d_128_coarse_aggregate_subdomains():
    n, i, j, v = P.laplace2d(128)
    info = _compare(n, i, j, v, radius1=2, strong1=0.08, smoothers=1, overlap=1, symmstruct=1,
                    symmnumeric=1, radius2=3)
    assert info["nsub"] > 1


def test_laplace2d_128_four_blocks():
    n, i, j, v = P.laplace2d(128)
    _compare(n, i, j, v, part=P.block_partition((128, 128), 64), radius1=2, strong1=0.08,
             smoothers=1, overlap=1, symmstruct=1, symmnumeric=1)


@pytest.mark.parametrize("seed", [1, 2])
def test_diffusion3d_setup(seed):
    n, i, j, v = P.diffusion3d(16, 1.0, seed)
    _compare(n, i, j, v, radius1=2, strong1=0.08, smoothers=1, overlap=1, symmstruct=1,
             symmnumeric=1, radius2=2)
    _compare(n, i, j, v, part=P.block_partition((16, 16, 16), 8), radius1=3, maxasize1=125,
             minasize1=40, strong1=0.08, smoothers=1, overlap=2, symmstruct=1, symmnumeric=1)


def test_row_scaling_path():
    # symmstruct/symmnumeric unset -> row scaling (SpMtx_arrangement.F90:337-344)
    n, i, j, v = P.diffusion3d(10, 1.0, 4)
    _compare(n, i, j, v, radius1=2, strong1=0.25, smoothers=1, overlap=1, radius2=2)


def test_hand_traced_chain_product():
    n, i, j, v = _path(7)
    with D.Solver(-1) as s:
        s.set_params(radius1=1, maxasize1=3, minasize1=2, strong1=0.67, symmstruct=1,
                     symmnumeric=1, smoothers=0, radius2=1)
        s.set_matrix(n, i, j, v, 1)
        s.set_partition([1, 8], np.arange(1, 8))
        s.setup_host()
        assert list(s.fine_aggregates()[0]) == [1, 1, 2, 2, 3, 3, 3]


def test_errors_are_reported():
    with D.Solver(-1) as s:
        with pytest.raises(D.DougError, match="STATE"):
            s.setup_host()
        with pytest.raises(D.DougError, match="ARG"):
            s.set_matrix(3, [1, 2, 4], [1, 2, 3], [1.0, 1.0, 1.0], 1)
        s.set_matrix(3, [1, 2, 3], [1, 2, 3], [1.0, 1.0, 1.0], 1)
        with pytest.raises(D.DougError, match="ARG"):
            s.set_partition([1, 3], [1, 1], 1)
        with pytest.raises(D.DougError, match="STATE"):
            s.setup()  # host-only context does no GPU work


@pytest.mark.parametrize("smoothers", [2, 3, 4, 5, 6, 7, 8, 9])
def test_smoothed_restriction_powers(smoothers):
    # smoothers >= 2 (CoarseMtx.F90:236-417); 9 takes the reference's
    # fallback to the smoothers=2 product
    n, i, j, v = P.diffusion2d(40, 2.0, seed=3)
    info = _compare(n, i, j, v, part=P.block_partition((40, 40), 20), radius1=9, strong1=0.6,
                    smoothers=smoothers, overlap=1, symmstruct=1, symmnumeric=1)
    assert info["smoothers"] == smoothers


def test_smoothers_clamped_to_radius():
    # CoarseMtx.F90:75-78: smoothers > radius1 is reduced to radius1
    n, i, j, v = P.diffusion2d(24, 1.0, seed=4)
    info = _compare(n, i, j, v, part=P.block_partition((24, 24), 12), radius1=2, strong1=0.3,
                    smoothers=5, overlap=1, symmstruct=1, symmnumeric=1)
    assert info["smoothers"] == 2
