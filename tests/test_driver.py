"""The Fortran doug_aggr driver (ISO_C_BINDING over the C-ABI): control-file
handling on the CPU, the reference's known-answer run on the GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, read_xdr_vector

DRIVER = os.path.join(ROOT, "doug_amd", "bin", "doug_aggr")


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "doug_amd", "csrc")], check=True)


def _run(cwd, *args, timeout=120):
    return subprocess.run([DRIVER, *args], cwd=cwd, capture_output=True, text=True,
                          timeout=timeout)


def test_missing_control_file_aborts(tmp_path):
    p = _run(tmp_path, "-f", "nope.ctl")
    assert p.returncode != 0 and "Unable to open DOUG control file" in p.stdout


def test_wrong_solver_aborts(tmp_path):
    (tmp_path / "DOUG.ctl").write_text("solver 1\nsolver 2\n")
    p = _run(tmp_path, "-f", "DOUG.ctl")
    # the first occurrence of a key wins (DOUG_utils.F90:728-732)
    assert p.returncode != 0 and "Wrong solution method" in p.stdout
    assert "repeated" in p.stdout


def test_unknown_argument_aborts(tmp_path):
    p = _run(tmp_path, "-x")
    assert p.returncode != 0 and "Unrecognised" in p.stdout


def test_unreadable_matrix_aborts(tmp_path):
    (tmp_path / "DOUG.ctl").write_text("solver 2\nassembled_mtx_file missing.txt\n")
    p = _run(tmp_path, "-f", "DOUG.ctl")
    assert p.returncode != 0 and "Unable to read assembled file" in p.stdout


@pytest.mark.gpu
def test_lap16x16_known_answer_run(tmp_path):
    src = os.path.join(GOLDEN, "lap16x16")
    for f in ("Lap16x16.txt", "rhs.xdr", "DOUG.ctl"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    p = _run(tmp_path, "-f", "DOUG.ctl", "-p")
    print(p.stdout[-2000:])
    assert p.returncode == 0, p.stdout + p.stderr
    x = read_xdr_vector(tmp_path / "solution.xdr")
    xr = read_xdr_vector(os.path.join(src, "csolution.xdr"))
    assert np.linalg.norm(x - xr) <= 1e-6  # scripts/doug/test.py:159-172
    prof = (tmp_path / "prof.00").read_text()
    assert "0:pcg iterations:" in prof


@pytest.mark.gpu
def test_virtual_ranks_run_matches_binding(tmp_path):
    # `virtual_ranks 4` with a rank map file (one rank per line, the layout
    # of the reference's domnums.txt): the driver's solution has the Python
    # binding's bits for the same ranks and reaches the known answer
    import doug_amd as D
    from conftest import LAP16_CTL, read_assembled_text
    src = os.path.join(GOLDEN, "lap16x16")
    for f in ("Lap16x16.txt", "rhs.xdr"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    ctl = open(os.path.join(src, "DOUG.ctl")).read()
    (tmp_path / "run.ctl").write_text("virtual_ranks 4\nrank_map_file ranks.txt\noverlap 1\n"
                                      "smoothers 1\n" + ctl)
    g = np.arange(225)
    rank = (1 + 2 * (g // 15 >= 8) + (g % 15 >= 8)).astype(np.int32)
    (tmp_path / "ranks.txt").write_text("\n".join(str(r) for r in rank) + "\n")
    print("driver virtual ranks lap16x16", flush=True)
    p = _run(tmp_path, "-f", "run.ctl")
    assert p.returncode == 0, p.stdout + p.stderr
    assert "virtual ranks: 4" in p.stdout
    x = read_xdr_vector(tmp_path / "solution.xdr")
    assert np.linalg.norm(x - read_xdr_vector(os.path.join(src, "csolution.xdr"))) <= 1e-6
    n, i, j, v = read_assembled_text(os.path.join(src, "Lap16x16.txt"))
    b = read_xdr_vector(os.path.join(src, "rhs.xdr"))
    with D.Solver(0) as s:
        s.set_params(**dict(LAP16_CTL, overlap=1, smoothers=1))
        s.set_matrix(n, i, j, v, 1)
        s.set_ranks(4, rank)
        s.setup(0)
        xb, it, _, _ = s.pcg(b, 1e-12, 300)
    np.testing.assert_array_equal(x.view(np.int64), xb.view(np.int64))
