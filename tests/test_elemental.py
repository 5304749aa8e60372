"""DOUG's elemental input (input_type 1): the Python reader/assembler and the
Fortran driver's, pinned by the reference's own e4x4 fixture
(tests/elemental/e4x4: element files plus the known solution csolution.xdr)
and by generated element files with repeated, zero and unused freedoms."""
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import doug_amd as D
import oracle as O
from conftest import GOLDEN, ROOT, read_xdr_vector
from doug_amd import elemental as E

DRIVER = os.path.join(ROOT, "doug_amd", "bin", "doug_aggr")
E4 = os.path.join(GOLDEN, "e4x4")
E4_FILES = ("doug_info.dat", "doug_element.dat", "doug_system.dat")


def _e4x4():
    return E.read_elemental(E4, *E4_FILES)


def test_e4x4_assembles_to_the_known_answer():
    n, i, j, v, b, lg = _e4x4()
    assert n == 9 and list(lg) == list(range(1, 10))
    A = sp.csr_matrix((v, (i - 1, j - 1)), shape=(n, n)).toarray()
    np.testing.assert_allclose(A, A.T, rtol=0, atol=1e-14)
    x = np.linalg.solve(A, b)
    xr = read_xdr_vector(os.path.join(E4, "csolution.xdr"))
    assert np.linalg.norm(x - xr) <= 1e-6  # scripts/doug/test.py:159-172
    # the COO comes out column by column (rows and columns swapped on output)
    np.testing.assert_array_equal(np.lexsort((i, j)), np.arange(len(i)))


def test_e4x4_through_the_oracle_path():
    # the assembled elemental matrix through doug_aggr's one-process setup
    n, i, j, v, b, _ = _e4x4()
    o = O.Oracle(n, i, j, v, O.params(radius1=1, strong1=0.08, smoothers=1, overlap=1,
                                      symmstruct=1, symmnumeric=1, radius2=1))
    x, it, _, _ = o.pcg(b, 1e-12, 100)
    xr = read_xdr_vector(os.path.join(E4, "csolution.xdr"))
    assert np.linalg.norm(x - xr) <= 1e-6


def _write_mesh(d, m, seed, unused=3, zero_elems=2):
    """P1 triangles on an m x m grid of cells; the (m-1)^2 interior nodes are
    freedoms 1.., plus `unused` freedoms no element lists; element matrices
    are stiffness matrices scaled by a random coefficient, a few elements
    entirely zero, element right-hand sides random."""
    g = np.random.default_rng(seed)
    k = m - 1
    free = -np.ones((m + 1, m + 1), np.int64)
    free[1:m, 1:m] = np.arange(1, k * k + 1).reshape(k, k)
    ngf = k * k + unused
    tris = []
    for r in range(m):
        for c in range(m):
            tris.append([(r, c), (r + 1, c), (r + 1, c + 1)])
            tris.append([(r, c), (r + 1, c + 1), (r, c + 1)])
    grad = {0: np.array([[-1, -1], [1, 0], [0, 1]]), 1: np.array([[0, -1], [1, 1], [-1, 0]])}
    lists, mats, rhss = [], [], []
    for t, tri in enumerate(tris):
        G = grad[t % 2].astype(float)
        K = 0.5 * (G @ G.T) * np.exp(g.normal())
        if t in range(5, 5 + zero_elems):
            K = np.zeros_like(K)
        keep = [q for q, (a, c) in enumerate(tri) if free[a, c] > 0]
        lists.append(np.array([free[tri[q]] for q in keep], np.int32))
        mats.append(K[np.ix_(keep, keep)])
        rhss.append(g.uniform(-1, 1, len(keep)))
    E.write_records(os.path.join(d, "info.dat"),
                    [np.array([len(tris), ngf, 2, 3, (m + 1) ** 2], "<i4").tobytes()])
    E.write_records(os.path.join(d, "elem.dat"),
                    [np.concatenate([[len(h)], h]).astype("<i4").tobytes() for h in lists])
    recs = []
    for K, f in zip(mats, rhss):
        recs += [np.ascontiguousarray(K[:, c], "<f8").tobytes() for c in range(K.shape[1])]
        recs.append(np.asarray(f, "<f8").tobytes())
    E.write_records(os.path.join(d, "system.dat"), recs)
    E.write_records(os.path.join(d, "freemap.dat"), [np.arange(1, ngf + 1, dtype="<i4").tobytes()])
    return ngf, lists, mats, rhss


def _ctl(d, extra=""):
    (d / "DOUG.ctl").write_text(
        "solver 2\nmethod 1\ninput_type 1\nlevels 2\ninfo_file info.dat\n"
        "freedom_lists_file elem.dat\nelemmat_rhs_file system.dat\nfreemap_file freemap.dat\n"
        "radius1 2\nstrong1 0.08\nsmoothers 1\noverlap 1\nsymmstruct T\nsymmnumeric T\n"
        "solve_tolerance 1.0e-10\nsolution_format 2\nsolution_file solution.xdr\n" + extra)


def test_generated_mesh_against_direct_assembly(tmp_path):
    m = 12
    ngf, lists, mats, rhss = _write_mesh(str(tmp_path), m, seed=4)
    n, i, j, v, b, lg = E.read_elemental(str(tmp_path), "info.dat", "elem.dat", "system.dat")
    assert n == (m - 1) ** 2 and n < ngf  # unused freedoms are not unknowns
    # brute force: dense sums of the element matrices (entry (lf(i), lf(j))
    # stored as A[lf(j), lf(i)]) and right-hand sides
    A = np.zeros((ngf + 1, ngf + 1))
    f = np.zeros(ngf + 1)
    for h, K, r in zip(lists, mats, rhss):
        A[np.ix_(h, h)] += K.T
        f[h] += r
    A = A[1:n + 1, 1:n + 1]
    got = sp.csr_matrix((v, (i - 1, j - 1)), shape=(n, n)).toarray()
    np.testing.assert_allclose(got, A, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(b, f[1:n + 1], rtol=1e-14, atol=1e-15)
    # structural zeros of all-zero elements never enter the pattern
    assert len(v) == np.count_nonzero(A)
    # a global right-hand side replaces the elements' one through lg
    xg = np.arange(1.0, ngf + 1)
    *_, b2, _ = E.read_elemental(str(tmp_path), "info.dat", "elem.dat", "system.dat", xg)
    np.testing.assert_array_equal(b2, xg[lg - 1])


def test_driver_dump_matches_python(tmp_path):
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "doug_amd", "csrc")], check=True)
    _write_mesh(str(tmp_path), 9, seed=5)
    _ctl(tmp_path, "dump_matrix_only T\ndump_matrix_file dump.txt\n")
    p = subprocess.run([DRIVER, "-f", "DOUG.ctl"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    n, i, j, v, _, _ = E.read_elemental(str(tmp_path), "info.dat", "elem.dat", "system.dat")
    lines = (tmp_path / "dump.txt").read_text().split("\n")
    hdr = lines[0].split()
    assert int(hdr[0]) == n and int(hdr[1]) == len(v)
    rows = np.array([[float(t) for t in ln.split()] for ln in lines[1:] if ln.strip()])
    # the dump lists the matrix arranged by rows (stable)
    order = np.argsort(i, kind="stable")
    np.testing.assert_array_equal(rows[:, 0], i[order])
    np.testing.assert_array_equal(rows[:, 1], j[order])
    np.testing.assert_allclose(rows[:, 2], v[order], rtol=1e-15, atol=0)


def test_driver_rejects_bad_freedom_lists(tmp_path):
    _write_mesh(str(tmp_path), 4, seed=6)
    recs = E.read_records(str(tmp_path / "elem.dat"))
    bad = np.frombuffer(recs[3], "<i4").copy()
    bad[1] = 999  # beyond ngf
    recs[3] = bad.tobytes()
    E.write_records(str(tmp_path / "elem.dat"), recs)
    with pytest.raises(ValueError, match="out of range"):
        E.read_elemental(str(tmp_path), "info.dat", "elem.dat", "system.dat")
    _ctl(tmp_path)
    p = subprocess.run([DRIVER, "-f", "DOUG.ctl"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode != 0 and "out of range" in p.stdout


@pytest.mark.gpu
def test_driver_elemental_run_matches_binding(tmp_path):
    # the driver on element files and the Python binding on the same
    # assembly give the same bits; e4x4 reaches the reference's solution
    _write_mesh(str(tmp_path), 16, seed=7)
    _ctl(tmp_path)
    print("elemental driver run", flush=True)
    p = subprocess.run([DRIVER, "-f", "DOUG.ctl"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    x = read_xdr_vector(tmp_path / "solution.xdr")
    n, i, j, v, b, _ = E.read_elemental(str(tmp_path), "info.dat", "elem.dat", "system.dat")
    with D.Solver(0) as s:
        s.set_params(radius1=2, strong1=0.08, smoothers=1, overlap=1, symmstruct=1, symmnumeric=1)
        s.set_matrix(n, i, j, v, 1)
        s.setup(0)
        xb, it, _, _ = s.pcg(b, 1e-10, 100000)
    np.testing.assert_array_equal(x.view(np.int64), xb.view(np.int64))
    e4 = tmp_path / "e4"
    shutil.copytree(E4, e4)
    (e4 / "run.ctl").write_text((e4 / "DOUG.ctl").read_text() +
                                "smoothers 1\nradius1 1\nradius2 1\nstrong1 0.08\n")
    p = subprocess.run([DRIVER, "-f", "run.ctl"], cwd=e4, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    x4 = read_xdr_vector(e4 / "solution.xdr")
    assert np.linalg.norm(x4 - read_xdr_vector(e4 / "csolution.xdr")) <= 1e-6
