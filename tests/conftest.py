"""Shared fixtures: reference test data (tests/golden) and builders."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libdoug_hip.so kernels)")
    # torch ships its own HIP runtime; let it initialise before libdoug_hip.so
    # loads /opt/rocm's so that torch tensors can be handed to the C-ABI
    try:
        import torch

        torch.cuda.is_available()
    except Exception:
        pass


def _ensure_built():
    lib = os.path.join(ROOT, "doug_amd", "libdoug_hip.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "doug_amd", "csrc")], check=True)
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


_ensure_built()


def read_assembled_text(path):
    """DOUG assembled text matrix: `n nnz` then `i j v` lines (1-based)."""
    with open(path) as f:
        n, nnz = map(int, f.readline().split())
        d = np.loadtxt(f, ndmin=2)
    assert len(d) == nnz
    return n, d[:, 0].astype(np.int32), d[:, 1].astype(np.int32), d[:, 2].copy()


def read_xdr_vector(path):
    """XDR vector: big-endian int32 length, then big-endian doubles."""
    raw = open(path, "rb").read()
    m = int(np.frombuffer(raw[:4], ">i4")[0])
    return np.frombuffer(raw[4:4 + 8 * m], ">f8").astype(np.float64)


def read_fortran_record(path):
    """One Fortran unformatted sequential record of doubles (4-byte markers)."""
    raw = open(path, "rb").read()
    ln = int(np.frombuffer(raw[:4], "<i4")[0])
    assert int(np.frombuffer(raw[4 + ln:8 + ln], "<i4")[0]) == ln
    return np.frombuffer(raw[4:4 + ln], "<f8").astype(np.float64)


@pytest.fixture(scope="session")
def lap16():
    d = os.path.join(GOLDEN, "lap16x16")
    n, i, j, v = read_assembled_text(os.path.join(d, "Lap16x16.txt"))
    return dict(n=n, i=i, j=j, v=v, b=read_xdr_vector(os.path.join(d, "rhs.xdr")),
                x=read_xdr_vector(os.path.join(d, "csolution.xdr")))


@pytest.fixture(scope="session")
def lap4():
    d = os.path.join(GOLDEN, "lap4x4")
    n, i, j, v = read_assembled_text(os.path.join(d, "Lap4x4.txt"))
    return dict(n=n, i=i, j=j, v=v, b=read_fortran_record(os.path.join(d, "rhs.dat")),
                x=read_fortran_record(os.path.join(d, "solution_correct.dat")))


# the control values of tests/assembled/lap16x16/DOUG.ctl
LAP16_CTL = dict(radius1=2, strong1=0.67, minasize1=2, radius2=1, strong2=0.67, minasize2=2,
                 overlap=-1, symmstruct=1, symmnumeric=1)
# tests/assembled/lap4x4/DOUG.ctl
LAP4_CTL = dict(radius1=2, strong1=0.67, minasize1=2, radius2=2, strong2=0.67, minasize2=2,
                overlap=-1, smoothers=0, symmstruct=1, symmnumeric=1)
