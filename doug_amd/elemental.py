"""DOUG's elemental input (input_type 1) for one process, assembled into the
COO matrix and right-hand side the solve path takes.

The files are Fortran unformatted sequential files (each record framed by
its byte count as a 4-byte integer):

- info file (Mesh_initFromFile, src/datatypes/Mesh.F90:293-322): one record
  of five int32, nell ngf nsd mfrelt nnode;
- freedom lists (Mesh_readFileFreelists, Mesh.F90:427-461): one record per
  element, nfrelt then its nfrelt global freedom numbers;
- element matrices and right-hand sides (ElemMtxs_readAndDistribute,
  src/datatypes/ElemMtxs/ElemMtxs_distribute.F90:552-557): per element,
  nfrelt records of nfrelt float64 (record j holds column j), then one record
  of nfrelt float64 for the element right-hand side.

Assembly follows ElemMtxsAssembleContext (ElemMtxs_assemble.F90:158-289)
with one partition: every freedom some element lists is local, numbered in
ascending global order (Mesh_MapsAndNghbrs, Mesh.F90:1098-1195); entry
(i, j) of element e, if nonzero, is added in element order, starting from
zero, to the slot (lf(i), lf(j)); the slots come out by ascending first,
then second index, stored with row and column swapped.  The element
right-hand sides are summed the same way unless a separate global
right-hand side is given (Vect_readAndBroadcastRHS, Vect.F90:603-632).
"""
from __future__ import annotations

import os

import numpy as np


def read_records(path: str) -> list[bytes]:
    """Records of a Fortran unformatted sequential file (4-byte markers)."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        if pos + 4 > len(data):
            raise ValueError(f"{path}: truncated record marker")
        m = int(np.frombuffer(data, "<i4", 1, pos)[0])
        end = pos + 4 + m
        if m < 0 or end + 4 > len(data) or int(np.frombuffer(data, "<i4", 1, end)[0]) != m:
            raise ValueError(f"{path}: bad record at byte {pos}")
        out.append(data[pos + 4:end])
        pos = end + 4
    return out


def write_records(path: str, records) -> None:
    """Write Fortran unformatted sequential records (inverse of read_records)."""
    with open(path, "wb") as f:
        for r in records:
            b = bytes(r)
            m = np.int32(len(b)).tobytes()
            f.write(m + b + m)


def read_info(path: str) -> dict:
    rec = read_records(path)[0]
    nell, ngf, nsd, mfrelt, nnode = np.frombuffer(rec, "<i4", 5)
    return dict(nell=int(nell), ngf=int(ngf), nsd=int(nsd), mfrelt=int(mfrelt), nnode=int(nnode))


def read_freelists(path: str, info: dict) -> list[np.ndarray]:
    recs = read_records(path)
    if len(recs) < info["nell"]:
        raise ValueError(f"{path}: {len(recs)} element records for {info['nell']} elements")
    lists = []
    for r in recs[:info["nell"]]:
        v = np.frombuffer(r, "<i4")
        nf = int(v[0])
        if nf < 0 or nf > info["mfrelt"] or len(v) < 1 + nf:
            raise ValueError(f"{path}: bad freedom list")
        h = v[1:1 + nf].astype(np.int64)
        if nf and (h.max() > info["ngf"] or h.min() < 1):
            raise ValueError("freedom index out of range")  # Mesh.F90:454-457
        lists.append(h)
    return lists


def assemble(info: dict, lists, elem_path: str):
    """(n, indi, indj, val, b, lg) with 1-based COO in DOUG's order and the
    local-to-global freedom map lg (1-based)."""
    used = np.zeros(info["ngf"] + 1, bool)
    for h in lists:
        used[h] = True
    lg = np.flatnonzero(used[1:]) + 1
    gl = np.zeros(info["ngf"] + 1, np.int64)
    gl[lg] = np.arange(1, len(lg) + 1)
    n = len(lg)
    recs = read_records(elem_path)
    b = np.zeros(n + 1)
    ti, tj, tv = [], [], []
    r = 0
    for h in lists:
        nf = len(h)
        if r + nf + 1 > len(recs):
            raise ValueError(f"{elem_path}: too few records")
        E = np.zeros((nf, nf))  # E[i, j]; record j holds column j
        for jj in range(nf):
            E[:, jj] = np.frombuffer(recs[r + jj], "<f8", nf)
        rhs = np.frombuffer(recs[r + nf], "<f8", nf)
        r += nf + 1
        lf = gl[h]
        for i in range(nf):
            for j in range(nf):
                if E[i, j] != 0.0:
                    ti.append(lf[i])
                    tj.append(lf[j])
                    tv.append(E[i, j])
            b[lf[i]] = b[lf[i]] + rhs[i]
    ti, tj, tv = np.array(ti, np.int64), np.array(tj, np.int64), np.array(tv)
    order = np.lexsort((tj, ti))  # stable: equal slots keep element order
    ti, tj, tv = ti[order], tj[order], tv[order]
    first = np.ones(len(ti), bool)
    first[1:] = (ti[1:] != ti[:-1]) | (tj[1:] != tj[:-1])
    slot = np.cumsum(first) - 1
    val = np.zeros(int(first.sum()))
    for k in range(len(tv)):  # sequential sums from zero, element order
        val[slot[k]] = val[slot[k]] + tv[k]
    indi = tj[first].astype(np.int32)  # rows and columns swapped on output
    indj = ti[first].astype(np.int32)
    return n, indi, indj, val, b[1:], lg


def read_elemental(ctl_dir: str, info_file: str, freedom_lists_file: str, elemmat_rhs_file: str,
                   assembled_rhs: np.ndarray | None = None):
    """Read and assemble DOUG elemental input from `ctl_dir`.  Returns
    (n, indi, indj, val, b) like the assembled problems (1-based COO) plus
    lg, the global freedom number of each unknown.  `assembled_rhs`, a
    global right-hand side of ngf values, replaces the elements' one."""
    p = lambda f: f if os.path.isabs(f) else os.path.join(ctl_dir, f)
    info = read_info(p(info_file))
    lists = read_freelists(p(freedom_lists_file), info)
    n, indi, indj, val, b, lg = assemble(info, lists, p(elemmat_rhs_file))
    if assembled_rhs is not None:
        x = np.asarray(assembled_rhs, np.float64)
        if x.shape != (info["ngf"],):
            raise ValueError("the global right-hand side must have ngf values")
        b = x[lg - 1].copy()
    return n, indi, indj, val, b, lg
