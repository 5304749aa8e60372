"""ctypes view of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is the checker the MI355X library is compared against.  Only
tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py
import this module; the product package never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Params(C.Structure):
    """Mirror of or_params (control values, src/utils/globals.F90:85-116)."""

    _fields_ = [
        ("radius1", C.c_int), ("radius2", C.c_int),
        ("minasize1", C.c_int), ("minasize2", C.c_int),
        ("maxasize1", C.c_int), ("maxasize2", C.c_int),
        ("strong1", C.c_double), ("strong2", C.c_double),
        ("smoothers", C.c_int), ("overlap", C.c_int), ("levels", C.c_int),
        ("method", C.c_int), ("fine_method", C.c_int),
        ("symmstruct", C.c_int), ("symmnumeric", C.c_int),
        ("omega", C.c_double),
        ("num_iters", C.c_int),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = bind(C.CDLL(_LIB))
    return _lib


def bind(L):
    """Declares the oracle's entry points on a loaded library (also used for
    the tests' mutant builds of doug_oracle.c)."""
    L.or_params_default.argtypes = [C.POINTER(Params)]
    L.or_reset_process.argtypes = []
    L.or_setup.restype = C.c_void_p
    L.or_setup.argtypes = [C.c_int, C.c_long, _i32p, _i32p, _f64p, C.POINTER(Params),
                           C.c_int, C.c_void_p, C.c_void_p]
    L.or_set_injection.argtypes = [C.c_long, C.c_void_p, C.c_int, C.c_long, C.c_void_p,
                                   C.c_void_p, C.c_void_p]
    L.or_free.argtypes = [C.c_void_p]
    L.or_get_strong.restype = C.c_long
    L.or_get_strong.argtypes = [C.c_void_p, C.c_void_p]
    L.or_free_ptr.argtypes = [C.c_void_p]
    L.or_setup_ranks.restype = C.c_void_p
    L.or_setup_ranks.argtypes = [C.c_int, C.c_long, _i32p, _i32p, _f64p, C.POINTER(Params),
                                 C.c_int, _i32p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_last_error.restype = C.c_char_p
    L.or_num_threads.restype = C.c_int
    for name in ("or_n", "or_nagr_fine", "or_nagr_coarse", "or_nsub", "or_overlap",
                 "or_smoothers"):
        getattr(L, name).argtypes = [C.c_void_p]
        getattr(L, name).restype = C.c_int
    for name in ("or_sub_total", "or_R_nnz", "or_AC_nnz", "or_A_nnz"):
        getattr(L, name).argtypes = [C.c_void_p]
        getattr(L, name).restype = C.c_long
    L.or_get_fine_aggr.argtypes = [C.c_void_p, _i32p, _i32p, _i32p]
    L.or_get_coarse_aggr.argtypes = [C.c_void_p, _i32p, _i32p, _i32p]
    L.or_get_subdomains.argtypes = [C.c_void_p, _i32p, _i32p]
    for name in ("or_get_R", "or_get_AC", "or_get_A"):
        getattr(L, name).argtypes = [C.c_void_p, _i32p, _i32p, _f64p]
    L.or_spmv.argtypes = [C.c_void_p, _f64p, _f64p]
    L.or_prec_apply.argtypes = [C.c_void_p, _f64p, _f64p]
    L.or_pcg.restype = C.c_int
    L.or_pcg.argtypes = [C.c_void_p, _f64p, _f64p, C.c_double, C.c_int,
                         C.POINTER(C.c_double), _f64p]
    L.or_spmv_csr.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
    L.or_lanczos_eigs.restype = C.c_int
    L.or_lanczos_eigs.argtypes = [C.c_int, _f64p, _f64p, _f64p]
    L.or_tql1.restype = C.c_int
    L.or_tql1.argtypes = [C.c_int, _f64p, _f64p]
    L.or_add_layers.restype = C.c_int
    L.or_add_layers.argtypes = [_i32p, _i32p, C.c_int, _i32p, C.c_int, C.c_int]
    L.or_aggregate.restype = C.c_int
    L.or_aggregate.argtypes = [C.c_int, C.c_long, _i32p, _i32p, _f64p, C.c_int, C.c_int, C.c_int,
                               C.c_int, C.c_double, C.c_int, _i32p, _i32p]
    return L


def params(**kw) -> Params:
    p = Params()
    lib().or_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class Oracle:
    """One doug_aggr process (numprocs == 1) restated on the CPU."""

    def __init__(self, n, indi, indj, val, p: Params | None = None, partition=None,
                 fresh_process=True, strong=None, restriction=None):
        """strong: flags per input entry replacing the computed strength;
        restriction: (nagr, ri, rj, rv), 1-based triples replacing the
        smoothed aggregation's R."""
        L = self.L = lib()
        if fresh_process:
            L.or_reset_process()
        self.p = p if p is not None else params()
        indi = np.ascontiguousarray(indi, dtype=np.int32)
        indj = np.ascontiguousarray(indj, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        keep = []
        if strong is not None or restriction is not None:
            sm = None if strong is None else np.ascontiguousarray(strong, dtype=np.uint8)
            if restriction is not None:
                nagr, ri, rj, rv = restriction
                ri = np.ascontiguousarray(ri, dtype=np.int32)
                rj = np.ascontiguousarray(rj, dtype=np.int32)
                rv = np.ascontiguousarray(rv, dtype=np.float64)
                keep += [ri, rj, rv]
                L.or_set_injection(len(val), None if sm is None else sm.ctypes.data, int(nagr),
                                   len(rv), ri.ctypes.data, rj.ctypes.data, rv.ctypes.data)
            else:
                L.or_set_injection(len(val), sm.ctypes.data, 0, 0, None, None, None)
            keep.append(sm)
        if partition is not None:
            ps, pn = partition
            self._ps = np.ascontiguousarray(ps, dtype=np.int32)
            self._pn = np.ascontiguousarray(pn, dtype=np.int32)
            h = L.or_setup(n, len(val), indi, indj, val, C.byref(self.p), len(self._ps) - 1,
                           self._ps.ctypes.data, self._pn.ctypes.data)
        else:
            h = L.or_setup(n, len(val), indi, indj, val, C.byref(self.p), 0, None, None)
        if not h:
            raise RuntimeError("oracle setup failed: " + L.or_last_error().decode())
        self.h = h
        self.n = n

    @classmethod
    def ranks(cls, n, indi, indj, val, p: Params, rank_of_node, L=None):
        """One DOUG run on several processes restated on the CPU:
        rank_of_node[i] in 1..nranks.  `self.rank_maps` holds each rank's
        (ninner, local-to-global map), 1-based.  L: another build of the
        oracle (the tests' planted-mistake builds)."""
        L = L or lib()
        self = cls.__new__(cls)
        self.L = L
        self.p = p
        self.n = n
        r = np.ascontiguousarray(rank_of_node, dtype=np.int32)
        nr = int(r.max())
        lgp = (C.POINTER(C.c_int) * nr)()
        nlf = np.zeros(nr, np.int32)
        nin = np.zeros(nr, np.int32)
        h = L.or_setup_ranks(n, len(val), np.ascontiguousarray(indi, dtype=np.int32),
                             np.ascontiguousarray(indj, dtype=np.int32),
                             np.ascontiguousarray(val, dtype=np.float64), C.byref(self.p), nr, r, 1,
                             C.cast(lgp, C.c_void_p), nlf.ctypes.data, nin.ctypes.data)
        if not h:
            raise RuntimeError("oracle setup failed: " + L.or_last_error().decode())
        self.h = h
        self.rank_maps = []
        for k in range(nr):
            self.rank_maps.append((int(nin[k]), np.ctypeslib.as_array(lgp[k], (int(nlf[k]),)).copy()))
            L.or_free_ptr(C.cast(lgp[k], C.c_void_p))
        return self

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_free(self.h)
            self.h = None

    # ---- setup products ----
    def strong(self):
        """The fine strength flags per input entry (one process)."""
        m = self.L.or_get_strong(self.h, None)
        out = np.zeros(m, np.uint8)
        self.L.or_get_strong(self.h, out.ctypes.data)
        return out

    def fine_aggregates(self):
        """fAggr%inner as (num, starts, nodes), 1-based."""
        L = self.L
        na = L.or_nagr_fine(self.h)
        num = np.zeros(self.n, np.int32)
        st = np.zeros(na + 1, np.int32)
        nodes = np.zeros(self.n, np.int32)
        L.or_get_fine_aggr(self.h, num, st, nodes)
        return num, st, nodes[: st[-1] - 1].copy()

    def coarse_aggregates(self):
        """cAggr%full over the coarse unknowns, 1-based."""
        L = self.L
        nc = L.or_nagr_fine(self.h)
        na = L.or_nagr_coarse(self.h)
        num = np.zeros(nc, np.int32)
        st = np.zeros(na + 1, np.int32)
        nodes = np.zeros(nc, np.int32)
        L.or_get_coarse_aggr(self.h, num, st, nodes)
        return num, st, nodes[: st[-1] - 1].copy()

    def subdomains(self):
        L = self.L
        ns = L.or_nsub(self.h)
        tot = L.or_sub_total(self.h)
        st = np.zeros(ns + 1, np.int32)
        inds = np.zeros(max(tot, 1), np.int32)
        L.or_get_subdomains(self.h, st, inds)
        return st, inds[:tot]

    def _coo(self, getter, nnz):
        i = np.zeros(max(nnz, 1), np.int32)
        j = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1), np.float64)
        getter(self.h, i, j, v)
        return i[:nnz], j[:nnz], v[:nnz]

    def R(self):
        L = self.L
        return self._coo(L.or_get_R, L.or_R_nnz(self.h))

    def AC(self):
        L = self.L
        return self._coo(L.or_get_AC, L.or_AC_nnz(self.h))

    def A(self):
        L = self.L
        return self._coo(L.or_get_A, L.or_A_nnz(self.h))

    @property
    def overlap(self):
        return self.L.or_overlap(self.h)

    @property
    def smoothers(self):
        return self.L.or_smoothers(self.h)

    # ---- operators ----
    def spmv(self, x):
        y = np.zeros(self.n)
        self.L.or_spmv(self.h, np.ascontiguousarray(x, np.float64), y)
        return y

    def prec(self, r):
        z = np.zeros(self.n)
        self.L.or_prec_apply(self.h, np.ascontiguousarray(r, np.float64), z)
        return z

    def pcg(self, b, tol, maxit, x0=None):
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64)
        cap = maxit if maxit > 0 else 100000
        hist = np.zeros(cap)
        cond = C.c_double(0.0)
        it = self.L.or_pcg(self.h, np.ascontiguousarray(b, np.float64), x, tol, maxit,
                          C.byref(cond), hist)
        return x, it, cond.value, hist[:it].copy()


def spmv_csr(rowptr, col, val, x):
    n = len(rowptr) - 1
    y = np.zeros(n)
    lib().or_spmv_csr(n, np.ascontiguousarray(rowptr, np.int32),
                      np.ascontiguousarray(col, np.int32),
                      np.ascontiguousarray(val, np.float64),
                      np.ascontiguousarray(x, np.float64), y)
    return y


def aggregate(n, indi, indj, val, radius, minasize, maxasize, alpha, coarse=False, symm=True,
              L=None):
    """SpMtx_find_strong + SpMtx_aggregate on one matrix inside the current
    process state (allmind, beta carry over from earlier calls until
    reset_process()).  Returns (nagr, inner numbers, full numbers)."""
    L = L or lib()
    ni = np.zeros(n, np.int32)
    nf = np.zeros(n, np.int32)
    na = L.or_aggregate(n, len(val), np.ascontiguousarray(indi, dtype=np.int32),
                        np.ascontiguousarray(indj, dtype=np.int32),
                        np.ascontiguousarray(val, dtype=np.float64), int(symm), radius, minasize,
                        maxasize, alpha, int(coarse), ni, nf)
    return na, ni, nf


def num_threads() -> int:
    """Threads of the oracle's parallel loops."""
    return lib().or_num_threads()


def reset_process(L=None):
    (L or lib()).or_reset_process()


def lanczos_eigs(alpha, beta):
    it = len(alpha)
    eig = np.zeros(it)
    ierr = lib().or_lanczos_eigs(it, np.ascontiguousarray(alpha, np.float64),
                                 np.ascontiguousarray(beta, np.float64), eig)
    return eig, ierr


def tql1(d, e):
    d = np.array(d, np.float64)
    e = np.array(e, np.float64)
    ierr = lib().or_tql1(len(d), d, e)
    return d, ierr


def add_layers(bounds1, adj1, n, nodes1, nlayers):
    """1-based CSR graph and node list; returns the expanded node list."""
    buf = np.zeros(n, np.int32)
    buf[: len(nodes1)] = nodes1
    on = lib().or_add_layers(np.ascontiguousarray(bounds1, np.int32),
                             np.ascontiguousarray(adj1, np.int32), n, buf, len(nodes1), nlayers)
    return buf[:on].copy()
