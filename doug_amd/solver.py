"""Python view of the MI355X Schwarz-PCG path, mirroring the doug_aggr flow.

The order of calls follows src/main/aggr.F90:123-292 of the reference:
matrix (Distribution_NewInit) -> setup (aggregates, coarse space, subdomains,
factorisations) -> pcg_weigs.  Every compute call goes through
libdoug_hip.so; errors raise DougError, the Python counterpart of DOUG_abort.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _lib

CTL_INT_KEYS = ("radius1", "radius2", "minasize1", "minasize2", "maxasize1", "maxasize2",
                "smoothers", "overlap", "levels", "method", "fine_method", "num_iters")


class DougError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_lib.ERRORS.get(code, code)}: {msg}")
        self.code = code


def default_params() -> _lib.DougParams:
    p = _lib.DougParams()
    _lib.load().doug_params_default(C.byref(p))
    return p


def make_params(**kw) -> _lib.DougParams:
    """doug_params from DOUG.ctl-style keyword values (booleans as 0/1)."""
    p = default_params()
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(f"unknown control {k!r}")
        setattr(p, k, int(v) if k in CTL_INT_KEYS or k in ("symmstruct", "symmnumeric") else float(v))
    return p


def _fence():
    """Finish torch's queued device work before the library reads or writes
    device pointers: its kernels run on their own stream, which does not wait
    for torch's stream (a torch copy or fill may otherwise still be running)."""
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_initialized():
        torch.cuda.synchronize()


class Solver:
    """One doug_ctx: one DOUG process with all its subdomains on one GPU."""

    def __init__(self, device: int = 0):
        """device: HIP ordinal, or -1 for a host-only context (setup products
        without GPU work)."""
        self._L = _lib.load()
        h = C.c_void_p()
        rc = self._L.doug_ctx_create(device, C.byref(h))
        if rc != 0:
            raise DougError(rc, "doug_ctx_create failed (no visible GPU or bad device)")
        self.h = h
        self.n = 0

    # -- plumbing --
    def _check(self, rc):
        if rc != 0:
            raise DougError(rc, self._L.doug_last_error(self.h).decode(errors="replace"))

    def close(self):
        if getattr(self, "h", None):
            self._L.doug_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- inputs --
    def set_params(self, params: _lib.DougParams | None = None, **kw):
        p = params if params is not None else make_params(**kw)
        self._check(self._L.doug_set_params(self.h, C.byref(p)))
        self.params = p

    def set_matrix(self, n, indi, indj, val, base=1):
        indi = np.ascontiguousarray(indi, np.int32)
        indj = np.ascontiguousarray(indj, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        self._check(self._L.doug_set_matrix_coo(self.h, int(n), len(val), _lib.ptr(indi),
                                                _lib.ptr(indj), _lib.ptr(val), int(base)))
        self.n = int(n)

    def set_partition(self, starts, nodes, base=1):
        starts = np.ascontiguousarray(starts, np.int32)
        nodes = np.ascontiguousarray(nodes, np.int32)
        self._check(self._L.doug_set_partition(self.h, len(starts) - 1, _lib.ptr(starts),
                                               _lib.ptr(nodes), int(base)))

    def set_num_subdomains(self, nsub: int):
        """doug_set_num_subdomains: subdomains from parts of the fine-aggregate graph."""
        self._check(self._L.doug_set_num_subdomains(self.h, int(nsub)))

    def set_aggregates(self, num):
        """doug_set_aggregates: the fine aggregate (1-based, 0 = none) of
        each unknown; None restores the aggregation."""
        if num is None:
            self._check(self._L.doug_set_aggregates(self.h, 0, None))
            return
        a = np.ascontiguousarray(num, np.int32)
        self._check(self._L.doug_set_aggregates(self.h, int(a.max()), _lib.ptr(a)))

    def set_strong_mask(self, strong):
        """doug_set_strong_mask: one strength flag per matrix entry, in the
        input order of set_matrix; None restores the computed flags."""
        if strong is None:
            self._check(self._L.doug_set_strong_mask(self.h, 0, None))
            return
        m = np.ascontiguousarray(strong, np.uint8)
        self._check(self._L.doug_set_strong_mask(self.h, len(m), _lib.ptr(m)))

    def set_restriction(self, nagr, ri, rj, rv, base=1):
        """doug_set_restriction_coo: R (nagr x n) as COO triples; nagr 0
        restores the smoothed aggregation's R."""
        if nagr <= 0:
            self._check(self._L.doug_set_restriction_coo(self.h, 0, 0, None, None, None, 1))
            return
        i = np.ascontiguousarray(ri, np.int32)
        j = np.ascontiguousarray(rj, np.int32)
        v = np.ascontiguousarray(rv, np.float64)
        self._check(self._L.doug_set_restriction_coo(self.h, int(nagr), len(v), _lib.ptr(i),
                                                     _lib.ptr(j), _lib.ptr(v), int(base)))

    def set_subdomains(self, starts, inds, base=1):
        """doug_set_subdomains: final index sets (overlap included)."""
        st = np.ascontiguousarray(starts, np.int32)
        ix = np.ascontiguousarray(inds, np.int32)
        self._check(self._L.doug_set_subdomains(self.h, len(st) - 1, _lib.ptr(st), _lib.ptr(ix),
                                                int(base)))

    def set_ranks(self, nranks: int, rank_of_node=None, base=1):
        """doug_set_ranks: the decomposition of a DOUG run on `nranks`
        processes (rank_of_node: owning rank of each unknown, or None for the
        project partitioner on the matrix graph); 0 restores one process."""
        if rank_of_node is None:
            self._check(self._L.doug_set_ranks(self.h, int(nranks), None, int(base)))
        else:
            r = np.ascontiguousarray(rank_of_node, np.int32)
            if len(r) != self.n:
                raise ValueError("rank_of_node needs one rank per unknown")
            self._check(self._L.doug_set_ranks(self.h, int(nranks), _lib.ptr(r), int(base)))

    def rank_map(self, rank: int):
        """(ninner, lg): a virtual rank's local numbering, 1-based global ids."""
        ni, nl = C.c_int32(0), C.c_int32(0)
        self._check(self._L.doug_get_rank_map(self.h, int(rank), C.byref(ni), C.byref(nl), None))
        lg = np.zeros(max(nl.value, 1), np.int32)
        self._check(self._L.doug_get_rank_map(self.h, int(rank), C.byref(ni), C.byref(nl),
                                              _lib.ptr(lg)))
        return ni.value, lg[:nl.value].copy()

    def setup(self, precision: int = 0) -> dict:
        self._check(self._L.doug_setup(self.h, int(precision)))
        return self.info()

    def setup_host(self) -> dict:
        """Aggregates, coarse space and subdomain index sets only (no GPU)."""
        self._check(self._L.doug_setup_host(self.h))
        return self.info()

    def info(self) -> dict:
        inf = _lib.DougInfo()
        self._check(self._L.doug_get_info(self.h, C.byref(inf)))
        return inf.as_dict()

    # -- setup products (1-based, as DOUG stores them) --
    def fine_aggregates(self):
        inf = self.info()
        num = np.zeros(inf["n"], np.int32)
        st = np.zeros(inf["nagr_fine"] + 1, np.int32)
        nodes = np.zeros(inf["n"], np.int32)
        self._check(self._L.doug_get_fine_aggregates(self.h, _lib.ptr(num), _lib.ptr(st), _lib.ptr(nodes)))
        return num, st, nodes[: st[-1] - 1].copy()

    def coarse_aggregates(self):
        inf = self.info()
        nc = inf["nagr_fine"]
        num = np.zeros(nc, np.int32)
        st = np.zeros(inf["nagr_coarse"] + 1, np.int32)
        nodes = np.zeros(max(nc, 1), np.int32)
        self._check(self._L.doug_get_coarse_aggregates(self.h, _lib.ptr(num), _lib.ptr(st), _lib.ptr(nodes)))
        return num, st, nodes[: st[-1] - 1].copy()

    def subdomains(self):
        inf = self.info()
        st = np.zeros(inf["nsub"] + 1, np.int32)
        inds = np.zeros(max(inf["sub_total"], 1), np.int32)
        self._check(self._L.doug_get_subdomains(self.h, _lib.ptr(st), _lib.ptr(inds)))
        return st, inds[: inf["sub_total"]].copy()

    def _csr(self, nnz_fn, get_fn, nrows):
        nnz = int(nnz_fn(self.h))
        rp = np.zeros(nrows + 1, np.int32)
        cols = np.zeros(max(nnz, 1), np.int32)
        vals = np.zeros(max(nnz, 1), np.float64)
        self._check(get_fn(self.h, _lib.ptr(rp), _lib.ptr(cols), _lib.ptr(vals)))
        return rp, cols[:nnz].copy(), vals[:nnz].copy()

    def restriction(self):
        """R as 0-based CSR (rows = aggregates, columns ascending)."""
        return self._csr(self._L.doug_restriction_nnz, self._L.doug_get_restriction,
                         self.info()["ncoarse"])

    def coarse_matrix(self):
        """A_c as 0-based CSR in DOUG's first-touch entry order."""
        return self._csr(self._L.doug_coarse_matrix_nnz, self._L.doug_get_coarse_matrix,
                         self.info()["ncoarse"])

    # -- operators (host arrays) --
    def spmv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n)
        self._check(self._L.doug_spmv(self.h, _lib.ptr(x), _lib.ptr(y)))
        return y

    def prec(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.n)
        self._check(self._L.doug_prec_apply(self.h, _lib.ptr(r), _lib.ptr(z)))
        return z

    def pcg(self, b, tol=1e-8, maxit=-1, x0=None):
        """pcg_weigs: returns (x, iterations, condition estimate, residual norms)."""
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64)
        cap = maxit if maxit > 0 else 100000
        hist = np.zeros(cap)
        it = C.c_int32(0)
        cond = C.c_double(0.0)
        self._check(self._L.doug_pcg(self.h, _lib.ptr(b), _lib.ptr(x), float(tol), int(maxit),
                                     C.byref(it), C.byref(cond), _lib.ptr(hist)))
        return x, it.value, cond.value, hist[: it.value].copy()

    # -- operators (device pointers, e.g. torch.Tensor.data_ptr()) --
    def spmv_device(self, dx: int, dy: int):
        _fence()
        self._check(self._L.doug_spmv_device(self.h, C.c_void_p(dx), C.c_void_p(dy)))

    def prec_device(self, dr: int, dz: int):
        _fence()
        self._check(self._L.doug_prec_apply_device(self.h, C.c_void_p(dr), C.c_void_p(dz)))

    def prec_device_noalloc(self, dr: int, dz: int):
        """doug_prec_apply_device_noalloc: the apply with its work vectors in
        the library's static per-GPU scratch array (no device allocation)."""
        _fence()
        self._check(self._L.doug_prec_apply_device_noalloc(self.h, C.c_void_p(dr), C.c_void_p(dz)))

    def pcg_device(self, db: int, dx: int, tol=1e-8, maxit=-1):
        _fence()
        cap = maxit if maxit > 0 else 100000
        hist = np.zeros(cap)
        it = C.c_int32(0)
        cond = C.c_double(0.0)
        self._check(self._L.doug_pcg_device(self.h, C.c_void_p(db), C.c_void_p(dx), float(tol),
                                            int(maxit), C.byref(it), C.byref(cond), _lib.ptr(hist)))
        return it.value, cond.value, hist[: it.value].copy()

    def set_option(self, option: int, value: int):
        """doug_set_option; option 1 = replay PCG iterations from a HIP graph."""
        self._check(self._L.doug_set_option(self.h, int(option), int(value)))

    def value_grad_device(self, dlam: int, dx: int, dgrad: int):
        """dgrad[e] = -lam[i_e] * x[j_e] for every input entry e (device pointers)."""
        _fence()
        self._check(self._L.doug_value_grad_device(self.h, C.c_void_p(dlam), C.c_void_p(dx),
                                                   C.c_void_p(dgrad)))

    # -- several right-hand sides --
    def prepare_batch(self):
        """Explicit inverse subdomain factors for batched applies (doug_prepare_batch)."""
        self._check(self._L.doug_prepare_batch(self.h))

    def pcg_batch(self, B, tol=1e-8, maxit=-1, X0=None):
        """Solve for every column of B (n x nrhs, nrhs <= 32) in one pass.
        Returns (X, iterations per column, residual norms (iterations x nrhs))."""
        B = np.ascontiguousarray(B, np.float64)
        if B.ndim == 1:
            B = B[:, None]
        nrhs = B.shape[1]
        X = np.zeros_like(B) if X0 is None else np.array(X0, np.float64, order="C")
        cap = maxit if maxit > 0 else 100000
        hist = np.zeros((cap, nrhs))
        its = np.zeros(nrhs, np.int32)
        self._check(self._L.doug_pcg_batch(self.h, nrhs, _lib.ptr(B), _lib.ptr(X), float(tol),
                                           int(maxit), _lib.ptr(its), _lib.ptr(hist)))
        return X, its, hist[: int(its.max(initial=0))]

    def pcg_batch_device(self, nrhs: int, dB: int, dX: int, tol=1e-8, maxit=-1):
        _fence()
        its = np.zeros(nrhs, np.int32)
        self._check(self._L.doug_pcg_batch_device(self.h, int(nrhs), C.c_void_p(dB), C.c_void_p(dX),
                                                  float(tol), int(maxit), _lib.ptr(its), None))
        return its

    def prec_batch_device(self, nrhs: int, dR: int, dZ: int):
        _fence()
        self._check(self._L.doug_prec_apply_batch_device(self.h, int(nrhs), C.c_void_p(dR), C.c_void_p(dZ)))

    def autotune(self, reps: int = 20) -> dict:
        """doug_autotune: time every SpMV launch shape and subdomain dispatch
        order, keep the fastest; returns the measurements and the choice."""
        tu = _lib.DougTuning()
        self._check(self._L.doug_autotune(self.h, int(reps), C.byref(tu)))
        return tu.as_dict()

    def tuning(self) -> dict:
        tu = _lib.DougTuning()
        self._check(self._L.doug_get_tuning(self.h, C.byref(tu)))
        return tu.as_dict()

    def time_operator(self, which: int, reps: int = 10) -> float:
        ms = C.c_double(0.0)
        self._check(self._L.doug_time_operator(self.h, int(which), int(reps), C.byref(ms)))
        return ms.value


def partition_graph(xadj, adjncy, nparts: int, vwgt=None) -> np.ndarray:
    """doug_partition_graph on a 0-based symmetric adjacency: part number per vertex."""
    L = _lib.load()
    xadj = np.ascontiguousarray(xadj, np.int32)
    adjncy = np.ascontiguousarray(adjncy, np.int32)
    n = len(xadj) - 1
    w = None if vwgt is None else np.ascontiguousarray(vwgt, np.int32)
    part = np.empty(n, np.int32)
    rc = L.doug_partition_graph(n, _lib.ptr(xadj), _lib.ptr(adjncy),
                                None if w is None else _lib.ptr(w), int(nparts), 0, _lib.ptr(part))
    if rc:
        raise DougError(rc, "doug_partition_graph failed")
    return part
