"""ctypes bindings of the CPU oracle (test infrastructure only).

Loaded by tests/, by the smoke check and by the CPU-baseline leg of bench.py
-- never by the product.  Builds oracle/liboracle.so with `make` on first use
if it is missing."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
# the same source with OpenMP: only the timed CPU baseline loads it
LIB_MT = os.path.join(HERE, "liboracle_mt.so")

_libs = {}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib(threaded=False):
    path = LIB_MT if threaded else LIB
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        _declare(L)
        _libs[path] = L
    return _libs[path]


class Opts(C.Structure):
    _fields_ = [("find_disconnected_comps", C.c_int), ("rsb_pre", C.c_int),
                ("max_iter", C.c_int), ("max_passes", C.c_int), ("tol", C.c_double),
                ("nparts", C.c_int), ("rsb_algo", C.c_int), ("mg_factor", C.c_int),
                ("mg_grammian", C.c_int)]


ML = 32


class Report(C.Structure):
    _fields_ = [("nlevels", C.c_int),
                ("niter_min", C.c_int * ML), ("niter_max", C.c_int * ML),
                ("ncomp_min", C.c_int * ML), ("ncomp_max", C.c_int * ML),
                ("nbrs_max", C.c_int * ML), ("nsegs", C.c_int * ML),
                ("nbalanced", C.c_int * ML),
                ("lambda_min", C.c_double * ML), ("lambda_max", C.c_double * ML),
                ("nbrs_min", C.c_int * ML), ("nbrs_mean", C.c_double * ML)]

    def as_dict(self):
        n = self.nlevels
        return {f: list(getattr(self, f))[:n] for f, _ in self._fields_ if f != "nlevels"}


P = C.c_void_p


def _declare(L):
    L.oracle_centroids.argtypes = [P, C.c_int, C.c_int, P]
    L.oracle_rib_order.argtypes = [P, C.c_int, P, C.c_int, C.c_int]
    L.oracle_conn_mesh.argtypes = [P, P, C.c_int, C.c_int, C.c_double]
    L.oracle_conn_mesh_periodic.argtypes = [P, P, C.c_int, C.c_int, C.c_double, P, C.c_int]
    L.oracle_rib_axis.argtypes = [P, C.c_int, P, C.c_int, P, P]
    L.oracle_sym_dominant.argtypes = [P, C.c_int, P]
    L.oracle_rank_count.argtypes = [C.c_int64, C.c_int, C.c_int]
    L.oracle_rank_count.restype = C.c_int64
    L.oracle_lower_count.argtypes = [C.c_int64, C.c_int]
    L.oracle_lower_count.restype = C.c_int64
    L.oracle_balance_expected.argtypes = [C.c_int64, C.c_int, C.c_int]
    L.oracle_balance_expected.restype = C.c_int64
    L.oracle_rcb_order.argtypes = [P, C.c_int, P, C.c_int, C.c_int]
    L.oracle_dual_graph.argtypes = [P, C.c_int, C.c_int, C.POINTER(P), C.POINTER(P),
                                    C.POINTER(P)]
    L.oracle_dual_graph.restype = C.c_int64
    L.oracle_free.argtypes = [P]
    L.oracle_set_verbose.argtypes = [C.c_int]
    L.oracle_threads.argtypes = [C.c_int]
    L.oracle_threads.restype = C.c_int
    L.oracle_laplacian_apply.argtypes = [P, C.c_int, P, C.c_int, P, P]
    L.oracle_tqli.argtypes = [P, P, C.c_int, P, P]
    L.oracle_tqli.restype = C.c_int
    L.oracle_fiedler.argtypes = [P, C.c_int, P, C.c_int, C.c_int, C.c_int, C.c_double, P,
                                 P, P]
    L.oracle_fiedler.restype = C.c_int
    L.oracle_fiedler_from.argtypes = [P, C.c_int, P, C.c_int, C.c_int, C.c_int,
                                      C.c_double, P, P, P, P]
    L.oracle_fiedler_from.restype = C.c_int
    L.oracle_lanczos_pass.argtypes = [P, C.c_int, P, C.c_int, C.c_int, C.c_double, P, P,
                                      P, P]
    L.oracle_lanczos_pass.restype = C.c_int
    L.oracle_components.argtypes = [P, C.c_int, P, C.c_int]
    L.oracle_fiedler_mg.argtypes = [P, C.c_int, P, C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_int, C.c_int, P, P]
    L.oracle_sym_smallest.argtypes = [P, C.c_int, P]
    L.oracle_sym_smallest.restype = C.c_double
    L.oracle_fiedler_mg.restype = C.c_int
    L.oracle_mg_levels.argtypes = [P, C.c_int, P, C.c_int, C.c_int, P, P]
    L.oracle_mg_levels.restype = C.c_int
    L.oracle_mg_level_matrix.argtypes = [P, C.c_int, P, C.c_int, C.c_int, C.c_int, P, P, P, P]
    L.oracle_mg_level_matrix.restype = C.c_int
    L.oracle_mg_vcycle.argtypes = [P, C.c_int, P, C.c_int, C.c_int, P, P]
    L.oracle_mg_project.argtypes = [P, C.c_int, P, C.c_int, C.c_int, C.c_int, C.c_double, P, P]
    L.oracle_mg_project.restype = C.c_int
    L.oracle_components.restype = C.c_int
    L.oracle_balance_step.argtypes = [P, C.c_int, P, P, C.c_int, C.c_int, C.c_int, C.c_int64]
    L.oracle_balance_step.restype = C.c_int
    L.oracle_part_mesh.argtypes = [P, P, P, C.c_int, C.c_int, C.POINTER(Opts),
                                   C.POINTER(Report)]
    L.oracle_part_mesh.restype = C.c_int
    L.oracle_part_mesh_tagged.argtypes = [P, P, P, P, C.c_int, C.c_int, C.POINTER(Opts)]
    L.oracle_part_mesh_tagged.restype = C.c_int
    L.oracle_quality.argtypes = [P, P, C.c_int, C.c_int, C.c_int, P]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def centroids(xyz, nv):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nel = xyz.shape[0]
    out = np.zeros((nel, 3), dtype=np.float64)
    lib().oracle_centroids(_p(xyz), nel, nv, _p(out))
    return out


def rank_count(n, p, r):
    return lib().oracle_rank_count(n, p, r)


def lower_count(n, p):
    return lib().oracle_lower_count(n, p)


def balance_expected(n, p, b):
    return lib().oracle_balance_expected(n, p, b)


def rcb_order(coord, ndim, order, p):
    coord = np.ascontiguousarray(coord, dtype=np.float64)
    order = np.array(order, dtype=np.int32, copy=True)
    lib().oracle_rcb_order(_p(coord), ndim, _p(order), order.size, p)
    return order


def rib_order(coord, ndim, order, p):
    coord = np.ascontiguousarray(coord, dtype=np.float64)
    order = np.array(order, dtype=np.int32, copy=True)
    lib().oracle_rib_order(_p(coord), ndim, _p(order), order.size, p)
    return order


def rib_axis(coord, ndim, order):
    coord = np.ascontiguousarray(coord, dtype=np.float64)
    order = np.ascontiguousarray(order, dtype=np.int32)
    m, ev = np.zeros(3), np.zeros(3)
    lib().oracle_rib_axis(_p(coord), ndim, _p(order), order.size, _p(m), _p(ev))
    return m, ev


def sym_dominant(A, n):
    A3 = np.zeros((3, 3))
    A3[:n, :n] = A
    ev = np.zeros(3)
    lib().oracle_sym_dominant(_p(np.ascontiguousarray(A3)), n, _p(ev))
    return ev[:n]


def conn_mesh(xyz, ndim, tol=0.2, periodic=None):
    """(rc, vtx[nel, nv]) from per-corner coordinates xyz[nel, nv, ndim];
    periodic = int64 records (element, face, element, face), 1-based."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nv = 8 if ndim == 3 else 4
    nel = xyz.size // (nv * ndim)
    vtx = np.zeros((nel, nv), dtype=np.int64)
    if periodic is None or len(periodic) == 0:
        rc = lib().oracle_conn_mesh(_p(vtx), _p(xyz), nel, ndim, tol)
    else:
        pi = np.ascontiguousarray(periodic, dtype=np.int64).reshape(-1, 4)
        rc = lib().oracle_conn_mesh_periodic(_p(vtx), _p(xyz), nel, ndim, tol, _p(pi),
                                             pi.shape[0])
    return rc, vtx


def dual_graph(vtx, nv):
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    nel = vtx.size // nv
    rp, cp, wp = P(), P(), P()
    nnz = lib().oracle_dual_graph(_p(vtx), nel, nv, C.byref(rp), C.byref(cp), C.byref(wp))
    rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int64)), (nel + 1,)).copy()
    col = np.ctypeslib.as_array(C.cast(cp, C.POINTER(C.c_int32)), (max(nnz, 1),))[:nnz].copy()
    w = np.ctypeslib.as_array(C.cast(wp, C.POINTER(C.c_int32)), (max(nnz, 1),))[:nnz].copy()
    for x in (rp, cp, wp):
        lib().oracle_free(x)
    return rowptr, col, w


def laplacian_apply(vtx, nv, elems, u):
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.zeros_like(u)
    lib().oracle_laplacian_apply(_p(vtx), nv, _p(elems), elems.size, _p(u), _p(v))
    return v


def tqli(diag, upper):
    d = np.ascontiguousarray(diag, dtype=np.float64)
    e = np.ascontiguousarray(upper, dtype=np.float64)
    n = d.size
    if e.size == 0:
        e = np.zeros(1)
    z = np.zeros((n, n), dtype=np.float64)
    ev = np.zeros(n, dtype=np.float64)
    rc = lib().oracle_tqli(_p(d), _p(e), n, _p(z), _p(ev))
    return rc, ev, z


def fiedler(vtx, nv, elems, max_iter=50, max_passes=50, tol=1e-5, init=None):
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    n = elems.size
    f = np.zeros(n, dtype=np.float64)
    lam = C.c_double()
    rn = np.zeros(3, dtype=np.float64)
    if init is None:
        it = lib().oracle_fiedler(_p(vtx), nv, _p(elems), n, max_iter, max_passes, tol,
                                  _p(f), C.byref(lam), _p(rn))
    else:
        init = np.ascontiguousarray(init, dtype=np.float64)
        it = lib().oracle_fiedler_from(_p(vtx), nv, _p(elems), n, max_iter, max_passes,
                                       tol, _p(init), _p(f), C.byref(lam), _p(rn))
    return f, it, lam.value, rn


def lanczos_pass(vtx, nv, elems, f, niter, tol=1e-5):
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = elems.size
    d = np.zeros(niter)
    e = np.zeros(max(niter, 1))
    basis = np.zeros((niter + 1, n))
    k = lib().oracle_lanczos_pass(_p(vtx), nv, _p(elems), n, niter, tol, _p(f), _p(d),
                                  _p(e), _p(basis))
    return k, d[:k], e[:max(k - 1, 0)], basis[:k + 1]


def components(vtx, nv, elems):
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    return lib().oracle_components(_p(vtx), nv, _p(elems), elems.size)


def balance_step(vtx, nv, elems, fied, nl, sender_lower, scount):
    """One balance step (rsb.c:118-229) on a segment in ascending Fiedler
    order: returns (new element order, permuted Fiedler values, new lower
    size)."""
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    el = np.ascontiguousarray(elems, dtype=np.int32).copy()
    fv = np.ascontiguousarray(fied, dtype=np.float64).copy()
    nl2 = lib().oracle_balance_step(_p(vtx), nv, _p(el), _p(fv), el.size, int(nl),
                                    int(sender_lower), int(scount))
    return el, fv, nl2


def sym_smallest(A):
    """Smallest eigenpair (lambda, v) of a symmetric matrix by cyclic Jacobi."""
    A = np.array(A, dtype=np.float64, order="C")
    n = A.shape[0]
    v = np.zeros(n)
    lam = lib().oracle_sym_smallest(_p(A), n, _p(v))
    return lam, v


def fiedler_mg(vtx, nv, elems, max_iter=50, max_passes=50, tol=1e-5, factor=2, grammian=0):
    """(fiedler, iterations, passes) of the MG inverse iteration."""
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    f = np.zeros(elems.size)
    npass = C.c_int()
    it = lib().oracle_fiedler_mg(_p(vtx), nv, _p(elems), elems.size, max_iter, max_passes, tol,
                                 factor, grammian, _p(f), C.byref(npass))
    return f, it, npass.value


def mg_levels(vtx, nv, elems, factor=2):
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    sizes = np.zeros(64, dtype=np.int32)
    nnz = np.zeros(64, dtype=np.int64)
    nl = lib().oracle_mg_levels(_p(vtx), nv, _p(elems), elems.size, factor, _p(sizes), _p(nnz))
    return sizes[:nl], nnz[:nl]


def mg_level_matrix(vtx, nv, elems, level, factor=2):
    """(rowptr, col, w, diag) of one MG level (positive off-diagonal weights)."""
    sizes, nnz = mg_levels(vtx, nv, elems, factor)
    n, m = int(sizes[level]), int(nnz[level])
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    rp = np.zeros(n + 1, dtype=np.int64)
    col = np.zeros(max(m, 1), dtype=np.int32)
    w = np.zeros(max(m, 1))
    d = np.zeros(max(n, 1))
    lib().oracle_mg_level_matrix(_p(vtx), nv, _p(elems), elems.size, factor, level, _p(rp),
                                 _p(col), _p(w), _p(d))
    return rp, col[:m], w[:m], d[:n]


def mg_vcycle(vtx, nv, elems, rhs, factor=2):
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    out = np.zeros_like(rhs)
    lib().oracle_mg_vcycle(_p(vtx), nv, _p(elems), elems.size, factor, _p(rhs), _p(out))
    return out


def mg_project(vtx, nv, elems, b, miter=50, tol=1e-5, factor=2):
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    it = lib().oracle_mg_project(_p(vtx), nv, _p(elems), elems.size, factor, miter, tol, _p(b),
                                 _p(x))
    return x, it


def part_mesh(vtx, xyz, nv, nparts, rsb_pre=1, find_comps=1, max_iter=50,
              max_passes=50, tol=1e-5, rsb_algo=0, mg_factor=2, threads=0, verbose=0,
              mg_grammian=0):
    """The reference algorithm over virtual ranks.  threads > 0 runs the
    OpenMP build on that many threads (the CPU baseline); its sums depend on
    the thread count, so parity checks keep threads = 0 (serial build)."""
    L = lib(threaded=threads > 0)
    if threads > 0:
        L.oracle_threads(int(threads))
    L.oracle_set_verbose(int(verbose))
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    nel = vtx.size // nv
    xp = None
    if xyz is not None:
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        xp = _p(xyz)
    part = np.full(nel, -1, dtype=np.int32)
    o = Opts(find_comps, rsb_pre, max_iter, max_passes, tol, nparts, rsb_algo, mg_factor,
             mg_grammian)
    rep = Report()
    rc = L.oracle_part_mesh(_p(part), _p(vtx), xp, nel, nv, C.byref(o), C.byref(rep))
    if rc != 0:
        raise RuntimeError(f"oracle_part_mesh returned {rc}")
    return part, rep.as_dict()


def part_mesh_tagged(vtx, xyz, nv, nparts, tag, rsb_pre=1, find_comps=1, max_iter=50,
                     max_passes=50, tol=1e-5, rsb_algo=0, mg_factor=2):
    """Tagged partitioning; returns part (raises on rc != 0)."""
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    nel = vtx.size // nv
    xp = None
    if xyz is not None:
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        xp = _p(xyz)
    tag = np.ascontiguousarray(tag, dtype=np.int32)
    part = np.full(nel, -1, dtype=np.int32)
    o = Opts(find_comps, rsb_pre, max_iter, max_passes, tol, nparts, rsb_algo, mg_factor)
    rc = lib().oracle_part_mesh_tagged(_p(part), _p(vtx), xp, _p(tag), nel, nv, C.byref(o))
    if rc != 0:
        raise RuntimeError(f"oracle_part_mesh_tagged returned {rc}")
    return part


QUALITY_KEYS = ("edge_cut", "cut_edges", "max_part", "min_part", "components",
                "max_components", "comm_volume", "empty_parts")


def quality(part, vtx, nv, nparts):
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    part = np.ascontiguousarray(part, dtype=np.int32)
    out = np.zeros(8, dtype=np.int64)
    lib().oracle_quality(_p(part), _p(vtx), part.size, nv, nparts, _p(out))
    return dict(zip(QUALITY_KEYS, (int(x) for x in out)))
