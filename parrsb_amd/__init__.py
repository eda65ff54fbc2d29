"""parrsb_amd -- recursive spectral bisection on one MI355X.

Python view of the C ABI in include/ (parRSB.h, parrsb_amd.h, rsbgpu.h),
loaded from the in-tree libparrsb_amd.so.  There is no fallback: without the
built library, or without a GPU, every call raises.

    part = parrsb_amd.part_mesh(vtx, xyz, nv, nparts)     # parrsb_part_mesh
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libparrsb_amd.so")

_lib = None


class ParrsbError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"parrsb returned {rc}: {msg}")
        self.rc = rc


class Options(C.Structure):
    """parrsb_options (parRSB.h) with the appended `nparts` field."""
    _fields_ = [("partitioner", C.c_int), ("tagged", C.c_int), ("levels", C.c_int),
                ("find_disconnected_comps", C.c_int), ("repair", C.c_int),
                ("verbose_level", C.c_int), ("profile_level", C.c_int),
                ("rsb_algo", C.c_int), ("rsb_pre", C.c_int), ("rsb_max_iter", C.c_int),
                ("rsb_max_passes", C.c_int), ("rsb_tol", C.c_double),
                ("rsb_dump_stats", C.c_int), ("rsb_mg_grammian", C.c_int),
                ("rsb_mg_factor", C.c_int), ("nparts", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


ML = 32


class Report(C.Structure):
    _fields_ = [("nlevels", C.c_int), ("nparts", C.c_int), ("nel", C.c_int),
                ("precision", C.c_int), ("seconds", C.c_double),
                ("phase_ms", C.c_double * 8), ("lanczos_bytes", C.c_double),
                ("lanczos_ms", C.c_double), ("lanczos_iters", C.c_double),
                ("spmv_bytes", C.c_double), ("spmv_ms", C.c_double),
                ("spmv_launches", C.c_double), ("graph_nnz", C.c_int64),
                ("graph_wsum", C.c_int64), ("nsegs", C.c_int * ML),
                ("niter_min", C.c_int * ML), ("niter_max", C.c_int * ML),
                ("npass_max", C.c_int * ML), ("ncomp_min", C.c_int * ML),
                ("ncomp_max", C.c_int * ML), ("nbalanced", C.c_int * ML),
                ("nconverged_fail", C.c_int * ML), ("lambda_min", C.c_double * ML),
                ("lambda_max", C.c_double * ML), ("resid_max", C.c_double * ML),
                ("resid_med", C.c_double * ML), ("niter_mean", C.c_double * ML),
                ("level_ms", C.c_double * ML), ("nfallback", C.c_int * ML),
                ("warm_start", C.c_int), ("nbrs_min", C.c_int * ML),
                ("nbrs_max", C.c_int * ML), ("nbrs_mean", C.c_double * ML),
                ("tol_init", C.c_double * ML), ("tol_target", C.c_double * ML),
                ("tol_final", C.c_double * ML), ("noalloc_bytes", C.c_double),
                ("noalloc_high", C.c_double)]

    def as_dict(self):
        d = {}
        n = max(0, min(self.nlevels, ML))
        for f, t in self._fields_:
            v = getattr(self, f)
            if f == "phase_ms":
                d[f] = dict(zip(("setup", "prepartition", "laplacian", "lanczos", "tqli_ritz",
                                 "sort", "components", "balance"), list(v)))
            elif hasattr(v, "__len__"):
                d[f] = list(v)[:n]
            else:
                d[f] = v
        return d


class FiedlerOut(C.Structure):
    _fields_ = [("niter", C.c_void_p), ("npass", C.c_void_p), ("lam", C.c_void_p),
                ("rnorm", C.c_void_p), ("resid", C.c_void_p), ("fallback", C.c_void_p)]


class CmdOpts(C.Structure):
    """parrsb_cmd_line_opts (parRSB.h)."""
    _fields_ = [("mesh", C.c_char_p), ("tol", C.c_double), ("test", C.c_int),
                ("dump", C.c_int), ("nactive", C.c_int), ("verbose", C.c_int)]


P = C.c_void_p
I = C.c_int

# name: (restype, argtypes) -- every function declared in include/*.h
SIGNATURES = {
    # parRSB.h
    "parrsb_part_mesh": (I, [P, P, P, P, I, I, C.POINTER(Options), I]),
    # parrsb_amd.h
    "parrsb_amd_part_mesh_dev": (I, [P, P, P, I, I, C.POINTER(Options)]),
    "parrsb_amd_last_report": (I, [C.POINTER(Report)]),
    "parrsb_amd_last_error": (C.c_char_p, []),
    "parrsb_amd_set_precision": (I, [I]),
    "parrsb_amd_set_profiling": (I, [I]),
    "parrsb_amd_warm_start_nparams": (I, []),
    "parrsb_part_mesh_seq": (I, [P, P, P, P, P, I, I, C.POINTER(Options), I]),
    "parrsb_conn_mesh": (I, [P, P, C.c_uint, C.c_uint, P, I, C.c_double, I]),
    "fparrsb_conn_mesh_": (None, [P, P, P, P, P, P, P, P, P]),
    "fparrsb_conn_mesh": (None, [P, P, P, P, P, P, P, P, P]),
    "parrsb_amd_conn_mesh_dev": (I, [P, P, I, I, C.c_double]),
    "parrsb_amd_set_warm_start": (I, [P, I, I]),
    "parrsb_amd_partition_quality": (I, [P, P, I, I, I, P]),
    "parrsb_amd_partition_quality_dev": (I, [P, P, I, I, I, P]),
    "parrsb_read_mesh": (I, [C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(P),
                             C.POINTER(P), C.POINTER(C.c_uint), C.POINTER(P), C.c_char_p, I,
                             I]),
    "parrsb_dump_con": (I, [C.c_char_p, C.c_uint, C.c_uint, P, I]),
    "parrsb_dump_map": (I, [C.c_char_p, C.c_uint, C.c_uint, P, I]),
    "parrsb_amd_dump_map_parts": (I, [C.c_char_p, C.c_uint, C.c_uint, P, P]),
    "parrsb_dist_mesh": (I, [C.POINTER(C.c_uint), C.POINTER(P), C.POINTER(P), P, C.c_uint,
                             I]),
    "parrsb_setup_mesh": (I, [C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(P),
                              C.POINTER(P), C.POINTER(CmdOpts), I]),
    "parrsb_parse_cmd_opts": (C.POINTER(CmdOpts), [I, C.POINTER(C.c_char_p)]),
    "parrsb_cmd_opts_free": (None, [C.POINTER(CmdOpts)]),
    "parrsb_amd_free": (None, [P]),
    # rsbgpu.h
    "rsbgpu_open": (I, [C.POINTER(P), I]),
    "rsbgpu_close": (None, [P]),
    "rsbgpu_error": (C.c_char_p, [P]),
    "rsbgpu_phase_times": (I, [P, P]),
    "rsbgpu_lanczos_traffic": (I, [P, P]),
    "rsbgpu_set_profiling": (I, [P, I]),
    "rsbgpu_test_tqli": (I, [P, P, I, P, P]),
    "rsbgpu_stream_test": (I, [I, C.c_size_t, I, P]),
    "rsbgpu_level_e12": (I, [P]),
    "rsbgpu_test_eigmin": (I, [P, P, I, P, P]),
    "rsbgpu_set_mesh": (I, [P, P, P, I, I, I]),
    "rsbgpu_graph_stats": (I, [P, P, P]),
    "rsbgpu_set_segments": (I, [P, I, P, P, P, P]),
    "rsbgpu_prepartition": (I, [P, I]),
    "rsbgpu_level_laplacian": (I, [P]),
    "rsbgpu_fiedler": (I, [P, I, I, I, C.c_double, I, C.POINTER(FiedlerOut)]),
    "rsbgpu_set_start": (I, [P, P]),
    "rsbgpu_fiedler_mg": (I, [P, I, I, C.c_double, I, I, C.POINTER(FiedlerOut)]),
    "rsbgpu_test_sym_smallest": (I, [P, I, P, P]),
    "rsbgpu_mg_level_csr": (I, [P, I, I, C.POINTER(I), C.POINTER(C.c_int64), P, P, P, P]),
    "rsbgpu_fiedler_sort": (I, [P]),
    "rsbgpu_bisect": (I, [P, P, P, P, I, P, P]),
    "rsbgpu_write_part": (I, [P, P, I]),
    "rsbgpu_get_order": (I, [P, P]),
    "rsbgpu_get_fiedler": (I, [P, P]),
    "rsbgpu_set_fiedler": (I, [P, P]),
    "rsbgpu_level_form": (I, [P, P, P]),
    "rsbgpu_noalloc_begin": (I, [I, P, I]),
    "rsbgpu_noalloc_end": (C.c_longlong, [P]),
    "parrsb_amd_part_mesh_dev_noalloc": (I, [P, P, P, I, I, C.POINTER(Options)]),
    "rsbgpu_get_level_csr": (I, [P, P, P, P, P, P, P]),
    "rsbgpu_get_graph": (I, [P, P, P, P]),
    "rsbgpu_get_centroids": (I, [P, P]),
    "rsbgpu_quality": (I, [P, P, I, P]),
    "rsbgpu_upload_ints": (I, [P, P, I, C.POINTER(P)]),
    "rsbgpu_upload": (I, [P, P, C.c_size_t, C.POINTER(P)]),
    "rsbgpu_download": (I, [P, P, C.c_size_t, P]),
    "rsbgpu_neighbors": (I, [P, P]),
    "rsbgpu_write_seq": (I, [P, P, I]),
    "rsbgpu_group_by_tag": (I, [P, P, I, I, P, P]),
    "rsbgpu_conn_mesh": (I, [I, P, P, I, I, C.c_double, P, I, I, C.c_char_p, I]),
    "rsbgpu_lws_nparams": (I, []),
    "rsbgpu_lws_set_params": (I, [P, P, C.c_float, I, I]),
    "rsbgpu_lws_predict": (I, [P, P, C.c_float, P]),
    "rsbgpu_lws_train_step": (I, [P, P, P, C.c_float, C.POINTER(C.c_double)]),
    "rsbgpu_lws_adam": (I, [P, P, P, P, I, C.c_float]),
    "rsbgpu_lws_test": (I, [I, P, C.c_float, P, P, P, C.c_double, P, P, P,
                            C.POINTER(C.c_double)]),
    "rsbgpu_free": (None, [P]),
}


def lib():
    """The loaded C library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run parrsb_amd/build.py "
                              "(or __graft_entry__.build()) first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def default_options(**kw) -> Options:
    o = Options.in_dll(lib(), "parrsb_default_options")
    opts = Options()
    C.memmove(C.byref(opts), C.byref(o), C.sizeof(Options))
    for k, v in kw.items():
        setattr(opts, k, v)
    return opts


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def last_report() -> dict:
    r = Report()
    lib().parrsb_amd_last_report(C.byref(r))
    return r.as_dict()


def last_error() -> str:
    return (lib().parrsb_amd_last_error() or b"").decode()


def set_precision(bits: int):
    rc = lib().parrsb_amd_set_precision(bits)
    if rc:
        raise ParrsbError(rc, "precision must be 32 or 64")


def set_profiling(on: bool):
    lib().parrsb_amd_set_profiling(1 if on else 0)


def part_mesh(vtx, xyz, nv, nparts, tag=None, comm=0, options=None, **kw):
    """parrsb_part_mesh on host arrays; returns (part, options_after_call)."""
    vtx = np.ascontiguousarray(vtx, dtype=np.int64) if vtx is not None else None
    xyz = np.ascontiguousarray(xyz, dtype=np.float64) if xyz is not None else None
    nel = (vtx.size if vtx is not None else xyz.size // (3 if nv == 8 else 2)) // nv
    opts = options if options is not None else default_options(verbose_level=0)
    opts.nparts = nparts
    for k, v in kw.items():
        setattr(opts, k, v)
    part = np.full(nel, -1, dtype=np.int32)
    tg = np.ascontiguousarray(tag, dtype=np.int32) if tag is not None else None
    rc = lib().parrsb_part_mesh(_p(part), _p(vtx), _p(xyz), _p(tg), nel, nv, C.byref(opts),
                                comm)
    if rc:
        raise ParrsbError(rc, last_error())
    return part, opts


def part_mesh_seq(vtx, xyz, nv, nparts, tag=None, comm=0, options=None, **kw):
    """parrsb_part_mesh_seq: (part, seq, options_after_call)."""
    vtx = np.ascontiguousarray(vtx, dtype=np.int64) if vtx is not None else None
    xyz = np.ascontiguousarray(xyz, dtype=np.float64) if xyz is not None else None
    nel = (vtx.size if vtx is not None else xyz.size // (3 if nv == 8 else 2)) // nv
    opts = options if options is not None else default_options(verbose_level=0)
    opts.nparts = nparts
    for k, v in kw.items():
        setattr(opts, k, v)
    part = np.full(nel, -1, dtype=np.int32)
    seq = np.full(nel, -1, dtype=np.int32)
    tg = np.ascontiguousarray(tag, dtype=np.int32) if tag is not None else None
    rc = lib().parrsb_part_mesh_seq(_p(part), _p(seq), _p(vtx), _p(xyz), _p(tg), nel, nv,
                                    C.byref(opts), comm)
    if rc:
        raise ParrsbError(rc, last_error())
    return part, seq, opts


def conn_mesh(xyz, ndim, tol=0.2, comm=0, periodic=None):
    """parrsb_conn_mesh: vtx[nel, nv] (1-based ids, symmetric corner order)
    from per-corner coordinates xyz[nel, nv, ndim] (preprocessor order);
    periodic = records (element, face, partner element, partner face),
    1-based, preprocessor face numbers (as parrsb_read_mesh returns them)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nv = 8 if ndim == 3 else 4
    nel = xyz.size // (nv * ndim)
    vtx = np.zeros((nel, nv), dtype=np.int64)
    pi, npi = None, 0
    if periodic is not None and len(periodic):
        pi = np.ascontiguousarray(periodic, dtype=np.int64).reshape(-1, 4)
        npi = pi.shape[0]
    rc = lib().parrsb_conn_mesh(_p(vtx), _p(xyz), nel, ndim, _p(pi), npi, tol, comm)
    if rc:
        raise ParrsbError(rc, last_error())
    return vtx


def _take(ptr, ctype, n, shape):
    """Copy n items of a library-allocated array into numpy and free it."""
    if not ptr.value:
        return None
    a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), (max(n, 1),))[:n].copy()
    lib().parrsb_amd_free(ptr)
    return a.reshape(shape)


def read_mesh(name, read=1, comm=0):
    """parrsb_read_mesh: dict with nel, nv, and (read & 1) coord[nel, nv, ndim]
    and the periodic records bcs[nbcs, 4], (read & 2) vl[nel, nv]."""
    nel, nv, nb = C.c_uint(), C.c_uint(), C.c_uint()
    vl, coord, bcs = P(), P(), P()
    rc = lib().parrsb_read_mesh(C.byref(nel), C.byref(nv), C.byref(vl) if read & 2 else None,
                                C.byref(coord) if read & 1 else None, C.byref(nb),
                                C.byref(bcs), os.fsencode(name), comm, read)
    if rc:
        for x in (vl, coord, bcs):
            if x.value:
                lib().parrsb_amd_free(x)
        raise ParrsbError(rc, last_error())
    n, v = nel.value, nv.value
    ndim = 3 if v == 8 else 2
    out = {"nel": n, "nv": v}
    if read & 1:
        out["coord"] = _take(coord, C.c_double, n * v * ndim, (n, v, ndim))
        out["bcs"] = _take(bcs, C.c_longlong, 4 * nb.value, (nb.value, 4))
        if out["bcs"] is None:
            out["bcs"] = np.zeros((0, 4), dtype=np.int64)
    if read & 2:
        out["vl"] = _take(vl, C.c_longlong, n * v, (n, v))
    return out


def dump_con(name, vl, comm=0):
    vl = np.ascontiguousarray(vl, dtype=np.int64)
    rc = lib().parrsb_dump_con(os.fsencode(name), vl.shape[0], vl.shape[1], _p(vl), comm)
    if rc:
        raise ParrsbError(rc, last_error())


def dump_map(name, vl, part=None, comm=0):
    """.ma2 writer: parrsb_dump_map (rank field 0), or with part the
    one-rank-holds-all-parts form (rank field = part id)."""
    vl = np.ascontiguousarray(vl, dtype=np.int64)
    if part is None:
        rc = lib().parrsb_dump_map(os.fsencode(name), vl.shape[0], vl.shape[1], _p(vl), comm)
    else:
        part = np.ascontiguousarray(part, dtype=np.int32)
        rc = lib().parrsb_amd_dump_map_parts(os.fsencode(name), vl.shape[0], vl.shape[1],
                                             _p(vl), _p(part))
    if rc:
        raise ParrsbError(rc, last_error())


def dist_mesh(vl, coord, part, comm=0):
    """parrsb_dist_mesh on copies: returns (vl, coord, part) grouped by part."""
    vl = np.array(vl, dtype=np.int64, order="C", copy=True)
    cd = None if coord is None else np.array(coord, dtype=np.float64, order="C", copy=True)
    part = np.array(part, dtype=np.int32, copy=True)
    n = C.c_uint(vl.shape[0])
    pv, pc = P(vl.ctypes.data), P(cd.ctypes.data if cd is not None else None)
    rc = lib().parrsb_dist_mesh(C.byref(n), C.byref(pv), C.byref(pc) if cd is not None else None,
                                _p(part), vl.shape[1], comm)
    if rc:
        raise ParrsbError(rc, last_error())
    return vl, cd, part


def part_mesh_dev(d_part, d_vtx, d_xyz, nel, nv, nparts, options=None, **kw):
    """Device-pointer variant (pointers as ints, e.g. torch tensor data_ptr())."""
    opts = options if options is not None else default_options(verbose_level=0)
    opts.nparts = nparts
    for k, v in kw.items():
        setattr(opts, k, v)
    rc = lib().parrsb_amd_part_mesh_dev(d_part, d_vtx, d_xyz, nel, nv, C.byref(opts))
    if rc:
        raise ParrsbError(rc, last_error())
    return opts


QUALITY_KEYS = ("edge_cut", "cut_edges", "max_part", "min_part", "components",
                "max_components", "comm_volume", "empty_parts")


def quality(part, vtx, nv, nparts) -> dict:
    part = np.ascontiguousarray(part, dtype=np.int32)
    vtx = np.ascontiguousarray(vtx, dtype=np.int64)
    out = np.zeros(8, dtype=np.int64)
    rc = lib().parrsb_amd_partition_quality(_p(part), _p(vtx), part.size, nv, nparts, _p(out))
    if rc:
        raise ParrsbError(rc, last_error())
    return dict(zip(QUALITY_KEYS, (int(x) for x in out)))


def quality_dev(d_part, d_vtx, nel, nv, nparts) -> dict:
    out = np.zeros(8, dtype=np.int64)
    rc = lib().parrsb_amd_partition_quality_dev(d_part, d_vtx, nel, nv, nparts, _p(out))
    if rc:
        raise ParrsbError(rc, last_error())
    return dict(zip(QUALITY_KEYS, (int(x) for x in out)))


class Device:
    """Step-by-step access to the device layer (rsbgpu.h), for parity tests."""

    def __init__(self, device: int = 0):
        self.h = P()
        rc = lib().rsbgpu_open(C.byref(self.h), device)
        if rc:
            raise ParrsbError(rc, "rsbgpu_open failed (no GPU?)")
        self.nel = 0

    def close(self):
        if self.h:
            lib().rsbgpu_close(self.h)
            self.h = P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc:
            raise ParrsbError(rc, (lib().rsbgpu_error(self.h) or b"").decode())

    def set_mesh(self, vtx, xyz, nv):
        self._vtx = np.ascontiguousarray(vtx, dtype=np.int64)
        self._xyz = np.ascontiguousarray(xyz, dtype=np.float64) if xyz is not None else None
        self.nel = self._vtx.size // nv
        self.nv = nv
        self._chk(lib().rsbgpu_set_mesh(self.h, _p(self._vtx), _p(self._xyz), self.nel, nv, 0))

    def graph(self):
        nnz, ws = C.c_int64(), C.c_int64()
        lib().rsbgpu_graph_stats(self.h, C.byref(nnz), C.byref(ws))
        rowptr = np.zeros(self.nel + 1, dtype=np.int64)
        col = np.zeros(max(nnz.value, 1), dtype=np.int32)
        w = np.zeros(max(nnz.value, 1), dtype=np.int32)
        self._chk(lib().rsbgpu_get_graph(self.h, _p(rowptr), _p(col), _p(w)))
        return rowptr, col[:nnz.value], w[:nnz.value]

    def centroids(self):
        c = np.zeros(3 * self.nel)
        self._chk(lib().rsbgpu_get_centroids(self.h, _p(c)))
        return c.reshape(3, self.nel).T.copy()

    def set_segments(self, cstart, n, p, r0):
        a = [np.ascontiguousarray(x, dtype=np.int32) for x in (cstart, n, p, r0)]
        self._chk(lib().rsbgpu_set_segments(self.h, a[0].size, *(_p(x) for x in a)))

    def prepartition(self, rsb_pre=1):
        self._chk(lib().rsbgpu_prepartition(self.h, rsb_pre))

    def level_laplacian(self):
        self._chk(lib().rsbgpu_level_laplacian(self.h))

    def level_csr(self):
        mpad, nnz = C.c_int64(), C.c_int64()
        self._chk(lib().rsbgpu_get_level_csr(self.h, C.byref(mpad), C.byref(nnz), None, None,
                                             None, None))
        rowptr = np.zeros(mpad.value + 1, dtype=np.int64)
        col = np.zeros(max(nnz.value, 1), dtype=np.int32)
        w = np.zeros(max(nnz.value, 1), dtype=np.int32)
        re = np.zeros(max(mpad.value, 1), dtype=np.int32)
        self._chk(lib().rsbgpu_get_level_csr(self.h, C.byref(mpad), C.byref(nnz), _p(rowptr),
                                             _p(col), _p(w), _p(re)))
        return rowptr, col[:nnz.value], w[:nnz.value], re[:mpad.value]

    def fiedler(self, nseg, precision=64, max_iter=50, max_passes=50, tol=1e-5, init_mode=0):
        """init_mode 2 needs lws_set_params first; out["fallback"][s] is 0 if
        the learned start met tol, else 1 + its iterations."""
        out = {k: np.zeros(nseg, dtype=np.int32) for k in ("niter", "npass", "fallback")}
        out["lambda"] = np.zeros(nseg)
        out["rnorm"] = np.zeros(3 * nseg)
        out["resid"] = np.zeros(nseg)
        fo = FiedlerOut(_p(out["niter"]), _p(out["npass"]), _p(out["lambda"]),
                        _p(out["rnorm"]), _p(out["resid"]), _p(out["fallback"]))
        self._chk(lib().rsbgpu_fiedler(self.h, precision, max_iter, max_passes, tol, init_mode,
                                       C.byref(fo)))
        return out

    def fiedler_mg(self, nseg, max_iter=50, max_passes=50, tol=1e-5, factor=2, grammian=0):
        """MG-preconditioned inverse iteration (rsb_algo = 1) on every segment;
        grammian=1 refines each next start vector by a Rayleigh-Ritz step."""
        out = {k: np.zeros(nseg, dtype=np.int32) for k in ("niter", "npass", "fallback")}
        out["lambda"] = np.zeros(nseg)
        out["rnorm"] = np.zeros(3 * nseg)
        out["resid"] = np.zeros(nseg)
        fo = FiedlerOut(_p(out["niter"]), _p(out["npass"]), _p(out["lambda"]),
                        _p(out["rnorm"]), _p(out["resid"]), _p(out["fallback"]))
        self._chk(lib().rsbgpu_fiedler_mg(self.h, max_iter, max_passes, tol, factor, grammian,
                                          C.byref(fo)))
        return out

    def mg_level_csr(self, level, seg):
        n, nnz = C.c_int(), C.c_int64()
        self._chk(lib().rsbgpu_mg_level_csr(self.h, level, seg, C.byref(n), C.byref(nnz), None,
                                            None, None, None))
        rp = np.zeros(n.value + 1, dtype=np.int64)
        col = np.zeros(max(nnz.value, 1), dtype=np.int32)
        w = np.zeros(max(nnz.value, 1))
        d = np.zeros(max(n.value, 1))
        self._chk(lib().rsbgpu_mg_level_csr(self.h, level, seg, C.byref(n), C.byref(nnz),
                                            _p(rp), _p(col), _p(w), _p(d)))
        return rp, col[:nnz.value], w[:nnz.value], d[:n.value]

    # learned warm start (device parameter arrays given as integer pointers,
    # e.g. torch tensor data_ptr())
    def lws_set_params(self, d_params, level, passes=1, max_n=0):
        self._chk(lib().rsbgpu_lws_set_params(self.h, d_params, level, passes, max_n))

    def lws_predict(self, d_params, level, d_rows):
        self._chk(lib().rsbgpu_lws_predict(self.h, d_params, level, d_rows))

    def upload(self, a):
        """Copy a numpy array to a new device buffer; returns its address
        (free with free())."""
        a = np.ascontiguousarray(a)
        d = P()
        self._chk(lib().rsbgpu_upload(self.h, _p(a), a.nbytes, C.byref(d)))
        return d.value

    def download(self, d, like):
        out = np.empty_like(like)
        self._chk(lib().rsbgpu_download(self.h, d, out.nbytes, _p(out)))
        return out

    @staticmethod
    def free(d):
        lib().rsbgpu_free(d)

    def lws_train_step(self, d_params, d_grad, level):
        loss = C.c_double()
        self._chk(lib().rsbgpu_lws_train_step(self.h, d_params, d_grad, level, C.byref(loss)))
        return loss.value

    def group_by_tag(self, tag, max_groups):
        """Stable grouping of the element order by tag; returns per-tag counts
        (tags ascending) or raises when there are more than max_groups tags."""
        tag = np.ascontiguousarray(tag, dtype=np.int32)
        ng = C.c_int()
        counts = np.zeros(max(max_groups, 1), dtype=np.int32)
        self._chk(lib().rsbgpu_group_by_tag(self.h, _p(tag), 0, max_groups, C.byref(ng),
                                            _p(counts)))
        if ng.value > max_groups:
            raise ValueError(f"{ng.value} tags > {max_groups}")
        return counts[:ng.value]

    def neighbors(self, nseg):
        nb = np.zeros(nseg, dtype=np.int32)
        self._chk(lib().rsbgpu_neighbors(self.h, _p(nb)))
        return nb

    def fiedler_sort(self):
        self._chk(lib().rsbgpu_fiedler_sort(self.h))

    def bisect(self, nlower, sender, scount, find_comps=1):
        nl = np.ascontiguousarray(nlower, dtype=np.int32).copy()
        sd = np.ascontiguousarray(sender, dtype=np.int32)
        sc = np.ascontiguousarray(scount, dtype=np.int32)
        ncomp = np.zeros(2 * nl.size, dtype=np.int32)
        moved = np.zeros(nl.size, dtype=np.int32)
        self._chk(lib().rsbgpu_bisect(self.h, _p(nl), _p(sd), _p(sc), find_comps, _p(ncomp),
                                      _p(moved)))
        return nl, ncomp, moved

    def order(self):
        o = np.zeros(self.nel, dtype=np.int32)
        self._chk(lib().rsbgpu_get_order(self.h, _p(o)))
        return o

    def fiedler_values(self):
        f = np.zeros(self.nel)
        self._chk(lib().rsbgpu_get_fiedler(self.h, _p(f)))
        return f

    def lanczos_traffic(self):
        """[iteration bytes (model), iteration ms, iterations, apply bytes
        (model), apply ms (profiling), apply launches (profiling)] over the
        handle's life."""
        v = np.zeros(6)
        self._chk(lib().rsbgpu_lanczos_traffic(self.h, _p(v)))
        return v

    def level_e12(self):
        """True if the level's tiles carry the 12-bit form."""
        return lib().rsbgpu_level_e12(self.h) == 1

    def level_form(self):
        """(tiles on the global path, tiles with the vertex form)."""
        a, b = C.c_int(), C.c_int()
        self._chk(lib().rsbgpu_level_form(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_fiedler_values(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.size == self.nel
        self._chk(lib().rsbgpu_set_fiedler(self.h, _p(f)))

    def write_part(self):
        part = np.zeros(self.nel, dtype=np.int32)
        self._chk(lib().rsbgpu_write_part(self.h, _p(part), 0))
        return part


# Largest segment the learned start is tried on by default (elements): above
# it the reference start needs fewer iterations than the network's
# prediction under the reference's stop rule (at 2048 elements a few percent
# of the segments already fall back).
WARM_MAX_N = 1024


def warm_start_nparams() -> int:
    return lib().parrsb_amd_warm_start_nparams()


def set_warm_start(params, passes=1, max_n=WARM_MAX_N):
    """Learned warm start for later part_mesh calls (None switches it off)."""
    if params is None:
        lib().parrsb_amd_set_warm_start(None, 1, 0)
        return
    params = np.ascontiguousarray(params, dtype=np.float32)
    if params.size != warm_start_nparams():
        raise ValueError(f"expected {warm_start_nparams()} parameters, got {params.size}")
    rc = lib().parrsb_amd_set_warm_start(_p(params), passes, max_n)
    if rc:
        raise ParrsbError(rc, "bad warm-start parameters")


def test_tqli(d, e, fast=False):
    """Tridiagonal eigen-solve on the device: (lambda_min_abs, y); fast=True
    runs the FP32 path's multisection solve instead of the QL solve."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64) if len(e) else np.zeros(1)
    y = np.zeros(d.size)
    lam = np.zeros(1)
    fn = lib().rsbgpu_test_eigmin if fast else lib().rsbgpu_test_tqli
    rc = fn(_p(d), _p(e), d.size, _p(y), _p(lam))
    if rc:
        raise ParrsbError(rc, "rsbgpu_test_tqli failed")
    return lam[0], y


def stream_test(nbytes=1 << 30, reps=10, device=0):
    """Best read / write / copy rates (GB/s) of plain streaming kernels on
    the device: the practical HBM ceiling."""
    gbs = np.zeros(3)
    rc = lib().rsbgpu_stream_test(device, nbytes, reps, _p(gbs))
    if rc:
        raise ParrsbError(rc, "rsbgpu_stream_test failed")
    return {"read": gbs[0], "write": gbs[1], "copy": gbs[2]}


def test_sym_smallest(A):
    """The MG Ritz step's Jacobi eigen-solve on the device: (lambda_min, v)
    of one symmetric n x n matrix, n <= 63."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    v = np.zeros(n)
    lam = np.zeros(1)
    rc = lib().rsbgpu_test_sym_smallest(_p(A), n, _p(v), _p(lam))
    if rc:
        raise ParrsbError(rc, "rsbgpu_test_sym_smallest failed")
    return lam[0], v


def lws_test(coord, level, rowptr, col, wgt, lam, params):
    """Network forward + backward on one segment on the device (parity hook)."""
    coord = np.ascontiguousarray(coord, dtype=np.float64)  # 3 x M
    M = coord.shape[1]
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    wgt = np.ascontiguousarray(wgt, dtype=np.int32)
    params = np.ascontiguousarray(params, dtype=np.float32)
    y = np.zeros(M, dtype=np.float32)
    grad = np.zeros(params.size, dtype=np.float32)
    loss = C.c_double()
    rc = lib().rsbgpu_lws_test(M, _p(coord), level, _p(rowptr), _p(col), _p(wgt), lam,
                               _p(params), _p(y), _p(grad), C.byref(loss))
    if rc:
        raise ParrsbError(rc, "rsbgpu_lws_test failed")
    return y, grad, loss.value
