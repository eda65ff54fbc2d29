"""The C-ABI library: loads, exports every function include/*.h declares, its
structs match the headers, and the host-side option logic behaves like the
reference's (parrsb.c:26-77, 244-250, 853-877).  No GPU needed: every call
here returns before the device is touched."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import parrsb_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def header_functions():
    names = set()
    for h in os.listdir(INC):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(INC, h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"#.*", "", txt)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", txt):
            name = m.group(1)
            if name.startswith(("parrsb_", "rsbgpu_", "fparrsb_")):
                names.add(name)
    return names


def test_library_exports_every_declared_function():
    L = pa.lib()
    declared = header_functions()
    assert len(declared) >= 30
    for name in sorted(declared):
        assert hasattr(L, name), f"{name} declared in include/ but not exported"
    # and the Python binding covers exactly the declared set
    assert set(pa.SIGNATURES) == declared
    nm = subprocess.run(["nm", "-D", "--defined-only", pa.LIB_PATH], capture_output=True,
                        text=True).stdout
    assert " parrsb_default_options" in nm


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "lay.c"
    src.write_text("""
#define PARRSB_NO_MPI
#include <stddef.h>
#include <stdio.h>
#include "parrsb_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\\n", sizeof(parrsb_options), offsetof(parrsb_options, rsb_tol),
         offsetof(parrsb_options, nparts), sizeof(parrsb_amd_report),
         offsetof(parrsb_amd_report, resid_max));
  return 0;
}
""")
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I" + INC, str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert out[0] == C.sizeof(pa.Options)
    assert out[1] == pa.Options.rsb_tol.offset
    assert out[2] == pa.Options.nparts.offset
    assert out[3] == C.sizeof(pa.Report)
    assert out[4] == pa.Report.resid_max.offset


def test_default_options_match_reference():
    o = pa.default_options().as_dict()
    assert o == dict(partitioner=0, tagged=0, levels=2, find_disconnected_comps=1, repair=0,
                     verbose_level=1, profile_level=0, rsb_algo=0, rsb_pre=1, rsb_max_iter=50,
                     rsb_max_passes=50, rsb_tol=1e-5, rsb_dump_stats=0, rsb_mg_grammian=0,
                     rsb_mg_factor=2, nparts=0)


def _call(vtx, xyz, nv, opts, tag=None):
    nel = 1
    part = np.zeros(nel, dtype=np.int32)
    return pa.lib().parrsb_part_mesh(pa._p(part), pa._p(vtx), pa._p(xyz), pa._p(tag), nel,
                                     nv, C.byref(opts), 0)


def test_env_overrides_mutate_callers_options(monkeypatch):
    env = {"PARRSB_RSB_MAX_ITER": "17", "PARRSB_RSB_TOL": "0.25", "PARRSB_NPARTS": "9",
           "PARRSB_VERBOSE_LEVEL": "0", "PARRSB_TAGGED": "1", "PARRSB_LEVELS": "1"}
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    opts = pa.default_options()
    vtx = np.arange(1, 9, dtype=np.int64)
    # tagged == 1 with tag == NULL returns 1 after the overrides (parrsb.c:873-877)
    assert _call(vtx, None, 8, opts) == 1
    assert (opts.rsb_max_iter, opts.rsb_tol, opts.nparts, opts.verbose_level, opts.levels) == \
        (17, 0.25, 9, 0, 1)
    assert "tag" in pa.last_error()


def test_input_errors_before_device(monkeypatch):
    for k in list(os.environ):
        if k.startswith("PARRSB_"):
            monkeypatch.delenv(k)
    opts = pa.default_options(verbose_level=0, nparts=2)
    assert _call(None, None, 8, opts) == 4           # both NULL (parrsb.c:244-249)
    vtx = np.arange(1, 9, dtype=np.int64)
    assert _call(vtx, None, 5, opts) == 3            # nv must be 4 or 8
    xyz = np.zeros(24)
    opts = pa.default_options(verbose_level=0, nparts=2, partitioner=3)
    assert _call(vtx, xyz, 8, opts) == 3             # no such partitioner
    opts = pa.default_options(verbose_level=0, nparts=2, partitioner=2)
    assert _call(vtx, None, 8, opts) == 3            # RIB needs coordinates
    opts = pa.default_options(verbose_level=0, nparts=2, rsb_algo=1, rsb_mg_grammian=1,
                              rsb_max_passes=65)
    assert _call(vtx, xyz, 8, opts) == 3             # Gram basis of at most 63 vectors
    assert "at most 64 passes" in pa.last_error()


def test_oracle_is_not_linked_by_the_product():
    nm = subprocess.run(["nm", "-D", pa.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle_" not in nm
    ldd = subprocess.run(["ldd", pa.LIB_PATH], capture_output=True, text=True).stdout
    assert "liboracle" not in ldd


def test_empty_tagged_input_returns_without_device(monkeypatch):
    """nel = 0 in tagged mode: the reference counts 0 tags and divides by
    that (parrsb.c:436); here the call returns 0 with nothing to write, and
    still zeroes verbose/profile on the caller's struct (parrsb.c:535-536)."""
    for k in list(os.environ):
        if k.startswith("PARRSB_"):
            monkeypatch.delenv(k)
    opts = pa.default_options(verbose_level=1, profile_level=1, nparts=4, tagged=1)
    part = np.zeros(1, dtype=np.int32)
    vtx = np.zeros(8, dtype=np.int64)
    xyz = np.zeros(24)
    tag = np.zeros(1, dtype=np.int32)
    rc = pa.lib().parrsb_part_mesh(pa._p(part), pa._p(vtx), pa._p(xyz), pa._p(tag), 0, 8,
                                   C.byref(opts), 0)
    assert rc == 0
    assert opts.verbose_level == 0 and opts.profile_level == 0


def test_fortran_conn_mesh_binding_checks_arguments(monkeypatch):
    """fparrsb_conn_mesh (con.c:297-307): arguments by reference, the return
    code in *err; a bad dimension is rejected before any device work."""
    monkeypatch.setenv("PARRSB_VERBOSE_LEVEL", "0")
    vtx = np.zeros(8, dtype=np.int64)
    xyz = np.zeros(24)
    ints = {k: C.c_int(v) for k, v in dict(nel=1, ndim=5, npf=0, comm=0, err=-7).items()}
    tol = C.c_double(0.2)
    for name in ("fparrsb_conn_mesh_", "fparrsb_conn_mesh"):
        fn = getattr(pa.lib(), name)
        fn.restype = None
        ints["err"].value = -7
        fn(pa._p(vtx), pa._p(xyz), C.byref(ints["nel"]), C.byref(ints["ndim"]), None,
           C.byref(ints["npf"]), C.byref(tol), C.byref(ints["comm"]), C.byref(ints["err"]))
        assert ints["err"].value == 3
        assert "nDim" in pa.last_error()


def test_noalloc_entry_point_rejects_sizes_beyond_its_maximum(monkeypatch):
    """parrsb_amd_part_mesh_dev_noalloc (no counterpart in the reference)
    holds every buffer in a fixed device array: meshes above its element
    limit, more parts than that, and tagged mode return 3 before any device
    work, with a message."""
    for k in list(os.environ):
        if k.startswith("PARRSB_"):
            monkeypatch.delenv(k)
    fn = pa.lib().parrsb_amd_part_mesh_dev_noalloc
    opts = pa.default_options(verbose_level=0, nparts=4)
    assert fn(None, None, None, 65537, 8, C.byref(opts)) == 3
    assert "at most 65536 elements" in pa.last_error()
    opts = pa.default_options(verbose_level=0, nparts=65537)
    assert fn(None, None, None, 100, 8, C.byref(opts)) == 3
    assert "at most 65536 parts" in pa.last_error()
    opts = pa.default_options(verbose_level=0, nparts=4, tagged=1)
    assert fn(None, None, None, 100, 8, C.byref(opts)) == 3
    opts = pa.default_options(verbose_level=0, nparts=4, rsb_algo=1)
    assert fn(None, None, None, 32769, 8, C.byref(opts)) == 3
    assert "rsb_algo = 1 takes at most 32768 elements" in pa.last_error()
    opts = pa.default_options(verbose_level=0, nparts=4, rsb_algo=1, rsb_mg_grammian=1)
    assert fn(None, None, None, 100, 8, C.byref(opts)) == 3
    assert "rsb_mg_grammian = 1" in pa.last_error()
