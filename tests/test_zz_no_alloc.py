"""The allocation-free entry point (parrsb_amd_part_mesh_dev_noalloc, no
counterpart in the reference) against the allocating one
(parrsb_amd_part_mesh_dev) through the C ABI: the same part ids, element for
element, on two calls in a row, the second on a smaller mesh (the fixed
device array keeps the first call's contents), with the Lanczos and the MG
solvers, and at the element limit."""
import ctypes as C

import numpy as np
import pytest

import parrsb_amd as pa
from parrsb_amd import meshgen as M

pytestmark = pytest.mark.gpu

NOALLOC_BYTES = 256 << 20
MAX_NEL = 65536
MAX_NEL_MG = 32768


def flush(msg):
    print(msg, flush=True)


def _call(entry, vtx, xyz, nv, P, **opt):
    n = vtx.shape[0]
    with pa.Device() as d:
        dv = d.upload(np.ascontiguousarray(vtx))
        dx = d.upload(np.ascontiguousarray(xyz))
        dp = d.upload(np.full(n, -1, dtype=np.int32))
        try:
            opts = pa.default_options(verbose_level=0, nparts=P, **opt)
            rc = getattr(pa.lib(), entry)(dp, dv, dx, n, nv, C.byref(opts))
            err = pa.last_error()
            part = d.download(dp, np.zeros(n, dtype=np.int32))
        finally:
            for p in (dv, dx, dp):
                pa.Device.free(p)
    return rc, err, part, pa.last_report()


def _both(vtx, xyz, nv, P, **opt):
    rc, err, a, ra = _call("parrsb_amd_part_mesh_dev", vtx, xyz, nv, P, **opt)
    assert rc == 0, err
    rc, err, b, rb = _call("parrsb_amd_part_mesh_dev_noalloc", vtx, xyz, nv, P, **opt)
    assert rc == 0, err
    used, high = rb["noalloc_bytes"], rb["noalloc_high"]
    flush(f"{vtx.shape[0]} elements into {P} parts {opt}: {used / 2**20:.1f} MiB held at most, "
          f"{high / 2**20:.1f} MiB of the array touched")
    assert 0 < used <= high <= NOALLOC_BYTES
    assert ra["noalloc_bytes"] == 0
    return a, b, ra, rb


def test_noalloc_equals_allocating_entry_point_twice():
    vtx, xyz, nv = M.box_hex(32, 24, 16, jitter=0.1, seed=3)
    a, b, ra, rb = _both(vtx, xyz, nv, 48)
    assert np.array_equal(a, b)
    assert ra["niter_max"] == rb["niter_max"] and ra["lambda_min"] == rb["lambda_min"]
    # a smaller mesh next: the array still holds the first call's data
    vtx, xyz, nv = M.holed_quads(40, 36, seed=7, frac=0.1, rmin=2, rmax=5)
    a, b, ra, rb = _both(vtx, xyz, nv, 12)
    assert np.array_equal(a, b)
    assert ra["ncomp_max"] == rb["ncomp_max"]


def test_noalloc_mg_solver_and_fp32_learned_start():
    import os
    from parrsb_amd import lws_train
    vtx, xyz, nv = M.box_hex(20, 16, 12, jitter=0.1, seed=5)
    a, b, _, _ = _both(vtx, xyz, nv, 16, rsb_algo=1, rsb_mg_factor=2)
    assert np.array_equal(a, b)
    params = lws_train.load_params(os.path.join(os.path.dirname(pa.__file__), "lws_weights.bin"))
    pa.set_precision(32)
    pa.set_warm_start(params, 1)
    try:
        vtx, xyz, nv = M.pipe_hex(4, 32, 64, seed=7)
        a, b, ra, rb = _both(vtx, xyz, nv, 32)
    finally:
        pa.set_warm_start(None)
        pa.set_precision(64)
    assert rb["warm_start"] == 1
    assert np.array_equal(a, b)


def test_noalloc_at_the_element_limit(monkeypatch):
    """The largest mesh allowed, once with the whole array and once with
    only part of it usable (PARRSB_AMD_NOALLOC_LIMIT), so that freed blocks
    are handed out again after the stream drains: both give the allocating
    entry point's parts."""
    vtx, xyz, nv = M.box_hex(64, 64, 16, jitter=0.1, seed=2)
    assert vtx.shape[0] == MAX_NEL
    a, b, _, rb = _both(vtx, xyz, nv, 256)
    assert np.array_equal(a, b)
    assert np.all(np.bincount(b, minlength=256) == 256)
    assert rb["noalloc_bytes"] <= 0.4 * NOALLOC_BYTES
    limit = int(rb["noalloc_bytes"] * 1.5) // 256 * 256
    assert limit < rb["noalloc_high"]  # the reuse path must run
    monkeypatch.setenv("PARRSB_AMD_NOALLOC_LIMIT", str(limit))
    a2, b2, _, rb2 = _both(vtx, xyz, nv, 256)
    assert rb2["noalloc_high"] <= limit
    assert np.array_equal(a2, b2) and np.array_equal(b2, b)


def test_noalloc_mg_at_its_element_limit():
    """The MG solver keeps rsb_max_iter stored directions per row, the most
    memory per element of any path, so it has the lower limit
    (PARRSB_AMD_NOALLOC_MAX_NEL_MG); at that limit it holds about half the
    array."""
    vtx, xyz, nv = M.box_hex(64, 32, 16, jitter=0.1, seed=2)
    assert vtx.shape[0] == MAX_NEL_MG
    a, b, _, rb = _both(vtx, xyz, nv, 16, rsb_algo=1, rsb_mg_factor=2)
    assert np.array_equal(a, b)
    assert rb["noalloc_bytes"] <= 0.6 * NOALLOC_BYTES
