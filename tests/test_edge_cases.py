"""Edge cases of parrsb_part_mesh on the device against the oracle: an empty
mesh, one part, more parts than elements, no coordinates (rsb_pre forced to
0, parrsb.c:250), the device-pointer entry point against the host one, and a
Lanczos forced not to converge (the warning of rsb.c:61-99)."""
import numpy as np
import pytest

import parrsb_amd as pa
from parrsb_amd import meshgen as M

pytestmark = pytest.mark.gpu


def flush(msg):
    print(msg, flush=True)


def test_empty_mesh_and_one_part():
    flush("empty mesh, one part")
    vtx = np.zeros((0, 8), dtype=np.int64)
    xyz = np.zeros((0, 8, 3))
    part, _ = pa.part_mesh(vtx, xyz, 8, 4)
    assert part.size == 0
    vtx, xyz, nv = M.box_hex(5, 4, 3, seed=1)
    part, _ = pa.part_mesh(vtx, xyz, nv, 1)
    assert np.all(part == 0)


@pytest.mark.parametrize("P", [16, 37])
def test_more_parts_than_elements(oracle, P):
    flush(f"10 elements into {P} parts")
    vtx, xyz, nv = M.box_hex(5, 2, 1, jitter=0.1, seed=3)
    part, _ = pa.part_mesh(vtx, xyz, nv, P)
    ref, _ = oracle.part_mesh(vtx, xyz, nv, P)
    assert part.min() >= 0 and part.max() < P
    assert np.array_equal(np.bincount(part, minlength=P), np.bincount(ref, minlength=P))


def test_no_coordinates_matches_oracle(oracle):
    flush("no coordinates")
    vtx, xyz, nv = M.box_hex(10, 8, 6, jitter=0.1, seed=4)
    part, opts = pa.part_mesh(vtx, None, nv, 8)
    assert opts.rsb_pre == 0  # parrsb.c:250
    ref, _ = oracle.part_mesh(vtx, None, nv, 8, rsb_pre=0)
    assert np.array_equal(np.bincount(part, minlength=8), np.bincount(ref, minlength=8))
    q, qo = pa.quality(part, vtx, nv, 8), oracle.quality(ref, vtx, nv, 8)
    assert abs(q["edge_cut"] - qo["edge_cut"]) <= 0.02 * qo["edge_cut"]


def test_device_entry_point_equals_host_entry_point():
    flush("device entry point")
    vtx, xyz, nv = M.box_hex(12, 10, 8, jitter=0.1, seed=6)
    n = vtx.shape[0]
    host, _ = pa.part_mesh(vtx, xyz, nv, 12)
    with pa.Device() as d:
        dv = d.upload(np.ascontiguousarray(vtx))
        dx = d.upload(np.ascontiguousarray(xyz))
        dp = d.upload(np.full(n, -1, dtype=np.int32))
        try:
            pa.part_mesh_dev(dp, dv, dx, n, nv, 12)
            dev = d.download(dp, np.zeros(n, dtype=np.int32))
        finally:
            for p in (dv, dx, dp):
                pa.Device.free(p)
    assert np.array_equal(dev, host)


def test_non_convergence_warning_and_residuals(oracle, monkeypatch, capfd):
    """PARRSB_RSB_MAX_ITER=2, PARRSB_RSB_MAX_PASSES=1: every segment stops at
    2 x 1 iterations without meeting rsb_tol, so every level reports the
    reference's warning with the minimum final residual and target over its
    segments (rsb.c:79-95; TOL_TGT / TOL_FNL of fiedler.c:451, 506); on the
    first cut (one segment, FP64) both equal the oracle's."""
    flush("forced non-convergence")
    monkeypatch.setenv("PARRSB_RSB_MAX_ITER", "2")
    monkeypatch.setenv("PARRSB_RSB_MAX_PASSES", "1")
    vtx, xyz, nv = M.box_hex(12, 10, 8, jitter=0.1, seed=6)
    n = vtx.shape[0]
    part, opts = pa.part_mesh(vtx, xyz, nv, 4)
    r = pa.last_report()
    assert (opts.rsb_max_iter, opts.rsb_max_passes) == (2, 1)
    assert r["niter_min"] == r["niter_max"] == [2, 2]
    assert r["nconverged_fail"] == [1, 2]
    err = capfd.readouterr().err
    for lv in range(2):
        assert 0 < r["tol_target"][lv] < r["tol_final"][lv]
        msg = ("Warning: Lanczos reached a residual of %f (target: %f) after 1 x 2 iterations "
               "in Level=%d!" % (r["tol_final"][lv], r["tol_target"][lv], lv))
        assert msg in err, (msg, err)
    order = oracle.rcb_order(oracle.centroids(xyz, nv), 3, np.arange(n, dtype=np.int32), 4)
    _, it, _, rn = oracle.fiedler(vtx, nv, order, max_iter=2, max_passes=1)
    assert it == 2
    assert abs(r["tol_init"][0] - rn[0]) <= 1e-9 * rn[0]
    assert abs(r["tol_target"][0] - rn[1]) <= 1e-9 * rn[1]
    assert abs(r["tol_final"][0] - rn[2]) <= 1e-9 * rn[2]
