"""The benchmark-sized configurations on the GPU, checked through properties
that do not depend on size (the CPU oracle cannot run them in test time):

  * 128 x 128 x 64 hex box into 1,024 parts (FP32 against FP64 Lanczos):
    every part holds exactly 1,024 elements, the two edge cuts agree within
    2%, the FP32 run is bit-reproducible;
  * 2048^2 quads minus seeded disks (~3.98M elements, many components) into
    4,096 parts: the part sizes are exactly those of balanced cuts
    (rsb.c:138-140 applied level by level), FP32 and FP64 cuts within 2%;
  * a 2M-element warped pipe into 2,048 parts with and without the learned
    start: equal part sizes, edge cuts within 2%;
  * the 256^3 warped hex box (16,777,216 elements) into 65,536 parts: batched
    FP32 Lanczos with the learned start, the same from the reference start,
    and plain FP64: every part exactly 256 elements in all three, the three
    edge cuts within 2% of one another.
The holed quads' per-part component count is compared with the oracle's
count on the same partition (the oracle's quality pass is linear and runs at
this size).
"""
import numpy as np
import pytest

import parrsb_amd as pa
from parrsb_amd import meshgen as M

pytestmark = pytest.mark.gpu


def flush(msg):
    print(msg, flush=True)


def _balance_expected(n, p, b):
    """Size of half b after balancing (rsb.c:138-140)."""
    q, rem = n // p, n % p
    nh = (p + 1) // 2 if b == 0 else p // 2
    return q * nh + rem // 2 + (rem % 2) * (1 - b)


def test_hexbox_128x128x64_fp32_matches_fp64():
    flush("hex box 128x128x64 into 1024 parts, fp32 / fp64")
    vtx, xyz, nv = M.box_hex(128, 128, 64, jitter=0.1, seed=2)
    P = 1024
    pa.set_precision(32)
    try:
        a, _ = pa.part_mesh(vtx, xyz, nv, P)
        b, _ = pa.part_mesh(vtx, xyz, nv, P)
    finally:
        pa.set_precision(64)
    c, _ = pa.part_mesh(vtx, xyz, nv, P)
    assert np.array_equal(a, b)
    assert np.all(np.bincount(a, minlength=P) == vtx.shape[0] // P)
    assert np.all(np.bincount(c, minlength=P) == vtx.shape[0] // P)
    qa, qc = pa.quality(a, vtx, nv, P), pa.quality(c, vtx, nv, P)
    assert abs(qa["edge_cut"] - qc["edge_cut"]) <= 0.02 * qc["edge_cut"]


def _expected_sizes(n, P):
    """Part sizes when every cut ends balanced: the lower half holds
    balance_expected(n, p, 0) elements (rsb.c:118-229); the oracle gives
    these sizes on holed quad meshes too."""
    out = []

    def rec(n, p):
        if p == 1:
            out.append(n)
            return
        nl = _balance_expected(n, p, 0)
        rec(nl, (p + 1) // 2)
        rec(n - nl, p // 2)
    rec(n, P)
    return sorted(out)


def test_holed_quads_2048_sizes_and_precisions(oracle):
    flush("holed quads 2048^2 into 4096 parts")
    vtx, xyz, nv = M.holed_quads(2048, 2048, seed=3)
    n, P = vtx.shape[0], 4096
    pa.set_precision(32)
    try:
        a, _ = pa.part_mesh(vtx, xyz, nv, P, find_disconnected_comps=1)
        ra = pa.last_report()
    finally:
        pa.set_precision(64)
    c, _ = pa.part_mesh(vtx, xyz, nv, P, find_disconnected_comps=1)
    for part in (a, c):
        assert sorted(np.bincount(part, minlength=P).tolist()) == _expected_sizes(n, P)
    qa, qc = pa.quality(a, vtx, nv, P), pa.quality(c, vtx, nv, P)
    assert abs(qa["edge_cut"] - qc["edge_cut"]) <= 0.02 * qc["edge_cut"]
    assert len(ra["ncomp_max"]) == 12
    # disconnected pieces per part on the same partition, device against the
    # oracle's union-find (components.c:160-328 counts the same adjacency)
    for part in (a, c):
        q, qo = pa.quality(part, vtx, nv, P), oracle.quality(part, vtx, nv, P)
        assert q["components"] == qo["components"] >= P
        assert q["max_components"] == qo["max_components"]


def test_pipe_2m_learned_start_parity():
    flush("2M-element pipe into 2048 parts, learned start / reference start")
    import os
    from parrsb_amd import lws_train
    vtx, xyz, nv = M.pipe_hex(8, 256, 1024, seed=7)
    P = 2048
    params = lws_train.load_params(os.path.join(os.path.dirname(pa.__file__), "lws_weights.bin"))
    pa.set_precision(32)
    try:
        b, _ = pa.part_mesh(vtx, xyz, nv, P)
        pa.set_warm_start(params, 1)
        a, _ = pa.part_mesh(vtx, xyz, nv, P)
        r = pa.last_report()
    finally:
        pa.set_warm_start(None)
        pa.set_precision(64)
    assert r["warm_start"] == 1
    assert np.array_equal(np.bincount(a, minlength=P), np.bincount(b, minlength=P))
    qa, qb = pa.quality(a, vtx, nv, P), pa.quality(b, vtx, nv, P)
    assert abs(qa["edge_cut"] - qb["edge_cut"]) <= 0.02 * qb["edge_cut"]


def test_warped_256cubed_65536_parts():
    """The headline workload at full size: 2^24 elements over 2^16 parts
    split evenly at every cut, so each part holds exactly 256 elements
    whichever solver ran; the learned start and plain FP64 change only where
    the cuts fall, within 2% in edge cut."""
    import os
    from parrsb_amd import lws_train
    flush("warped hex 256^3 into 65536 parts: generating")
    vtx, xyz, nv = M.warped_box_hex(256, seed=5)
    P = 65536
    params = lws_train.load_params(os.path.join(os.path.dirname(pa.__file__), "lws_weights.bin"))
    parts, reports = {}, {}
    pa.set_precision(32)
    try:
        pa.set_warm_start(params, 1)
        flush("fp32 learned+batched")
        parts["learned"], _ = pa.part_mesh(vtx, xyz, nv, P)
        reports["learned"] = pa.last_report()
        pa.set_warm_start(None)
        flush("fp32 reference start")
        parts["reference"], _ = pa.part_mesh(vtx, xyz, nv, P)
    finally:
        pa.set_warm_start(None)
        pa.set_precision(64)
    flush("fp64 plain")
    parts["fp64"], _ = pa.part_mesh(vtx, xyz, nv, P)
    reports["fp64"] = pa.last_report()
    assert reports["learned"]["warm_start"] == 1
    cuts = {}
    for k, part in parts.items():
        assert np.all(np.bincount(part, minlength=P) == 256), k
        q = pa.quality(part, vtx, nv, P)
        assert q["max_part"] == q["min_part"] == 256 and q["empty_parts"] == 0
        cuts[k] = q["edge_cut"]
        flush(f"{k}: edge cut {cuts[k]}")
    for k in ("learned", "reference"):
        assert abs(cuts[k] - cuts["fp64"]) <= 0.02 * cuts["fp64"], k
    assert len(reports["fp64"]["niter_max"]) == 16
