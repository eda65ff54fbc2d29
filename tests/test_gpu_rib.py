"""RIB (rsb_pre = 2 and partitioner = 2) on the device against the oracle.

Both take the exact dominant covariance eigenvector with the same sign rule;
they differ only in the rounding of the node sums (compensated on the
device, plain in the oracle), so on meshes whose projections are distinct
the orders agree element for element."""
import numpy as np
import pytest

import parrsb_amd as pa
from parrsb_amd import meshgen as M

pytestmark = pytest.mark.gpu


def flush(msg):
    print(msg, flush=True)


def jittered_quads(nx, ny, seed):
    vtx, xyz, nv = M.quad_grid(nx, ny)
    rng = np.random.default_rng(seed)
    h = 1.0 / max(nx, ny)
    xyz = xyz + rng.uniform(-0.1 * h, 0.1 * h, size=(xyz.shape[0], 1, 2))
    return vtx, xyz, nv


MESHES = {
    "hex": lambda: M.box_hex(11, 7, 5, jitter=0.1, seed=3),
    "pipe": lambda: M.pipe_hex(3, 16, 20, seed=5, amp=0.1),
    "quad": lambda: jittered_quads(29, 17, seed=4),
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("p", [2, 3, 8, 13])
def test_rib_order_matches_oracle(oracle, name, p):
    flush(f"rib {name} p={p}")
    vtx, xyz, nv = MESHES[name]()
    n = vtx.shape[0]
    ndim = 3 if nv == 8 else 2
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        d.set_segments([0], [n], [p], [0])
        d.prepartition(2)
        got = d.order()
    ref = oracle.rib_order(oracle.centroids(xyz, nv), ndim, np.arange(n, dtype=np.int32), p)
    assert np.array_equal(got, ref)


def test_rib_many_segments_match_oracle(oracle):
    vtx, xyz, nv = M.box_hex(20, 9, 7, jitter=0.1, seed=6)
    n = vtx.shape[0]
    c = oracle.centroids(xyz, nv)
    cs, ln, ps = [0, 300, 301, 700, 900], [300, 1, 399, 200, n - 900], [4, 2, 7, 1, 5]
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        d.set_segments(cs, ln, ps, [0, 4, 6, 13, 14])
        d.prepartition(2)
        got = d.order()
    ref = np.arange(n, dtype=np.int32)
    for s0, l, p in zip(cs, ln, ps):
        if p > 1:
            ref[s0:s0 + l] = oracle.rib_order(c, 3, ref[s0:s0 + l], p)
    assert np.array_equal(got, ref)


def test_rsb_with_rib_prepartition_matches_oracle(oracle):
    vtx, xyz, nv = M.box_hex(12, 10, 8, jitter=0.1, seed=9)
    P = 8
    part, _ = pa.part_mesh(vtx, xyz, nv, P, rsb_pre=2)
    ref, _ = oracle.part_mesh(vtx, xyz, nv, P, rsb_pre=2)
    q, qr = pa.quality(part, vtx, nv, P), oracle.quality(ref, vtx, nv, P)
    assert q["max_part"] == qr["max_part"] and q["min_part"] == qr["min_part"]
    assert abs(q["edge_cut"] - qr["edge_cut"]) <= 0.02 * qr["edge_cut"]


def test_rib_partitioner_matches_oracle_ranks(oracle):
    vtx, xyz, nv = M.pipe_hex(3, 20, 24, seed=8, amp=0.15)
    n = vtx.shape[0]
    P = 7
    part, _ = pa.part_mesh(vtx, xyz, nv, P, partitioner=2)
    order = oracle.rib_order(oracle.centroids(xyz, nv), 3, np.arange(n, dtype=np.int32), P)
    ref = np.zeros(n, dtype=np.int32)

    def ranges(start, length, p, r0):
        if p == 1:
            ref[order[start:start + length]] = r0
            return
        l = oracle.lower_count(length, p)
        ranges(start, l, (p + 1) // 2, r0)
        ranges(start + l, length - l, p // 2, r0 + (p + 1) // 2)
    ranges(0, n, P, 0)
    assert np.array_equal(part, ref)
