"""Tagged partitioning (parrsb.c:379-561): every distinct tag is partitioned
on its own into nparts / ntags parts, the i-th smallest tag taking parts
[i * chunk, (i + 1) * chunk).  Tags are ordered as unsigned 32-bit values (the
reference sorts a uint field, parrsb.c:388-402): -3 comes after 100.

The oracle's restatement is pinned on meshes made of disjoint copies of one
box: each copy's tagged parts must be the untagged partition of the box
alone, shifted by its tag's rank -- a wrong tag order, a lost offset or a
reordering of the elements inside a tag shows up there.  The device path
must then give the oracle's sizes and quality (RSB, FP64) or its exact part
ids (RCB)."""
import numpy as np
import pytest

import parrsb_amd as pa
from parrsb_amd import meshgen as M


def flush(msg):
    print(msg, flush=True)


def copies(ncopy, nx, ny, nz, seed=3):
    """ncopy disjoint copies of one jittered box (vertex ids shifted per copy),
    elements interleaved round-robin in the input; returns the mesh and the
    copy index of every element."""
    vtx, xyz, nv = M.box_hex(nx, ny, nz, jitter=0.1, seed=seed)
    n = vtx.shape[0]
    off = int(vtx.max())
    V = np.concatenate([vtx + c * off for c in range(ncopy)])
    X = np.concatenate([xyz for _ in range(ncopy)])
    copy = np.repeat(np.arange(ncopy), n)
    perm = np.arange(ncopy * n).reshape(ncopy, n).T.ravel()  # round-robin
    return V[perm], X[perm], nv, copy[perm], (vtx, xyz)


def layered(nx, ny, nz, per_layer, seed=5):
    vtx, xyz, nv = M.box_hex(nx, ny, nz, jitter=0.1, seed=seed)
    k = np.arange(vtx.shape[0]) // (nx * ny)
    return vtx, xyz, nv, (k // per_layer).astype(np.int32)


def test_oracle_tagged_copies_are_shifted_untagged_partitions(oracle):
    V, X, nv, copy, (vtx, xyz) = copies(3, 6, 5, 4)
    # unsigned order: 7 -> 0, 100 -> 1, -3 (0xfffffffd) -> 2
    tags = np.array([7, -3, 100], dtype=np.int32)
    rank = {0: 0, 1: 2, 2: 1}
    chunk = 4
    part = oracle.part_mesh_tagged(V, X, nv, 3 * chunk, tags[copy])
    alone, _ = oracle.part_mesh(vtx, xyz, nv, chunk)
    for c in range(3):
        assert np.array_equal(part[copy == c], alone + rank[c] * chunk)


def test_oracle_tagged_single_tag_is_untagged(oracle):
    vtx, xyz, nv = M.box_hex(8, 6, 5, jitter=0.1, seed=2)
    part = oracle.part_mesh_tagged(vtx, xyz, nv, 5, np.full(vtx.shape[0], 42, np.int32))
    ref, _ = oracle.part_mesh(vtx, xyz, nv, 5)
    assert np.array_equal(part, ref)


def test_oracle_tagged_needs_a_multiple_of_the_tag_count(oracle):
    vtx, xyz, nv, tag = layered(4, 4, 6, 2)
    with pytest.raises(RuntimeError, match="returned 3"):
        oracle.part_mesh_tagged(vtx, xyz, nv, 4, tag)  # 3 tags


@pytest.mark.gpu
def test_tagged_layers_match_oracle(oracle):
    flush("tagged layers 12x10x6, 3 tags, P=6")
    vtx, xyz, nv, tag = layered(12, 10, 6, 2)
    P = 6
    part, _ = pa.part_mesh(vtx, xyz, nv, P, tag=tag, tagged=1)
    ref = oracle.part_mesh_tagged(vtx, xyz, nv, P, tag)
    assert np.array_equal(np.bincount(part, minlength=P), np.bincount(ref, minlength=P))
    for t in range(3):
        assert set(np.unique(part[tag == t])) <= {2 * t, 2 * t + 1}
    q, qo = pa.quality(part, vtx, nv, P), oracle.quality(ref, vtx, nv, P)
    assert q == qo


@pytest.mark.gpu
def test_tagged_copies_match_oracle(oracle):
    flush("tagged copies 3 x 8x7x6, P=12")
    V, X, nv, copy, _ = copies(3, 8, 7, 6)
    tag = np.array([7, -3, 100], dtype=np.int32)[copy]
    P = 12
    part, _ = pa.part_mesh(V, X, nv, P, tag=tag, tagged=1)
    ref = oracle.part_mesh_tagged(V, X, nv, P, tag)
    assert np.array_equal(np.bincount(part, minlength=P), np.bincount(ref, minlength=P))
    for c, r in ((0, 0), (1, 2), (2, 1)):  # tags 7, -3, 100 in unsigned order
        sel = copy == c
        assert set(np.unique(part[sel])) == set(range(4 * r, 4 * r + 4))
        q = pa.quality(part[sel] - 4 * r, V[sel], nv, 4)
        qo = oracle.quality(ref[sel] - 4 * r, V[sel], nv, 4)
        assert q == qo


@pytest.mark.gpu
def test_tagged_rcb_partitioner_matches_oracle_ranks(oracle):
    flush("tagged RCB")
    vtx, xyz, nv, tag = layered(9, 8, 8, 2)  # 4 tags
    P = 12
    part, _ = pa.part_mesh(vtx, xyz, nv, P, tag=tag, tagged=1, partitioner=1)
    c = oracle.centroids(xyz, nv)
    ref = np.full(vtx.shape[0], -1, dtype=np.int32)

    def ranges(order, start, length, p, r0):
        if p == 1:
            ref[order[start:start + length]] = r0
            return
        l = oracle.lower_count(length, p)
        ranges(order, start, l, (p + 1) // 2, r0)
        ranges(order, start + l, length - l, p // 2, r0 + (p + 1) // 2)

    for t in range(4):
        el = np.flatnonzero(tag == t).astype(np.int32)
        order = oracle.rcb_order(c, 3, el, 3)
        ranges(order, 0, el.size, 3, 3 * t)
    assert np.array_equal(part, ref)


@pytest.mark.gpu
def test_tagged_single_tag_equals_untagged():
    vtx, xyz, nv = M.box_hex(10, 9, 7, jitter=0.1, seed=4)
    a, _ = pa.part_mesh(vtx, xyz, nv, 8)
    b, o = pa.part_mesh(vtx, xyz, nv, 8, tag=np.full(vtx.shape[0], -5, np.int32), tagged=1)
    assert np.array_equal(a, b)


@pytest.mark.gpu
def test_tagged_errors_and_option_mutation():
    vtx, xyz, nv, tag = layered(4, 4, 6, 2)  # 3 tags
    with pytest.raises(pa.ParrsbError) as e:
        pa.part_mesh(vtx, xyz, nv, 4, tag=tag, tagged=1)
    assert e.value.rc == 3 and "multiple" in str(e.value)
    opts = pa.default_options(verbose_level=1, profile_level=1)
    part, o = pa.part_mesh(vtx, xyz, nv, 6, tag=tag, options=opts, tagged=1)
    assert o.verbose_level == 0 and o.profile_level == 0  # parrsb.c:535-536
    part2, seq, _ = pa.part_mesh_seq(vtx, xyz, nv, 6, tag=tag, tagged=1)
    assert np.array_equal(part, part2)
    for q in range(6):
        assert np.array_equal(np.sort(seq[part2 == q]), np.arange(np.count_nonzero(part2 == q)))


@pytest.mark.gpu
def test_group_by_tag_is_a_stable_sort():
    vtx, xyz, nv = M.box_hex(9, 7, 6, jitter=0.1, seed=1)
    rng = np.random.default_rng(11)
    tag = rng.choice(np.array([-2**31, -7, 0, 3, 2**31 - 1], dtype=np.int64), vtx.shape[0])
    tag = tag.astype(np.int32)
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        counts = d.group_by_tag(tag, 8)
        order = d.order()
        with pytest.raises(ValueError):
            d.group_by_tag(tag, 4)
    ref = np.argsort(tag.view(np.uint32), kind="stable")
    assert np.array_equal(order, ref)
    assert np.array_equal(counts, np.unique(tag.view(np.uint32), return_counts=True)[1])
