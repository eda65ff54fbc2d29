"""Connectivity (parrsb_conn_mesh): vertex ids from corner coordinates.

Periodic faces (con-periodic.c) are pinned the same way: boxes and quad
grids periodic along one, two or all axes (corner classes that chain across
several periodic pairs) must give the wrapped lattice's vertex classes, with
ids 1..#vertices; faces whose partner does not fit fail.

The oracle is pinned by the reference's own acceptance test (gencon.c:21-63:
the ids define the same vertex classes as a known connectivity, both ways)
on generated meshes whose connectivity is known, including a periodic
wrap-around (pipe) and corners perturbed within the tolerance, and by its
two failure checks on meshes built to fail them.  The device path must give
the oracle's ids exactly (same refinement, same numbering order)."""
import numpy as np
import pytest

from parrsb_amd import meshgen as M

PRE_TO_SYM = [0, 1, 3, 2, 4, 5, 7, 6]


def same_classes(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    # a -> b and b -> a are both functions
    pa = np.unique(np.stack([a, b], 1), axis=0)
    return len(np.unique(pa[:, 0])) == len(pa) and len(np.unique(pa[:, 1])) == len(pa)


def known_mesh(name):
    if name == "box":
        vtx, xyz, nv = M.box_hex(6, 5, 4, jitter=0.0)
        return vtx, xyz, 3
    if name == "pipe":  # periodic around the axis: ids wrap
        vtx, xyz, nv = M.pipe_hex(2, 12, 5, seed=3, amp=0.1)
        return vtx, xyz, 3
    if name == "warped":
        vtx, xyz, nv = M.warped_box_hex(6, seed=2, amp=0.05)
        return vtx, xyz, 3
    if name == "quads":
        vtx, xyz, nv = M.quad_grid(9, 7)
        return vtx, xyz, 2
    raise KeyError(name)


@pytest.mark.parametrize("name", ["box", "pipe", "warped", "quads"])
def test_oracle_ids_define_the_known_vertices(oracle, name):
    vtx, xyz, ndim = known_mesh(name)
    nv = 8 if ndim == 3 else 4
    rc, out = oracle.conn_mesh(xyz, ndim, 0.2)
    assert rc == 0
    assert same_classes(out, vtx[:, PRE_TO_SYM[:nv]])
    # ids are 1..#vertices
    assert out.min() == 1 and out.max() == len(np.unique(vtx))


def test_oracle_merges_within_tolerance_only(oracle):
    """Corners moved by 5% of the edge still merge; a corner moved by 50%
    splits off (and then the face check catches the broken face)."""
    vtx, xyz, nv = M.box_hex(4, 3, 3, jitter=0.0)
    rng = np.random.default_rng(0)
    h = 1.0 / 4
    x2 = xyz + rng.uniform(-0.05 * h, 0.05 * h, size=xyz.shape)
    rc, out = oracle.conn_mesh(x2, 3, 0.2)
    assert rc == 0 and same_classes(out, vtx[:, PRE_TO_SYM])
    x3 = xyz.copy()
    x3[5, 6] += 0.5 * h  # one corner of element 5 moves away
    rc, out = oracle.conn_mesh(x3, 3, 0.2)
    assert rc == 1
    assert len(np.unique(out)) == len(np.unique(vtx)) + 1


def test_oracle_element_check(oracle):
    """An element with two coinciding corners fails the element check."""
    vtx, xyz, nv = M.box_hex(2, 2, 1, jitter=0.0)
    x = xyz.copy()
    x[0, 1] = x[0, 0]
    rc, _ = oracle.conn_mesh(x, 3, 0.2)
    assert rc == 1


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box", "pipe", "warped", "quads"])
def test_device_ids_equal_oracle(oracle, name):
    import parrsb_amd as pa
    vtx, xyz, ndim = known_mesh(name)
    rc, ref = oracle.conn_mesh(xyz, ndim, 0.2)
    out = pa.conn_mesh(xyz, ndim, 0.2)
    assert np.array_equal(out, ref)


@pytest.mark.gpu
def test_fortran_binding_equals_c_entry(oracle):
    """fparrsb_conn_mesh (con.c:297-307), the entry Nek5000 calls with every
    argument by reference, gives the C entry's ids and return code."""
    import ctypes as C
    import parrsb_amd as pa
    vtx, xyz, ndim = known_mesh("pipe")
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nel = xyz.shape[0]
    out = np.zeros((nel, 8), dtype=np.int64)
    a = [C.c_int(nel), C.c_int(ndim), C.c_int(0), C.c_int(0), C.c_int(-1)]
    tol = C.c_double(0.2)
    pa.lib().fparrsb_conn_mesh_(pa._p(out), pa._p(xyz), C.byref(a[0]), C.byref(a[1]), None,
                                C.byref(a[2]), C.byref(tol), C.byref(a[3]), C.byref(a[4]))
    assert a[4].value == 0
    rc, ref = oracle.conn_mesh(xyz, ndim, 0.2)
    assert np.array_equal(out, ref) and np.array_equal(out, pa.conn_mesh(xyz, ndim, 0.2))


@pytest.mark.gpu
def test_device_checks_and_partition_of_generated_ids(oracle):
    import parrsb_amd as pa
    vtx, xyz, nv = M.box_hex(4, 3, 3, jitter=0.0)
    x3 = xyz.copy()
    x3[5, 6] += 0.5 / 4
    with pytest.raises(pa.ParrsbError) as ei:
        pa.conn_mesh(x3, 3, 0.2)
    assert ei.value.rc == 1
    # ids from coordinates partition exactly like the generator's ids
    vtx, xyz, nv = M.warped_box_hex(8, seed=4, amp=0.05)
    ids = pa.conn_mesh(xyz, 3, 0.2)
    # back to the preprocessor corner order of xyz
    ids_pre = ids[:, PRE_TO_SYM]
    p1, _ = pa.part_mesh(ids_pre, xyz, 8, 16)
    p2, _ = pa.part_mesh(vtx, xyz, 8, 16)
    q1, q2 = pa.quality(p1, vtx, 8, 16), pa.quality(p2, vtx, 8, 16)
    assert q1 == q2


PERIODIC = {
    "hex-x": dict(nx=5, ny=4, nz=3, periodic=(True, False, False)),
    "hex-xyz": dict(nx=5, ny=4, nz=3, periodic=(True, True, True)),
    "hex-yz-warped": dict(nx=4, ny=6, nz=5, periodic=(False, True, True), warp=0.15),
    "quad-xy": dict(nx=7, ny=5, nz=0, periodic=(True, True, False), warp=0.1),
    "quad-y": dict(nx=6, ny=5, nz=0, periodic=(False, True, False)),
}


@pytest.mark.parametrize("name", sorted(PERIODIC))
def test_oracle_periodic_faces_give_the_wrapped_lattice(oracle, name):
    vtx, xyz, rec = M.periodic_box(**PERIODIC[name])
    ndim = xyz.shape[2]
    nv = 8 if ndim == 3 else 4
    rc, out = oracle.conn_mesh(xyz, ndim, 0.2, rec)
    assert rc == 0
    assert same_classes(out, vtx[:, PRE_TO_SYM[:nv]])
    assert out.min() == 1 and out.max() == len(np.unique(vtx))
    # without the records the periodic faces stay apart
    rc0, out0 = oracle.conn_mesh(xyz, ndim, 0.2)
    assert rc0 == 0 and out0.max() > out.max()


def test_oracle_periodic_faces_that_do_not_fit_fail(oracle):
    vtx, xyz, rec = M.periodic_box(5, 4, 3, periodic=(True, False, False))
    bad = rec.copy()
    bad[0, 1], bad[1, 3] = 1, 1  # element face y=0 paired with a face x=1
    rc, _ = oracle.conn_mesh(xyz, 3, 0.2, bad)
    assert rc == 1
    out_of_range = rec.copy()
    out_of_range[0, 0] = 10 ** 6
    rc, _ = oracle.conn_mesh(xyz, 3, 0.2, out_of_range)
    assert rc == 3


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(PERIODIC))
def test_device_periodic_ids_equal_oracle(oracle, name):
    import parrsb_amd as pa
    vtx, xyz, rec = M.periodic_box(**PERIODIC[name])
    ndim = xyz.shape[2]
    rc, ref = oracle.conn_mesh(xyz, ndim, 0.2, rec)
    out = pa.conn_mesh(xyz, ndim, 0.2, periodic=rec)
    assert rc == 0 and np.array_equal(out, ref)


@pytest.mark.gpu
def test_device_periodic_mismatch_and_range_errors():
    import parrsb_amd as pa
    vtx, xyz, rec = M.periodic_box(5, 4, 3, periodic=(True, False, False))
    bad = rec.copy()
    bad[0, 1], bad[1, 3] = 1, 1
    with pytest.raises(pa.ParrsbError) as e:
        pa.conn_mesh(xyz, 3, 0.2, periodic=bad)
    assert e.value.rc == 1 and "did not match" in str(e.value)
    out_of_range = rec.copy()
    out_of_range[0, 0] = 10 ** 6
    with pytest.raises(pa.ParrsbError) as e:
        pa.conn_mesh(xyz, 3, 0.2, periodic=out_of_range)
    assert e.value.rc == 3
