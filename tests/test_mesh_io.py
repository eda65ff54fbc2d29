"""Nek5000 mesh files (io.c) and the helpers around them (helpers.c).

Fixtures: tests/golden/ethier.re2 is the reference's own example mesh
(examples/ethier.re2: 32 hexes, 24 'EXO' boundary faces, no periodic ones).
Periodic meshes are written here by `write_re2` in the .re2 layout, so the
reader must return exactly the coordinates and the 'P  ' records that went
in.  The .co2 / .ma2 writers are checked byte for byte against the layouts
io.c:339-470 write, and by reading the .co2 back."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import parrsb_amd as pa
from parrsb_amd import meshgen as M

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ETHIER = os.path.join(GOLDEN, "ethier")
PRE_TO_SYM = [0, 1, 3, 2, 4, 5, 7, 6]


def write_re2(path, xyz, bcs=(), other=()):
    """.re2 writer for tests: xyz[nel, nv, ndim] preprocessor order; bcs are
    periodic records (element, face, partner element, partner face); other
    are (element, face, type) non-periodic boundary faces."""
    nel, nv, ndim = xyz.shape
    hdr = f"#v003{nel:12d}{ndim:3d}{nel:12d} this is the hdr".ljust(80).encode()
    out = [hdr, struct.pack("<f", 6.54321)]
    for e in range(nel):
        out.append(struct.pack("<d", 0.0))
        for d in range(ndim):
            out.append(xyz[e, :, d].astype("<f8").tobytes())
    out.append(struct.pack("<d", 2.0))  # two curve records, skipped by the reader
    out.append(np.arange(16, dtype="<f8").tobytes())
    recs = []
    for e, f, e2, f2 in bcs:
        recs.append(struct.pack("<7d", e, f, e2, f2, 0, 0, 0) + b"P       ")
    for e, f, t in other:
        recs.append(struct.pack("<7d", e, f, 0, 0, 0, 0, 0) + t.ljust(8).encode())
    out.append(struct.pack("<d", float(len(recs))))
    out += recs
    with open(path, "wb") as fh:
        fh.write(b"".join(out))


def test_read_ethier_fixture(oracle):
    m = pa.read_mesh(ETHIER, read=1)
    assert m["nel"] == 32 and m["nv"] == 8
    xyz = m["coord"]
    assert xyz.shape == (32, 8, 3) and m["bcs"].shape == (0, 4)
    assert xyz.min() == -1.0 and xyz.max() == 1.0
    # corners of the example mesh coincide exactly where they are shared: the
    # oracle's connectivity must find exactly the distinct coordinates
    rc, vl = oracle.conn_mesh(xyz, 3, 0.2)
    assert rc == 0
    pts = xyz[:, PRE_TO_SYM, :].reshape(-1, 3)
    _, cls = np.unique(pts, axis=0, return_inverse=True)
    assert vl.max() == cls.max() + 1 == 53
    a, b = vl.ravel(), cls.ravel()
    pairs = np.unique(np.stack([a, b], 1), axis=0)
    assert len(pairs) == 53


def test_read_written_periodic_re2(tmp_path):
    vtx, xyz, rec = M.periodic_box(4, 3, 2, periodic=(True, False, True), warp=0.1)
    other = [(1, 1, "W"), (2, 1, "v"), (5, 3, "O")]
    write_re2(tmp_path / "box.re2", xyz, rec, other)
    m = pa.read_mesh(str(tmp_path / "box"), read=1)
    assert m["nel"] == 24 and m["nv"] == 8
    assert np.array_equal(m["coord"], xyz)
    assert np.array_equal(m["bcs"], rec)


def test_read_errors(tmp_path):
    with pytest.raises(pa.ParrsbError, match="cannot open"):
        pa.read_mesh(str(tmp_path / "missing"), read=1)
    vtx, xyz, rec = M.periodic_box(2, 2, 2, periodic=(False, False, False))
    write_re2(tmp_path / "b.re2", xyz)
    raw = bytearray(open(tmp_path / "b.re2", "rb").read())
    raw[80:84] = struct.pack(">f", 6.54321)  # the other byte order
    open(tmp_path / "c.re2", "wb").write(bytes(raw))
    with pytest.raises(pa.ParrsbError, match="byte test"):
        pa.read_mesh(str(tmp_path / "c"), read=1)
    raw[80:84] = struct.pack("<f", 6.54321)
    open(tmp_path / "d.re2", "wb").write(bytes(raw[:84 + 100]))
    with pytest.raises(pa.ParrsbError, match="truncated"):
        pa.read_mesh(str(tmp_path / "d"), read=1)


def test_co2_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    vl = rng.integers(1, 1000, size=(7, 8)).astype(np.int64)
    name = str(tmp_path / "m")
    pa.dump_con(name, vl)
    raw = open(name + ".co2", "rb").read()
    hdr = f"{'#v001':>5s}{7:12d}{7:12d}{8:12d}".ljust(132).encode()
    body = b"".join(struct.pack("<9i", e + 1, *vl[e]) for e in range(7))
    assert raw == hdr + struct.pack("<f", 6.54321) + body
    m = pa.read_mesh(name, read=2)
    assert m["nel"] == 7 and m["nv"] == 8 and np.array_equal(m["vl"], vl)
    # read & 3: nel and nv must agree between .re2 and .co2
    vtx, xyz, rec = M.periodic_box(2, 2, 2, periodic=(False, False, False))
    write_re2(tmp_path / "m.re2", xyz)
    with pytest.raises(pa.ParrsbError, match="don't match"):
        pa.read_mesh(name, read=3)


def test_ma2_layout(tmp_path):
    vl = np.arange(1, 1 + 5 * 4, dtype=np.int64).reshape(5, 4)
    part = np.array([2, 0, 1, 0, 2], dtype=np.int32)
    name = str(tmp_path / "q")
    pa.dump_map(name, vl)
    raw = open(name + ".ma2", "rb").read()
    # version, nelgt, nactive, depth = floor(log2 5), 2^depth, npts, nrnk, noutflow
    hdr = ("%5s%12d%12d%12d%12d%12d%12d%12d" % ("#v001", 5, 5, 2, 4, 20, 5, 0)).ljust(132)
    body = b"".join(struct.pack("<5i", 0, *vl[e]) for e in range(5))
    assert raw == hdr.encode() + struct.pack("<f", 6.54321) + body
    pa.dump_map(name, vl, part)
    raw = open(name + ".ma2", "rb").read()
    body = b"".join(struct.pack("<5i", part[e], *vl[e]) for e in range(5))
    assert raw[136:] == body


def test_dist_mesh_groups_by_part():
    rng = np.random.default_rng(3)
    n = 50
    vl = rng.integers(1, 99, size=(n, 8)).astype(np.int64)
    xyz = rng.uniform(size=(n, 8, 3))
    part = rng.integers(0, 6, size=n).astype(np.int32)
    v2, x2, p2 = pa.dist_mesh(vl, xyz, part)
    order = np.argsort(part, kind="stable")
    assert np.array_equal(v2, vl[order]) and np.array_equal(x2, xyz[order])
    assert np.array_equal(p2, part[order])
    with pytest.raises(pa.ParrsbError):
        pa.dist_mesh(vl, xyz, np.full(n, -1, np.int32))


def test_cmd_line_options():
    import ctypes as C
    args = [b"genmap", b"--mesh=abc", b"--tol=0.05", b"--test", b"--dump", b"--nactive=3",
            b"--verbose=2"]
    argv = (C.c_char_p * len(args))(*args)
    o = pa.lib().parrsb_parse_cmd_opts(len(args), argv)
    assert o and o.contents.mesh == b"abc" and abs(o.contents.tol - 0.05) < 1e-15
    assert (o.contents.test, o.contents.dump, o.contents.nactive, o.contents.verbose) == \
        (1, 1, 3, 2)
    pa.lib().parrsb_cmd_opts_free(o)
    args = [b"genmap", b"--tol=0.1"]
    argv = (C.c_char_p * len(args))(*args)
    assert not pa.lib().parrsb_parse_cmd_opts(len(args), argv)


@pytest.mark.gpu
def test_genmap_flow_on_periodic_re2(tmp_path, oracle):
    """read .re2 -> connectivity with periodic faces -> partition ->
    redistribute -> .ma2, as examples/genmap.c does; the connectivity equals
    the oracle's and the partition lowers the communication volume of the
    initial even distribution (genmap.c:75-78)."""
    vtx, xyz, rec = M.periodic_box(12, 8, 6, periodic=(True, True, False), warp=0.1)
    write_re2(tmp_path / "pb.re2", xyz, rec, [(1, 5, "W")])
    m = pa.read_mesh(str(tmp_path / "pb"), read=1)
    vl = pa.conn_mesh(m["coord"], 3, 0.2, periodic=m["bcs"])
    rc, ref = oracle.conn_mesh(m["coord"], 3, 0.2, m["bcs"])
    assert rc == 0 and np.array_equal(vl, ref)
    P = 8
    vl_pre = vl[:, PRE_TO_SYM]  # back to the preprocessor corner order of coord
    part, _ = pa.part_mesh(vl_pre, m["coord"], 8, P)
    nel = vl.shape[0]
    even = np.repeat(np.arange(P), [nel // P + (r >= P - nel % P) for r in range(P)])
    before = pa.quality(even.astype(np.int32), vl_pre, 8, P)["comm_volume"]
    after = pa.quality(part, vl_pre, 8, P)["comm_volume"]
    assert after <= before
    v2, x2, p2 = pa.dist_mesh(vl_pre, m["coord"], part)
    pa.dump_map(str(tmp_path / "pb"), v2, p2)
    raw = open(tmp_path / "pb.ma2", "rb").read()
    ranks = np.frombuffer(raw[136:], dtype="<i4").reshape(nel, 9)[:, 0]
    assert np.array_equal(ranks, np.sort(part))


@pytest.mark.gpu
def test_setup_mesh_on_ethier(tmp_path, monkeypatch):
    """parrsb_setup_mesh: read, connectivity, partition (PARRSB_NPARTS parts),
    redistribution of the example mesh."""
    import ctypes as C
    monkeypatch.setenv("PARRSB_NPARTS", "4")
    monkeypatch.setenv("PARRSB_VERBOSE_LEVEL", "0")
    o = pa.CmdOpts(mesh=ETHIER.encode(), tol=0.2)
    nel, nv = C.c_uint(), C.c_uint()
    vl, coord = C.c_void_p(), C.c_void_p()
    rc = pa.lib().parrsb_setup_mesh(C.byref(nel), C.byref(nv), C.byref(vl), C.byref(coord),
                                    C.byref(o), 0)
    assert rc == 0, pa.last_error()
    assert nel.value == 32 and nv.value == 8
    v = np.ctypeslib.as_array(C.cast(vl, C.POINTER(C.c_longlong)), (256,)).copy()
    pa.lib().parrsb_amd_free(vl), pa.lib().parrsb_amd_free(coord)
    assert v.min() == 1 and v.max() == 53
