"""Seeded synthetic meshes in the layout parrsb_part_mesh takes.

Every generator returns ``(vtx, xyz, nv)``:
  vtx  int64 [nel, nv]        global vertex ids, 1-based (id 0 means "no
                              connection" to the gather-scatter the reference
                              builds on, so ids start at 1)
  xyz  float64 [nel, nv, ndim] per-vertex coordinates (ndim = 3 for hexes,
                              2 for quads), as parrsb_part_mesh reads them
                              (the reference forms centroids itself,
                              parrsb.c:139-145)
Elements are numbered with x fastest.  Hex corners are listed in the order
(0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1) (1,0,1) (1,1,1) (0,1,1).

Workloads used by the tests and the benchmark:
  box_hex(16, 16, 16, jitter=0.1, seed=1)      4,096 hexes, 8 parts
  box_hex(128, 128, 64, jitter=0.1, seed=2)    1,048,576 hexes, 1,024 parts
  holed_quads(2048, 2048, seed=3)              ~3.98M quads, 4,096 parts
  warped_box_hex(256, seed=5)                  16,777,216 hexes, 65,536 parts
  pipe_hex(16, 128, 1024, seed=7)              2,097,152 hexes, 2,048 parts
                                               (held-out learned-start mesh)
  warm_start_mesh(i)                           20k-200k hexes, the training set
                                               of the learned warm start
"""
from __future__ import annotations

import numpy as np

_HEX = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)
_QUAD = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.int64)


def _hex_vtx(nx: int, ny: int, nz: int) -> np.ndarray:
    i = np.arange(nx, dtype=np.int64)
    j = np.arange(ny, dtype=np.int64)
    k = np.arange(nz, dtype=np.int64)
    kk, jj, ii = np.meshgrid(k, j, i, indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    sx, sy = nx + 1, (nx + 1) * (ny + 1)
    vtx = np.empty((ii.size, 8), dtype=np.int64)
    for c, (dx, dy, dz) in enumerate(_HEX):
        vtx[:, c] = 1 + (ii + dx) + sx * (jj + dy) + sy * (kk + dz)
    return vtx


def box_hex(nx: int, ny: int, nz: int, jitter: float = 0.1, seed: int = 1,
            h: float | None = None):
    """Structured hex box on [0, nx h] x [0, ny h] x [0, nz h] with h = 1/max(n).

    Each element's 8 lattice corners are shifted rigidly by a seeded
    U[-jitter h, jitter h)^3 offset, so its centroid is the lattice centroid
    plus that offset (the vertex ids stay those of the lattice)."""
    if h is None:
        h = 1.0 / max(nx, ny, nz)
    vtx = _hex_vtx(nx, ny, nz)
    nel = nx * ny * nz
    sx, sy = nx + 1, (nx + 1) * (ny + 1)
    v0 = vtx - 1
    lat = np.stack([v0 % sx, (v0 // sx) % (ny + 1), v0 // sy], axis=-1).astype(np.float64) * h
    if jitter:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter * h, jitter * h, size=(nel, 1, 3))
        lat += d
    return vtx, lat, 8


def warped_box_hex(n: int = 256, seed: int = 5, amp: float = 0.05):
    """n^3 hex box whose lattice is moved by a seeded smooth map
    X_d = x_d + sum_e a_de sin(2 pi k_de x_e + phi_de), a_de <= amp, k in {1,2}.
    The map acts on vertices, so shared corners keep one position."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.0, amp, size=(3, 3))
    kf = rng.integers(1, 3, size=(3, 3)).astype(np.float64)
    ph = rng.uniform(0.0, 2 * np.pi, size=(3, 3))
    h = 1.0 / n
    m = n + 1
    g = np.arange(m, dtype=np.float64) * h
    vtx = _hex_vtx(n, n, n)
    # vertex positions: separable sums, evaluated per vertex
    nvert = m ** 3
    vid = np.arange(nvert, dtype=np.int64)
    x = np.stack([g[vid % m], g[(vid // m) % m], g[vid // (m * m)]], axis=-1)
    X = x.copy()
    for d in range(3):
        for e in range(3):
            X[:, d] += a[d, e] * np.sin(2 * np.pi * kf[d, e] * x[:, e] + ph[d, e])
    del x
    xyz = X[vtx - 1]
    return vtx, xyz, 8


def _lattice(nx: int, ny: int, nz: int, wrap_y: bool = False):
    """Hex connectivity of an nx x ny x nz lattice and the lattice index
    (i, j, k) of every vertex id.  With wrap_y the j direction is periodic
    (j = ny is j = 0), as around a pipe."""
    my = ny if wrap_y else ny + 1
    i = np.arange(nx, dtype=np.int64)
    j = np.arange(ny, dtype=np.int64)
    k = np.arange(nz, dtype=np.int64)
    kk, jj, ii = np.meshgrid(k, j, i, indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    sx, sy = nx + 1, (nx + 1) * my
    vtx = np.empty((ii.size, 8), dtype=np.int64)
    for c, (dx, dy, dz) in enumerate(_HEX):
        vtx[:, c] = 1 + (ii + dx) + sx * ((jj + dy) % my) + sy * (kk + dz)
    nvert = sx * my * (nz + 1)
    vid = np.arange(nvert, dtype=np.int64)
    ijk = np.stack([vid % sx, (vid // sx) % my, vid // sy], axis=-1).astype(np.float64)
    return vtx, ijk


def _warp(X, rng, amp, size):
    """X_d += size * sum_e a_de sin(2 pi k_de u_e + phi_de), u = X scaled to
    [0, 1] per axis, a_de in [0, amp], k_de in {1, 2}."""
    lo, hi = X.min(0), X.max(0)
    u = (X - lo) / np.where(hi > lo, hi - lo, 1.0)
    a = rng.uniform(0.0, amp, size=(3, 3))
    kf = rng.integers(1, 3, size=(3, 3)).astype(np.float64)
    ph = rng.uniform(0.0, 2 * np.pi, size=(3, 3))
    Y = X.copy()
    for d in range(3):
        for e in range(3):
            Y[:, d] += size * a[d, e] * np.sin(2 * np.pi * kf[d, e] * u[:, e] + ph[d, e])
    return Y


def pipe_hex(nr: int, nt: int, nz: int, r_in: float = 0.5, r_out: float = 1.0,
             length: float | None = None, seed: int = 7, amp: float = 0.05):
    """Annular pipe section: nr cells across the wall, nt around (periodic),
    nz along the axis; length defaults to the one that makes the cells about
    as long as they are wide around the mean radius.  Vertex positions are
    warped by a seeded sinusoidal map of amplitude <= amp times the outer
    diameter."""
    rng = np.random.default_rng(seed)
    if length is None:
        length = nz * 2 * np.pi * 0.5 * (r_in + r_out) / nt
    vtx, ijk = _lattice(nr, nt, nz, wrap_y=True)
    r = r_in + (r_out - r_in) * ijk[:, 0] / nr
    th = 2 * np.pi * ijk[:, 1] / nt
    X = np.stack([r * np.cos(th), r * np.sin(th), length * ijk[:, 2] / nz], axis=-1)
    X = _warp(X, rng, amp, 2 * r_out)
    return vtx, X[vtx - 1], 8


def channel_hex(nx: int, ny: int, nz: int, width: float = 1.0, bend: float = 1.0,
                seed: int = 11, amp: float = 0.05):
    """Rectangular channel of nx cells along its centreline (cells square in
    plan), ny across and nz deep, bent through `bend` radians along a circular
    arc, then warped like pipe_hex."""
    rng = np.random.default_rng(seed)
    vtx, ijk = _lattice(nx, ny, nz)
    h = width / ny
    s = ijk[:, 0] * h
    u = ijk[:, 1] * h - 0.5 * width
    w = ijk[:, 2] * h
    L = nx * h
    if abs(bend) > 1e-9:
        R = L / bend
        X = np.stack([(R + u) * np.sin(s / R), R - (R + u) * np.cos(s / R), w], axis=-1)
    else:
        X = np.stack([s, u, w], axis=-1)
    X = _warp(X, rng, amp, width)
    return vtx, X[vtx - 1], 8


def warm_start_mesh(i: int):
    """Mesh i of the learned warm start's training set: a seeded annular pipe
    section or bent channel of 20k-200k hexes, aspect ratio 1-16, warped with
    amplitude <= 0.2.  Returns (vtx, xyz, nv, info)."""
    rng = np.random.default_rng(1000003 * (i + 1))
    target = int(np.exp(rng.uniform(np.log(20000), np.log(200000))))
    aspect = float(np.exp(rng.uniform(0.0, np.log(16.0))))
    amp = float(rng.uniform(0.0, 0.2))
    seed = int(rng.integers(1 << 30))
    if rng.uniform() < 0.5:
        # pipe: length / outer diameter = aspect; cells ~cubic
        r_in = float(rng.uniform(0.3, 0.8))
        nr_over_nt = (1 - r_in) / (np.pi * (1 + r_in))
        nz_over_nt = 2 * aspect / (np.pi * (1 + r_in))
        nt = max(8, int(round((target / (nr_over_nt * nz_over_nt)) ** (1 / 3))))
        nr = max(1, int(round(nr_over_nt * nt)))
        nz = max(1, int(round(nz_over_nt * nt)))
        vtx, xyz, nv = pipe_hex(nr, nt, nz, r_in=r_in, r_out=1.0, length=2.0 * aspect,
                                seed=seed, amp=amp)
        info = dict(kind="pipe", nr=nr, nt=nt, nz=nz, r_in=r_in, aspect=aspect, amp=amp)
    else:
        depth = float(rng.uniform(0.2, 1.0))  # depth / width
        ny = max(2, int(round((target / (aspect * depth)) ** (1 / 3))))
        nx, nz = max(1, int(round(aspect * ny))), max(1, int(round(depth * ny)))
        bend = float(rng.uniform(0.0, np.pi))
        vtx, xyz, nv = channel_hex(nx, ny, nz, bend=bend, seed=seed, amp=amp)
        info = dict(kind="channel", nx=nx, ny=ny, nz=nz, bend=bend, aspect=aspect, amp=amp)
    info["nel"] = int(vtx.shape[0])
    return vtx, xyz, nv, info


def quad_grid(nx: int, ny: int, h: float | None = None):
    """Structured nx x ny quad grid, vid = 1 + x + (nx+1) y, no jitter."""
    if h is None:
        h = 1.0 / max(nx, ny)
    i = np.arange(nx, dtype=np.int64)
    j = np.arange(ny, dtype=np.int64)
    jj, ii = np.meshgrid(j, i, indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    vtx = np.empty((ii.size, 4), dtype=np.int64)
    xyz = np.empty((ii.size, 4, 2), dtype=np.float64)
    for c, (dx, dy) in enumerate(_QUAD):
        vtx[:, c] = 1 + (ii + dx) + (nx + 1) * (jj + dy)
        xyz[:, c, 0] = (ii + dx) * h
        xyz[:, c, 1] = (jj + dy) * h
    return vtx, xyz, 4


def holed_quads(nx: int = 2048, ny: int = 2048, seed: int = 3, frac: float = 0.05,
                rmin: float = 2.0, rmax: float = 40.0):
    """Quad grid with elements removed inside seeded random disks (radius
    U[rmin, rmax] cells, uniform centres) until at least `frac` of the area is
    gone.  The surviving elements keep their lattice vertex ids (not dense)
    and their order; the mesh splits into components of uneven size."""
    rng = np.random.default_rng(seed)
    keep = np.ones((ny, nx), dtype=bool)
    cy, cx = np.mgrid[0:ny, 0:nx].astype(np.float64) + 0.5
    target = int(np.ceil(frac * nx * ny))
    removed = 0
    while removed < target:
        r = rng.uniform(rmin, rmax)
        x0, y0 = rng.uniform(0, nx), rng.uniform(0, ny)
        ylo, yhi = max(int(y0 - r) - 1, 0), min(int(y0 + r) + 2, ny)
        xlo, xhi = max(int(x0 - r) - 1, 0), min(int(x0 + r) + 2, nx)
        sub = (cx[ylo:yhi, xlo:xhi] - x0) ** 2 + (cy[ylo:yhi, xlo:xhi] - y0) ** 2 <= r * r
        kept = keep[ylo:yhi, xlo:xhi]
        removed += int(np.count_nonzero(sub & kept))
        kept &= ~sub
    vtx, xyz, nv = quad_grid(nx, ny)
    sel = keep.ravel()
    return np.ascontiguousarray(vtx[sel]), np.ascontiguousarray(xyz[sel]), nv


def ncuts(nparts: int) -> int:
    """Number of bisection levels, ceil(log2 P) (rsb.c:253-255)."""
    c = 0
    while (1 << c) < nparts:
        c += 1
    return c


def periodic_box(nx: int, ny: int, nz: int = 0, periodic=(True, False, False),
                 warp: float = 0.0, seed: int = 1):
    """Structured hex box (nz > 0) or quad grid (nz == 0) whose faces on the
    chosen axes are periodic.

    Returns (vtx, xyz, records): vtx[nel, nv] are the lattice ids with the
    periodic wrap applied (what a connectivity must find), xyz[nel, nv, ndim]
    the corner coordinates in preprocessor order (not wrapped), and
    records[nrec, 4] the periodic face records (element, face, partner
    element, partner face), 1-based, preprocessor face numbers (hex: 1 y=0,
    2 x=1, 3 y=1, 4 x=0, 5 z=0, 6 z=1; quad: 1 y=0, 2 x=1, 3 y=1, 4 x=0), one
    record per face as a .re2 file lists them.  warp moves the interior
    lattice points smoothly (periodic boundaries stay planar)."""
    ndim = 3 if nz > 0 else 2
    n = (nx, ny, nz)[:ndim]
    corners = _HEX if ndim == 3 else _QUAD
    grids = np.meshgrid(*[np.arange(m, dtype=np.int64) for m in reversed(n)], indexing="ij")
    idx = [g.ravel() for g in reversed(grids)]  # idx[0] = i (fastest)
    nel = idx[0].size
    h = 1.0 / max(n)
    lat = np.stack([np.stack([idx[d] + corners[c, d] for d in range(ndim)], -1)
                    for c in range(len(corners))], 1)  # [nel, nv, ndim] lattice ints
    x = lat.astype(np.float64)
    if warp:
        rng = np.random.default_rng(seed)
        for d in range(ndim):
            ph = rng.uniform(0, 2 * np.pi, size=ndim)
            s = np.ones(lat.shape[:2])
            for e in range(ndim):
                s = s * np.sin(np.pi * lat[..., e] / n[e])
            x[..., d] += warp * s * np.cos(ph[d])
    xyz = x * h
    wrap = lat.copy()
    for d in range(ndim):
        if periodic[d]:
            wrap[..., d] %= n[d]
    m = [n[d] + (0 if periodic[d] else 1) for d in range(ndim)]
    vid = wrap[..., 0] + m[0] * wrap[..., 1]
    if ndim == 3:
        vid = vid + m[0] * m[1] * wrap[..., 2]
    vtx = 1 + vid
    # faces: (axis, low face, high face) in preprocessor numbering
    faces = {0: (4, 2), 1: (1, 3), 2: (5, 6)}
    recs = []
    stride = [1, n[0], n[0] * (n[1] if ndim > 1 else 1)]
    for d in range(ndim):
        if not periodic[d]:
            continue
        lo, hi = faces[d]
        low = np.flatnonzero(idx[d] == 0)
        for e0 in low:
            e1 = e0 + (n[d] - 1) * stride[d]
            recs.append((e0 + 1, lo, e1 + 1, hi))
            recs.append((e1 + 1, hi, e0 + 1, lo))
    rec = np.array(recs, dtype=np.int64).reshape(-1, 4)
    return vtx.astype(np.int64), xyz, rec
