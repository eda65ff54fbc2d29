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
