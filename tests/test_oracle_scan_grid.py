"""Oracle pinning: exclusive scan and uniform-grid counting sort.

Checks: the reference's scan KATs (tests/golden/scan_kat.json), the
reference's Blelloch tree restated (power-of-two sizes) against the serial
scan, numpy cumsum on ragged sizes, grid-build invariants (permutation,
sorted cells, stable order, offsets = scan(counts)), the closed-form cell
hash on edge cases, and neighbour sets through the grid against brute force.
"""
import json
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).with_name("golden") / "scan_kat.json"


def test_scan_kat(oracle):
    kat = json.loads(GOLD.read_text())
    for case in kat.values():
        x = np.array(case["input"], np.int32)
        exp = case["expected_exclusive"]
        assert oracle.exclusive_scan(x)[:-1].tolist() == exp
        assert oracle.exclusive_scan(x)[-1] == x.sum()
        assert oracle.blelloch_scan(x).tolist() == exp


@pytest.mark.parametrize("n", [0, 1, 2, 3, 7, 100, 2047, 2048, 2049, 100_003])
def test_scan_random_sizes(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 50, n).astype(np.int32)
    out = oracle.exclusive_scan(x)
    ref = np.concatenate([[0], np.cumsum(x)]).astype(np.int64)
    assert out.tolist() == ref.tolist()


@pytest.mark.parametrize("n", [2, 4, 16, 1024, 1 << 14])
def test_blelloch_matches_serial(oracle, n):
    rng = np.random.default_rng(n + 1)
    x = rng.integers(0, 9, n).astype(np.int32)
    assert oracle.blelloch_scan(x).tolist() == oracle.exclusive_scan(x)[:-1].tolist()


def _params(cwa, dims=(6, 5, 4), cell=0.1, gmin=(0.0, 0.0, 0.0), h=None):
    p = cwa.reference_params()
    p.cell_size = cell
    p.h = cell if h is None else h
    for a in range(3):
        p.grid_min[a] = gmin[a]
        p.grid_dims[a] = dims[a]
    return p


def _check_grid_invariants(g, n, nc, key_ref):
    key, count, offset, lst = g["key"], g["count"], g["offset"], g["list"]
    assert key.tolist() == key_ref.tolist()
    assert count.sum() == n
    assert offset.tolist() == np.concatenate([[0], np.cumsum(count)]).tolist()
    assert sorted(lst.tolist()) == list(range(n))          # permutation
    ks = key[lst]
    assert np.all(np.diff(ks) >= 0)                         # cells non-decreasing
    same = np.diff(ks) == 0
    assert np.all(np.diff(lst)[same] > 0)                   # stable: index ascending
    for c in np.unique(key):
        assert offset[c + 1] - offset[c] == np.count_nonzero(key == c)


TILE_Z = 16


def _column(ix, iz, nx, nz):
    """Memory column of cell column (ix, iz): z-tiles of 16 columns, x-major
    within a tile, the last tile narrower when 16 does not divide nz."""
    t = iz // TILE_Z
    w = np.minimum(TILE_Z, nz - t * TILE_Z)
    return t * TILE_Z * nx + ix * w + (iz - t * TILE_Z)


def _hash_ref(pos, p):
    """Closed-form cell hash, y fastest: column(ix, iz)*ny + iy."""
    d = np.array(p.grid_dims)
    cell = np.float32(p.cell_size)
    q = np.floor((pos[:, :3] - np.array(p.grid_min, np.float32)) / cell)
    q = np.clip(q, 0, d - 1).astype(np.int64)
    return (_column(q[:, 0], q[:, 2], d[0], d[2]) * d[1] + q[:, 1]).astype(np.int32)


@pytest.mark.parametrize("nx,nz", [(1, 1), (3, 2), (5, 16), (4, 17), (7, 37), (2, 48)])
def test_column_order_is_a_tiled_bijection(nx, nz):
    """The column order covers 0..nx*nz-1 once; within a tile the x-major
    order holds (consecutive columns step z inside the tile), and tiles
    follow each other: every column of tile t comes before tile t+1."""
    ix, iz = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    col = _column(ix, iz, nx, nz)
    assert sorted(col.ravel().tolist()) == list(range(nx * nz))
    t = iz // TILE_Z
    for a in range(nx):
        for b in range(nz - 1):
            if t[a, b] == t[a, b + 1]:
                assert col[a, b + 1] == col[a, b] + 1
    for tt in range(t.max()):
        assert col[t == tt].max() < col[t == tt + 1].min()


def test_cell_hash_tiled_columns(oracle, cwa):
    """Keys of a grid with 37 columns along z (tiles of 16, 16 and 5) and
    14 along x, spelled out at tile edges and checked on random points."""
    p = _params(cwa, dims=(14, 3, 37), cell=0.1, gmin=(0.0, 0.0, 0.0))
    pts = np.array([[0.05, 0.05, 0.05],      # (0, 0, 0)
                    [0.05, 0.05, 1.55],      # (0, 0, 15): last column of tile 0
                    [0.05, 0.05, 1.65],      # (0, 0, 16): first column of tile 1
                    [0.15, 0.25, 1.65],      # (1, 2, 16)
                    [1.35, 0.05, 3.65]],     # (13, 0, 36): the last column
                   np.float32)
    pos = np.concatenate([pts, np.ones((len(pts), 1), np.float32)], 1)
    g = oracle.grid_build(p, pos)
    ny = 3
    assert g["key"].tolist() == [0, 15 * ny, (16 * 14) * ny, (16 * 14 + 16 + 0) * ny + 2,
                                 (2 * 16 * 14 + 13 * 5 + 4) * ny]
    rng = np.random.default_rng(5)
    pos = np.zeros((3000, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.1, 3.8, (3000, 3)).astype(np.float32)
    g = oracle.grid_build(p, pos)
    _check_grid_invariants(g, 3000, 14 * 3 * 37, _hash_ref(pos, p))


@pytest.mark.parametrize("n", [0, 1, 5, 257, 5000])
def test_grid_build_invariants(oracle, cwa, n):
    rng = np.random.default_rng(7 + n)
    p = _params(cwa)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.1, 0.7, (n, 3)).astype(np.float32)   # some outside the grid
    g = oracle.grid_build(p, pos)
    nc = 6 * 5 * 4
    _check_grid_invariants(g, n, nc, _hash_ref(pos, p))


def test_cell_hash_edge_cases(oracle, cwa):
    p = _params(cwa, dims=(4, 3, 2), cell=0.25, gmin=(-0.5, 0.0, 0.0))
    pts = np.array([
        [-0.5, 0.0, 0.0],       # on the min corner
        [-0.25, 0.25, 0.25],    # exactly on interior faces
        [-0.5000001, -1e-8, -3.0],  # just below min (negative floor, clamped)
        [0.5, 0.75, 0.5],       # on the max faces (clamped to last cell)
        [10.0, 10.0, 10.0],     # far outside
        [-0.2, 0.1, 0.3],
    ], np.float32)
    pos = np.concatenate([pts, np.ones((len(pts), 1), np.float32)], 1)
    g = oracle.grid_build(p, pos)
    assert g["key"].tolist() == _hash_ref(pos, p).tolist()
    # spelled out: (ix, iy, iz) -> (ix*nz + iz)*ny + iy with nz=2, ny=3 (one
    # z-tile narrower than 16 columns)
    lin = lambda ix, iy, iz: (ix * 2 + iz) * 3 + iy  # noqa: E731
    assert g["key"].tolist() == [lin(0, 0, 0), lin(1, 1, 1), lin(0, 0, 0), lin(3, 2, 1),
                                 lin(3, 2, 1), lin(1, 0, 1)]


@pytest.mark.parametrize("ydiv", [1, 3, 4])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_neighbor_sets_grid_equals_all_pairs(oracle, cwa, seed, ydiv):
    """The 27-cell search returns exactly the all-pairs set r < h
    (rho_pres_comp.glsl:60-64), including cells of size > h, y rows of
    cell/ydiv, and points outside the grid (clamped into edge cells)."""
    rng = np.random.default_rng(seed)
    n = 3000
    h = 0.05
    p = _params(cwa, dims=(8, 9 * ydiv, 7), cell=0.06 + 0.01 * seed, h=h)
    p.grid_ydiv = ydiv
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.05, 0.55, (n, 3)).astype(np.float32)
    # near-duplicates at distance ~h to exercise the strict comparison
    pos[1:200:2, :3] = pos[0:199:2, :3] + np.float32(h) * np.array([1, 0, 0], np.float32)
    a = oracle.neighbor_counts(p, pos, all_pairs=False)
    b = oracle.neighbor_counts(p, pos, all_pairs=True)
    assert a.tolist() == b.tolist()
    assert a.min() >= 1  # self included


def test_neighbor_sets_across_z_tiles(oracle, cwa):
    """The grid search stays exact where neighbour columns lie in another
    z-tile of the column order: 37 columns along z (tiles of 16, 16, 5)."""
    rng = np.random.default_rng(9)
    n = 4000
    h = 0.03
    p = _params(cwa, dims=(6, 8, 37), cell=0.032, h=h)
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = rng.uniform([0.0, 0.0, 0.0], [0.19, 0.25, 1.18], (n, 3)).astype(np.float32)
    a = oracle.neighbor_counts(p, pos, all_pairs=False)
    b = oracle.neighbor_counts(p, pos, all_pairs=True)
    assert a.tolist() == b.tolist()
    assert a.mean() > 3


# ---------------------------------------------------------------- box grid
def _box_cases():
    """(name, centres, half-widths, grid_min, cell, dims)."""
    rng = np.random.default_rng(21)
    gmin, dims = (-1.0, -1.0, -1.0), (16, 16, 16)
    cell = (2.0 / 16, 2.0 / 16, 2.0 / 16)
    out = []
    n = 4000
    x = np.zeros((n, 4), np.float32)
    x[:, :3] = rng.uniform(-0.95, 0.95, (n, 3))
    hw = np.zeros((n, 4), np.float32)
    hw[:, :3] = rng.uniform(0.0, 0.5 * cell[0], (n, 3))    # the reference's <= 8 cells
    out.append(("small", x, hw, gmin, cell, dims))
    for span in (1, 2, 4, 8):                                # boxes over 1, 2, 4, 8 cells
        m = 500
        xs = np.zeros((m, 4), np.float32)
        xs[:, :3] = rng.uniform(-0.9, 0.9, (m, 3))
        hs = np.zeros((m, 4), np.float32)
        hs[:, :3] = 0.5 * span * cell[0] * rng.uniform(0.3, 0.95, (m, 3))
        out.append((f"span{span}", xs, hs, gmin, cell, dims))
    m = 300                                                  # partly or fully outside
    xs = np.zeros((m, 4), np.float32)
    xs[:, :3] = rng.uniform(-1.6, 1.6, (m, 3))
    hs = np.zeros((m, 4), np.float32)
    hs[:, :3] = rng.uniform(0.0, 0.4, (m, 3))
    out.append(("edges", xs, hs, gmin, cell, dims))
    xs = np.zeros((6, 4), np.float32)                        # inverted / degenerate
    xs[:, :3] = [[0.1, 0.2, 0.3], [0.0, 0.0, 0.0], [-0.5, 0.7, 0.1], [0.3, 0.3, 0.3],
                 [0.9, -0.9, 0.5], [5.0, 5.0, 5.0]]
    hs = np.zeros((6, 4), np.float32)
    hs[:, :3] = [[-0.2, 0.1, 0.1], [0.0, 0.0, 0.0], [0.05, -0.3, 0.05], [0.3, 0.3, 0.3],
                 [0.0, 0.0, 0.0], [0.1, 0.1, 0.1]]
    out.append(("degenerate", xs, hs, gmin, cell, dims))
    m = 2000                                                 # non-cubic grid, cells per axis
    xs = np.zeros((m, 4), np.float32)
    xs[:, :3] = rng.uniform([0, 0, 0], [3.0, 1.0, 2.0], (m, 3))
    hs = np.zeros((m, 4), np.float32)
    hs[:, :3] = rng.uniform(0.0, 0.15, (m, 3))
    out.append(("aniso", xs, hs, (0.0, 0.0, 0.0), (0.125, 0.05, 0.1), (24, 20, 20)))
    out.append(("empty", np.zeros((0, 4), np.float32), np.zeros((0, 4), np.float32), gmin,
                cell, dims))
    return out


def _box_brute_force(x, hw, gmin, cell, dims):
    """Per cell, the boxes whose closed interval [x - hw, x + hw] (corners
    rounded to fp32) meets the cell's half-open interval on every axis; the
    edge cells reach to infinity (points outside the grid clamp into them).
    Ascending box order within each cell."""
    n = x.shape[0]
    a = (x[:, :3] - hw[:, :3]).astype(np.float64)
    b = (x[:, :3] + hw[:, :3]).astype(np.float64)
    per_axis = []
    for ax in range(3):
        i = np.arange(dims[ax])
        lo = np.float64(np.float32(gmin[ax])) + i * np.float64(np.float32(cell[ax]))
        hi = lo + np.float64(np.float32(cell[ax]))
        lo[0], hi[-1] = -np.inf, np.inf
        per_axis.append((a[:, ax:ax + 1] < hi[None, :]) & (b[:, ax:ax + 1] >= lo[None, :]))
    cells = {}
    for k in range(n):
        for ix in np.nonzero(per_axis[0][k])[0]:
            for iy in np.nonzero(per_axis[1][k])[0]:
                for iz in np.nonzero(per_axis[2][k])[0]:
                    c = (ix * dims[2] + iz) * dims[1] + iy
                    cells.setdefault(c, []).append(k)
    nc = dims[0] * dims[1] * dims[2]
    count = np.zeros(nc, np.int64)
    lst = []
    for c in range(nc):
        count[c] = len(cells.get(c, []))
        lst += cells.get(c, [])
    return count, np.array(lst, np.int64)


@pytest.mark.parametrize("case", [c[0] for c in _box_cases()])
def test_box_grid_against_brute_force(oracle, case):
    """The oracle's box grid (UniformGridPbdGpu3D: ComputeCounts / InsertBox
    over ComputeCellIndex(x -+ hw)) equals a brute-force overlap of every box
    with every cell, and lists the boxes of a cell in ascending order."""
    name, x, hw, gmin, cell, dims = next(c for c in _box_cases() if c[0] == case)
    g = oracle.grid_build_boxes(x, hw, gmin, cell, dims)
    count, lst = _box_brute_force(x, hw, gmin, cell, dims)
    assert g["count"].tolist() == count.tolist()
    assert g["offset"].tolist() == np.concatenate([[0], np.cumsum(count)]).tolist()
    assert g["total"] == len(lst) and g["list"].tolist() == lst.tolist()
    if name.startswith("span"):
        span = int(name[4:])
        per_box = np.bincount(lst, minlength=x.shape[0])
        assert per_box.max() <= (span + 1) ** 3 and per_box.min() >= 1
