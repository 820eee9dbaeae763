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


def _hash_ref(pos, p):
    """Closed-form cell hash, y fastest: ((ix*nz + iz)*ny + iy)."""
    d = np.array(p.grid_dims)
    cell = np.float32(p.cell_size)
    q = np.floor((pos[:, :3] - np.array(p.grid_min, np.float32)) / cell)
    q = np.clip(q, 0, d - 1).astype(np.int64)
    return ((q[:, 0] * d[2] + q[:, 2]) * d[1] + q[:, 1]).astype(np.int32)


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
    # spelled out: (ix, iy, iz) -> (ix*nz + iz)*ny + iy with nz=2, ny=3
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
