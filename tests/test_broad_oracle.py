"""Pinning the broad-phase oracle (CPU only).

The reference's grid (visualization/integrate_final_gjk.cu:467-570, 916-1002)
has no tests or golden vectors of its own, so the oracle is pinned by what the
mathematics fixes:
  * the pair set equals an all-pairs brute force over the same FP32 test
    whenever cells are at least one largest diameter wide (the reference's
    sizing, :813-819), including objects clamped into the border cells;
  * the order is by first index, then neighbour cell (z, y, x from -1 to +1),
    then second index, recomputed here from the cell definition;
  * hand cases: tangency is not overlap (strict <), the per-cell cap keeps the
    lowest indices, the max_pairs clamp keeps a prefix, NaN never pairs;
  * sub-mesh expansion: a hand-written example, the identity layout, counts.
"""
import numpy as np
import pytest

import oracle
import ogjk
from ogjk import gen


def brute_pairs(s):
    """All (i, j), i < j, with ((dx*dx + dy*dy) + dz*dz) < (ri + rj)^2 in FP32."""
    s = np.asarray(s, np.float32)
    n = len(s)
    i, j = np.triu_indices(n, 1)
    d = s[i, :3] - s[j, :3]
    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    rs = s[i, 3] + s[j, 3]
    m = d2 < rs * rs
    return set(zip(i[m].tolist(), j[m].tolist()))


def cells_of(s, grid):
    f = np.floor((s[:, :3] + np.float32(grid.boundary)) / np.float32(grid.cell_size))
    f = np.where(np.isnan(f), 0, f)
    return np.clip(f, 0, grid.grid_size - 1).astype(np.int64)


def check_order(pairs, s, grid):
    """First index non-decreasing; within one, (neighbour rank, j) increasing."""
    if len(pairs) == 0:
        return
    i, j = pairs[:, 0].astype(np.int64), pairs[:, 1].astype(np.int64)
    assert np.all(i < j)
    c = cells_of(s, grid)
    d = c[j] - c[i]
    assert np.all(np.abs(d) <= 1)
    rank = (d[:, 2] + 1) * 9 + (d[:, 1] + 1) * 3 + (d[:, 0] + 1)
    key = (i * 27 + rank) * (len(s) + 1) + j
    assert np.all(np.diff(key) > 0)


@pytest.mark.parametrize("n,box,rmax,outside,max_grid", [
    (1500, 8.0, 0.6, 0.0, 30), (3000, 12.0, 1.0, 0.1, 30), (2000, 5.0, 0.3, 0.2, 64),
    (800, 3.0, 1.5, 0.0, 4)])
def test_oracle_broad_phase_equals_brute_force(n, box, rmax, outside, max_grid):
    s = gen.spheres(n, seed=n, box=box, rmin=0.1, rmax=rmax, outside=outside)
    grid = ogjk.Grid.for_spheres(float(s[:, 3].max()), box, max_grid)
    pairs, found = oracle.broad_phase(s, grid.cell_size, grid.boundary, grid.grid_size)
    assert found == len(pairs) > 0
    assert set(map(tuple, pairs.tolist())) == brute_pairs(s)
    check_order(pairs, s, grid)


def test_oracle_tangent_spheres_do_not_pair():
    # centres 1 apart on each axis, radii 0.5: d^2 == (r_i + r_j)^2 exactly
    s = np.array([[0, 0, 0, 0.5], [1, 0, 0, 0.5], [0, 1, 0, 0.5], [0, 0, 1, 0.5],
                  [0.5, 0.5, 0.5, 0.5]], np.float32)   # 0.866 from each of 0-3
    pairs, found = oracle.broad_phase(s, 2.0, 4.0, 4)
    assert [tuple(p) for p in pairs] == [(0, 4), (1, 4), (2, 4), (3, 4)]
    s[1, 0] = np.nextafter(np.float32(1), np.float32(0))
    pairs, _ = oracle.broad_phase(s, 2.0, 4.0, 4)
    assert (0, 1) in set(map(tuple, pairs.tolist()))


def test_oracle_order_hand_case():
    # object 0 at the centre cell; neighbours placed in known cells
    cs = 1.0
    s = np.array([[0.5, 0.5, 0.5, 0.6],     # 0: cell (2,2,2) of a 5-grid from -2
                  [1.5, 0.5, 0.5, 0.6],     # 1: +x
                  [0.5, 0.5, -0.4, 0.6],    # 2: -z
                  [0.5, -0.4, 0.5, 0.6],    # 3: -y
                  [0.6, 0.6, 0.6, 0.6]],    # 4: same cell
                 np.float32)
    pairs, _ = oracle.broad_phase(s, cs, 2.0, 5)
    # object 0 visits -z (2), then -y (3), then its own cell (4), then +x (1)
    assert [tuple(p) for p in pairs[:4]] == [(0, 2), (0, 3), (0, 4), (0, 1)]


def test_oracle_cap_keeps_lowest_indices():
    s = np.zeros((6, 4), np.float32)
    s[:, 3] = 1.0
    s[:, 0] = np.linspace(0.0, 0.5, 6)        # all in one cell, all overlapping
    pairs, found = oracle.broad_phase(s, 4.0, 4.0, 2, max_per_cell=3)
    assert [tuple(p) for p in pairs] == [(0, 1), (0, 2), (1, 2)]
    pairs, found = oracle.broad_phase(s, 4.0, 4.0, 2, max_per_cell=0)
    assert found == 15


def test_oracle_max_pairs_keeps_prefix_and_nan_never_pairs():
    s = gen.spheres(1000, seed=9, box=4.0, rmin=0.2, rmax=0.6)
    s[7, 1] = np.nan
    grid = ogjk.Grid.for_spheres(0.6, 4.0)
    full, found = oracle.broad_phase(s, grid.cell_size, grid.boundary, grid.grid_size)
    assert not np.any(full == 7)
    part, found2 = oracle.broad_phase(s, grid.cell_size, grid.boundary, grid.grid_size,
                                      max_pairs=found // 3)
    assert found2 == found and np.array_equal(part, full[:found // 3])
    empty, f0 = oracle.broad_phase(s[:0], grid.cell_size, grid.boundary, grid.grid_size)
    assert f0 == 0 and len(empty) == 0


def test_oracle_expand_pairs():
    sub_off = np.array([0, 2, 3, 6], np.int32)
    sub_cnt = np.array([2, 1, 3, 0], np.int32)
    bp = np.array([[0, 1], [1, 2], [0, 3], [5, 1], [0, 2]], np.int32)
    out, found = oracle.expand_pairs(bp, sub_off, sub_cnt)
    assert [tuple(p) for p in out] == [(0, 2), (1, 2), (2, 3), (2, 4), (2, 5),
                                       (0, 3), (0, 4), (0, 5), (1, 3), (1, 4), (1, 5)]
    assert found == 11
    part, f2 = oracle.expand_pairs(bp, sub_off, sub_cnt, max_sub=4)
    assert f2 == 11 and np.array_equal(part, out[:4])
    # one sub-mesh per body at offset = body index: the identity
    rng = np.random.default_rng(3)
    bp = np.sort(rng.integers(0, 100, (500, 2)), 1).astype(np.int32)
    ident, f3 = oracle.expand_pairs(bp, np.arange(100, dtype=np.int32),
                                    np.ones(100, np.int32))
    assert f3 == 500 and np.array_equal(ident, bp)
