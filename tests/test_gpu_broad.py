"""GPU broad phase against the oracle: the device grid (ogjk_broad_phase,
ogjk_expand_pairs through the C ABI) must produce the oracle's pair list
exactly, order included, on the same seeded spheres."""
import numpy as np
import pytest

import oracle
import ogjk
from ogjk import gen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert ogjk.device_count() >= 1
    return torch


def gpu_pairs(torch, s, grid, max_pairs=None):
    d = torch.from_numpy(np.ascontiguousarray(s)).cuda()
    pairs, found = ogjk.broad_phase(d, grid, max_pairs=max_pairs)
    torch.cuda.synchronize()
    return pairs.cpu().numpy(), found


def check_equal(torch, name, s, grid, max_pairs=None):
    print(f"\n[{name}] n={len(s)} {grid} max_pairs={max_pairs}", flush=True)
    ref, rfound = oracle.broad_phase(s, grid.cell_size, grid.boundary, grid.grid_size,
                                     grid.max_per_cell, max_pairs=max_pairs)
    got, found = gpu_pairs(torch, s, grid, max_pairs=max_pairs if max_pairs is not None
                           else max(rfound, 1))
    print(f"[{name}] found {found} (oracle {rfound}), written {len(got)}", flush=True)
    assert found == rfound
    assert got.shape == ref.shape and np.array_equal(got, ref)
    return found


SCENES = {
    "sparse_30": dict(n=20_000, box=20.0, rmax=0.8, outside=0.05, max_grid=30, cap=0),
    "dense_capped": dict(n=20_000, box=5.0, rmax=0.4, outside=0.0, max_grid=10, cap=8),
    "dense_uncapped": dict(n=20_000, box=5.0, rmax=0.4, outside=0.0, max_grid=10, cap=0),
    "fine_128": dict(n=100_000, box=40.0, rmax=0.3, outside=0.02, max_grid=128, cap=0),
    "one_cell": dict(n=3_000, box=3.0, rmax=0.2, outside=0.0, max_grid=1, cap=0),
    "cap_one": dict(n=5_000, box=2.0, rmax=0.3, outside=0.0, max_grid=8, cap=1),
    # crowded cells (more than 64 objects) are ordered by the radix-sort path
    "crowded_cell_20k": dict(n=20_000, box=3.0, rmax=0.1, outside=0.0, max_grid=1, cap=0),
    "crowded_capped": dict(n=20_000, box=3.0, rmax=0.1, outside=0.3, max_grid=2, cap=512),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_broad_phase_matches_oracle(torch_cuda, name):
    c = SCENES[name]
    s = gen.spheres(c["n"], seed=20 + list(SCENES).index(name), box=c["box"], rmin=0.05, rmax=c["rmax"],
                    outside=c["outside"])
    g = ogjk.Grid.for_spheres(float(s[:, 3].max()), c["box"], c["max_grid"], c["cap"])
    assert check_equal(torch_cuda, name, s, g) > 0


def test_broad_phase_special_values_and_clamp(torch_cuda):
    s = gen.spheres(10_000, seed=11, box=6.0, rmin=0.1, rmax=0.5, outside=0.1)
    s[5, 0] = np.nan
    s[17, 2] = np.inf
    s[18, 1] = -np.inf
    s[40, :3] = 1e30
    g = ogjk.Grid.for_spheres(0.5, 6.0)
    full = check_equal(torch_cuda, "special", s, g)
    check_equal(torch_cuda, "clamped", s, g, max_pairs=full // 3)
    check_equal(torch_cuda, "zero_max", s, g, max_pairs=0)


def test_broad_phase_tiny_inputs(torch_cuda):
    g = ogjk.Grid(1.0, 2.0, 4)
    s = np.array([[0, 0, 0, 0.5], [0.9, 0, 0, 0.5]], np.float32)
    got, found = gpu_pairs(torch_cuda, s, g)
    assert found == 1 and got.tolist() == [[0, 1]]
    got, found = gpu_pairs(torch_cuda, s[:1], g)
    assert found == 0 and len(got) == 0
    got, found = gpu_pairs(torch_cuda, s[:0], g)
    assert found == 0 and len(got) == 0


def test_broad_phase_one_million(torch_cuda):
    s = gen.spheres(1_000_000, seed=12, box=100.0, rmin=0.2, rmax=1.0, outside=0.01)
    g = ogjk.Grid(2.0, 100.0, 100)
    assert check_equal(torch_cuda, "1m", s, g) > 100_000


def test_broad_phase_object_reuse(torch_cuda):
    """A BroadPhase object reused over moving spheres gives each frame's list."""
    torch = torch_cuda
    s = gen.spheres(30_000, seed=13, box=15.0, rmin=0.1, rmax=0.6)
    g = ogjk.Grid.for_spheres(0.6, 15.0)
    bp = ogjk.BroadPhase(len(s), g, max_pairs=2_000_000)
    rng = np.random.default_rng(0)
    for frame in range(3):
        s[:, :3] += rng.normal(scale=0.2, size=(len(s), 3)).astype(np.float32)
        d = torch.from_numpy(s).cuda()
        pairs, found = bp(d)
        ref, rfound = oracle.broad_phase(s, g.cell_size, g.boundary, g.grid_size)
        print(f"\n[frame {frame}] found {found}", flush=True)
        assert found == rfound and np.array_equal(pairs.cpu().numpy(), ref)


def test_expand_pairs_matches_oracle(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(14)
    nb = 5000
    sub_cnt = rng.integers(0, 4, nb).astype(np.int32)
    sub_cnt[rng.random(nb) < 0.05] = -1
    sub_off = np.concatenate([[0], np.cumsum(np.maximum(sub_cnt, 0))[:-1]]).astype(np.int32)
    bp = rng.integers(0, nb, (100_000, 2)).astype(np.int32)
    bp[rng.random(len(bp)) < 0.01, 1] = nb + 3
    bp[rng.random(len(bp)) < 0.01, 0] = -2
    ref, rfound = oracle.expand_pairs(bp, sub_off, sub_cnt)
    T = lambda a: torch.from_numpy(a).cuda()
    got, found = ogjk.expand_pairs(T(bp), T(sub_off), T(sub_cnt))
    print(f"\n[expand] found {found} (oracle {rfound})", flush=True)
    assert found == rfound and np.array_equal(got.cpu().numpy(), ref)
    got, found = ogjk.expand_pairs(T(bp), T(sub_off), T(sub_cnt), max_sub_pairs=rfound // 2)
    assert found == rfound and np.array_equal(got.cpu().numpy(), ref[:rfound // 2])


def test_broad_phase_large_sparse_grid(torch_cuda):
    """64M cells, 200K objects: most cells empty, start table and histogram far
    larger than the object count."""
    s = gen.spheres(200_000, seed=15, box=200.0, rmin=0.1, rmax=0.5, outside=0.01)
    g = ogjk.Grid(1.0, 200.0, 400)
    assert check_equal(torch_cuda, "sparse_400", s, g) > 200
