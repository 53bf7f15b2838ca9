"""Broad phase feeding the narrow phase: bodies of a resident scene placed by
one pose each, bounding spheres through the device grid, the pair list
through the posed GJK+EPA.  The pair list must equal the oracle's broad
phase, and every pair's results the FP64 oracle's on the world vertices."""
import numpy as np
import pytest

import oracle
import ogjk
from ogjk import gen
from parity_utils import compare_f64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("tbox,max_grid", [(12.0, 30), (20.0, 64)])
def test_scene_step_parity(torch_cuda, tbox, max_grid):
    torch = torch_cuda
    b = gen.make(gen.ELLIPSOID, 1000, seed=51, box=0.0, nmin=8, nmax=512)
    nb = len(b.cnt)
    radii = gen.body_radii(b.coords, b.off, b.cnt)
    poses = gen.body_poses(nb, seed=52, tbox=tbox)
    grid = ogjk.Grid.for_spheres(float(radii.max()), tbox, max_grid)
    sc = ogjk.DeviceScene(b)
    step = ogjk.SceneStep(sc, radii, grid, max_pairs=20 * nb, f64_record=True)
    pairs, found, out = step(torch.from_numpy(poses).cuda())
    torch.cuda.synchronize()
    got = pairs.cpu().numpy()
    spheres = np.concatenate([poses[:, 4:7], radii[:, None]], 1)
    ref_pairs, rfound = oracle.broad_phase(spheres, grid.cell_size, grid.boundary, grid.grid_size)
    print(f"\n[scene step tbox={tbox}] {nb} bodies, {found} pairs (oracle {rfound})", flush=True)
    assert found == rfound and np.array_equal(got, ref_pairs) and found > nb // 2
    n = len(got)
    side = np.concatenate([poses[got[:, 0]], poses[got[:, 1]]], 1)
    pool, off, cnt, pa, pb = oracle.posed_pool(b.coords, b.off, b.cnt, got, side)
    ref = oracle.run(oracle.GJK_EPA, pool, off, cnt, pa, pb, bits=64, nthreads=8)
    rec = out["f64"][:n].cpu().numpy()
    simp = ogjk.simplices_from_device(out["simplices"][:n])
    res = compare_f64(rec, simp, ref, oracle.GJK_EPA)
    print(f"[scene step tbox={tbox}] {res}", flush=True)
    assert (ref.distances < 0).mean() > 0.05
    assert res["exact"] == res["n"], res["inexact"]
