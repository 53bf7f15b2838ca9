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


@pytest.mark.parametrize("tbox,max_grid,lanes",
                         [(12.0, 30, (8, 32)), (20.0, 64, (8, 32)), (12.0, 30, (64, 64))])
def test_scene_step_parity(torch_cuda, tbox, max_grid, lanes):
    """lanes (64, 64): one pose per body in wave-wide groups, whose support
    queries run one after the other."""
    torch = torch_cuda
    ogjk.set_config(lanes[0], lanes[1], 0)
    b = gen.make(gen.ELLIPSOID, 1000, seed=51, box=0.0, nmin=8, nmax=512)
    nb = len(b.cnt)
    radii = gen.body_radii(b.coords, b.off, b.cnt)
    poses = gen.body_poses(nb, seed=52, tbox=tbox)
    grid = ogjk.Grid.for_spheres(float(radii.max()), tbox, max_grid)
    sc = ogjk.DeviceScene(b)
    step = ogjk.SceneStep(sc, radii, grid, max_pairs=20 * nb, f64_record=True)
    try:
        pairs, found, out = step(torch.from_numpy(poses).cuda())
        torch.cuda.synchronize()
    finally:
        ogjk.set_config(8, 32, 0)  # the library's default launch shape
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


def test_sub_mesh_pipeline_parity(torch_cuda):
    """The reference's whole step with bodies made of several convex sub-meshes
    (visualization/integrate_final_gjk.cu:916-1035): broad phase over body
    spheres, body pairs expanded to sub-mesh pairs, each sub-mesh placed by
    its body's pose, GJK+EPA on every sub-mesh pair.  Pair lists and results
    must equal the oracle chain's."""
    torch = torch_cuda
    rng = np.random.default_rng(55)
    b = gen.make(gen.ELLIPSOID, 400, seed=56, box=0.0, nmin=8, nmax=256)
    n_sub = len(b.cnt)                                   # 800 sub-meshes
    sub_cnt = []
    while sum(sub_cnt) < n_sub:
        sub_cnt.append(int(min(rng.integers(1, 5), n_sub - sum(sub_cnt))))
    sub_cnt = np.array(sub_cnt, np.int32)
    nb = len(sub_cnt)
    sub_off = np.concatenate([[0], np.cumsum(sub_cnt)[:-1]]).astype(np.int32)
    sub_body = np.repeat(np.arange(nb), sub_cnt)
    r_sub = gen.body_radii(b.coords, b.off, b.cnt)
    radii = np.maximum.reduceat(r_sub, sub_off).astype(np.float32)
    poses = gen.body_poses(nb, seed=57, tbox=10.0)
    grid = ogjk.Grid.for_spheres(float(radii.max()), 10.0, 30)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    spheres = np.concatenate([poses[:, 4:7], radii[:, None]], 1).astype(np.float32)
    body_pairs, found = ogjk.broad_phase(T(spheres), grid, max_pairs=50 * nb)
    sub_pairs, nsub = ogjk.expand_pairs(body_pairs, T(sub_off), T(sub_cnt))
    sub_poses = T(poses[sub_body])
    sc = ogjk.DeviceScene(b)
    out = sc.collide_posed_bodies(oracle.GJK_EPA, sub_pairs, sub_poses, f64_record=True)
    torch.cuda.synchronize()
    ref_bp, rfound = oracle.broad_phase(spheres, grid.cell_size, grid.boundary, grid.grid_size)
    ref_sp, rsub = oracle.expand_pairs(ref_bp, sub_off, sub_cnt)
    got_sp = sub_pairs.cpu().numpy()
    print(f"\n[sub-mesh step] {nb} bodies, {found} body pairs, {nsub} sub-mesh pairs", flush=True)
    assert found == rfound and np.array_equal(body_pairs.cpu().numpy(), ref_bp)
    assert nsub == rsub and np.array_equal(got_sp, ref_sp) and nsub > found
    sp = poses[sub_body]
    side = np.concatenate([sp[got_sp[:, 0]], sp[got_sp[:, 1]]], 1)
    pool, off, cnt, pa, pb = oracle.posed_pool(b.coords, b.off, b.cnt, got_sp, side)
    ref = oracle.run(oracle.GJK_EPA, pool, off, cnt, pa, pb, bits=64, nthreads=8)
    res = compare_f64(out["f64"].cpu().numpy(), ogjk.simplices_from_device(out["simplices"]),
                      ref, oracle.GJK_EPA)
    print(f"[sub-mesh step] {res}", flush=True)
    assert (ref.distances < 0).mean() > 0.05
    assert res["exact"] == res["n"], res["inexact"]
