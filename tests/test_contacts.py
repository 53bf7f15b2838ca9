"""Contact records (ogjk_collide_scene_contacts, ogjk_collide_posed_bodies_contacts):
each 40-byte record must equal the FP64 oracle's distance, contact normal and
witnesses rounded once to float32 (the fields collision_response_kernel reads,
visualization/integrate_final_gjk.cu:572-689), and the gkSimplex witnesses,
distance and normal that the ABI outputs of the same launch hold, bit for bit."""
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
    return torch


def as_records(out):
    simp = ogjk.simplices_from_device(out["simplices"])
    d = out["distances"].cpu().numpy()
    nrm = out["normals"].cpu().numpy().reshape(-1, 3)
    w = simp["witnesses"].reshape(-1, 6)
    return np.concatenate([d[:, None], nrm, w], 1).astype(np.float32)


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                          np.ascontiguousarray(b).view(np.uint32))


@pytest.mark.parametrize("mode", [oracle.GJK, oracle.GJK_EPA])
def test_scene_contacts_equal_abi_outputs(torch_cuda, mode):
    torch = torch_cuda
    b = gen.make(gen.SPHERE, 8192, seed=61, nverts=64, box=1.3)
    sc = ogjk.DeviceScene(b)
    n = b.n_pairs
    out = sc.collide(mode, n, f64_record=True)
    con = sc.collide_contacts(mode, n, f64_record=True)
    torch.cuda.synchronize()
    ref = as_records(out)
    got = con["contacts"][:n].cpu().numpy()
    if mode == oracle.GJK:
        ref[:, 1:4] = 0.0  # no normal without EPA
    print(f"\n[contacts mode {mode}] {n} pairs, {(ref[:, 0] <= 0).mean():.2f} penetrating",
          flush=True)
    assert same_bits(got, ref)
    assert np.array_equal(con["f64"][:n].cpu().numpy().view(np.uint64),
                          out["f64"][:n].cpu().numpy().view(np.uint64))


def oracle_records(ref, mode):
    """The FP64 oracle's results as contact records, rounded once to float32:
    distance, normal (zero without EPA), witness on A, witness on B."""
    n = len(ref.distances)
    nrm = ref.normals if mode & oracle.EPA else np.zeros((n, 3))
    w = ref.simplices["witnesses"].reshape(n, 6)
    return np.concatenate([ref.distances[:, None], nrm, w], 1).astype(np.float32)


@pytest.mark.parametrize("mode", [oracle.GJK, oracle.GJK_EPA])
def test_scene_contacts_equal_oracle(torch_cuda, mode):
    """Uneven sizes (buckets and clusters) and tie-heavy bodies."""
    import cases
    torch = torch_cuda
    for name, b in (("uneven", gen.make(gen.ELLIPSOID, 3000, seed=67, box=1.5, nmin=4, nmax=4096)),
                    ("degenerate", cases.degenerate_batch(2000, seed=68))):
        ref = oracle.run(mode, b.coords, b.off, b.cnt, b.pa, b.pb, bits=64, nthreads=8)
        sc = ogjk.DeviceScene(b)
        sc.build_clusters(256)
        ogjk.set_buckets([128], [8, 64], [32, 64])
        try:
            con = sc.collide_contacts(mode, b.n_pairs)
            torch.cuda.synchronize()
        finally:
            ogjk.set_config(8, 32, False)
        got = con["contacts"][:b.n_pairs].cpu().numpy()
        want = oracle_records(ref, mode)
        same = (got == want) | (np.isnan(got) & np.isnan(want))
        print(f"\n[contacts vs oracle, {name}, mode {mode}] {b.n_pairs} pairs, "
              f"{(want[:, 0] <= 0).mean():.2f} penetrating", flush=True)
        assert same.all(), f"{name}: records differ at pairs {np.where(~same.all(1))[0][:10]}"


def test_scene_step_contacts_equal_oracle(torch_cuda):
    """Posed bodies from the broad phase: records against the oracle run on
    the FP64 world vertices of each pair."""
    torch = torch_cuda
    b = gen.make(gen.ELLIPSOID, 1000, seed=69, box=0.0, nmin=8, nmax=512)
    nb = len(b.cnt)
    radii = gen.body_radii(b.coords, b.off, b.cnt)
    poses_h = gen.body_poses(nb, seed=70, tbox=12.0)
    grid = ogjk.Grid.for_spheres(float(radii.max()), 12.0, 30)
    sc = ogjk.DeviceScene(b)
    step = ogjk.SceneStep(sc, radii, grid, max_pairs=20 * nb, contacts=True)
    pairs, found, con = step(torch.from_numpy(poses_h).cuda())
    torch.cuda.synchronize()
    pr = pairs.cpu().numpy()
    n = len(pr)
    side = np.concatenate([poses_h[pr[:, 0]], poses_h[pr[:, 1]]], 1)
    pool, off, cnt, pa, pb = oracle.posed_pool(b.coords, b.off, b.cnt, pr, side)
    ref = oracle.run(oracle.GJK_EPA, pool, off, cnt, pa, pb, bits=64, nthreads=8)
    got = con["contacts"][:n].cpu().numpy()
    want = oracle_records(ref, oracle.GJK_EPA)
    print(f"\n[scene step contacts vs oracle] {n} pairs, {(want[:, 0] <= 0).mean():.2f} "
          "penetrating", flush=True)
    assert n > nb // 2 and np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_scene_contacts_invalid_pairs_are_nan(torch_cuda):
    torch = torch_cuda
    b = gen.make(gen.SPHERE, 64, seed=62, nverts=32, box=1.3)
    sc = ogjk.DeviceScene(b)
    pairs = np.stack([np.arange(0, 128, 2), np.arange(1, 128, 2)], 1).astype(np.int32)
    pairs[5, 1] = 10_000
    pairs[9, 0] = -1
    con = sc.collide_contacts(oracle.GJK_EPA, len(pairs), pairs=torch.from_numpy(pairs).cuda())
    torch.cuda.synchronize()
    got = con["contacts"][:len(pairs)].cpu().numpy()
    assert np.isnan(got[[5, 9]]).all()
    ok = np.ones(len(pairs), bool)
    ok[[5, 9]] = False
    assert np.isfinite(got[ok]).all()


def test_scene_step_contacts_equal_abi_outputs(torch_cuda):
    torch = torch_cuda
    b = gen.make(gen.ELLIPSOID, 1000, seed=63, box=0.0, nmin=8, nmax=512)
    nb = len(b.cnt)
    radii = gen.body_radii(b.coords, b.off, b.cnt)
    poses = torch.from_numpy(gen.body_poses(nb, seed=64, tbox=12.0)).cuda()
    grid = ogjk.Grid.for_spheres(float(radii.max()), 12.0, 30)
    sc = ogjk.DeviceScene(b)
    full = ogjk.SceneStep(sc, radii, grid, max_pairs=20 * nb)
    comp = ogjk.SceneStep(sc, radii, grid, max_pairs=20 * nb, contacts=True)
    pairs, found, out = full(poses)
    pairs2, found2, con = comp(poses)
    torch.cuda.synchronize()
    n = len(pairs)
    print(f"\n[scene step contacts] {n} pairs", flush=True)
    assert found == found2 and torch.equal(pairs, pairs2) and n > nb // 2
    ref = as_records({k: v[:n] for k, v in out.items() if k != "_workspace" and v is not None})
    assert same_bits(con["contacts"][:n].cpu().numpy(), ref)


def test_posed_bodies_invalid_pairs(torch_cuda):
    """Out-of-range body indices in a per-body-pose pair list give NaN results
    (gkSimplex outputs and contact records alike), the other pairs are
    untouched by them."""
    torch = torch_cuda
    b = gen.make(gen.ELLIPSOID, 100, seed=65, box=0.0, nmin=8, nmax=64)
    nb = len(b.cnt)
    sc = ogjk.DeviceScene(b)
    poses = torch.from_numpy(gen.body_poses(nb, seed=66, tbox=1.0)).cuda()
    pairs = np.stack([np.arange(0, nb, 2), np.arange(1, nb, 2)], 1).astype(np.int32)
    pairs[3, 0] = nb
    pairs[7, 1] = -5
    dp = torch.from_numpy(pairs).cuda()
    out = sc.collide_posed_bodies(oracle.GJK_EPA, dp, poses)
    con = sc.collide_posed_bodies_contacts(oracle.GJK_EPA, dp, poses)
    ref = sc.collide_posed_bodies(oracle.GJK_EPA, dp[[0, 1, 2, 4]].contiguous(), poses)
    torch.cuda.synchronize()
    d = out["distances"].cpu().numpy()
    c = con["contacts"][:len(pairs)].cpu().numpy()
    simp = ogjk.simplices_from_device(out["simplices"])
    assert np.isnan(d[[3, 7]]).all() and (simp["nvrtx"][[3, 7]] == 0).all()
    assert np.isnan(c[[3, 7]]).all()
    ok = np.ones(len(pairs), bool)
    ok[[3, 7]] = False
    assert np.isfinite(d[ok]).all() and np.isfinite(c[ok]).all()
    assert np.array_equal(d[[0, 1, 2, 4]].view(np.uint32), ref["distances"].cpu().numpy().view(np.uint32))
