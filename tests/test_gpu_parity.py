"""GPU parity: the HIP path, called through the C ABI, against the FP64 oracle
on the same seeded inputs.  Tolerances are defined in parity_utils."""
import numpy as np
import pytest

import oracle
import ogjk
from ogjk import gen
import cases
from parity_utils import compare_f32, compare_f64, oracle_run

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert ogjk.device_count() >= 1
    return torch


def run_scene(torch, batch, mode, plain=False, cluster_min=None):
    sc = ogjk.DeviceScene(batch, f64=plain)
    if cluster_min is not None:
        sc.build_clusters(cluster_min)
    out = sc.collide(mode, batch.n_pairs, f64_record=True, plain_f64=plain)
    torch.cuda.synchronize()
    rec = out["f64"].cpu().numpy()
    simp = ogjk.simplices_from_device(out["simplices"])
    return rec, simp, out["distances"].cpu().numpy(), out["normals"].cpu().numpy()


def check(res, name, min_exact=1.0):
    print(f"\n[{name}] {res}", flush=True)
    assert res["within_tol"] == res["n"], f"{name}: out of tolerance at pairs {res['bad']}"
    assert res["exact"] >= min_exact * res["n"], f"{name}: inexact pairs {res['inexact']}"


WORKLOADS = {
    "c0_sphere32_box2": dict(family=gen.SPHERE, n_pairs=1000, seed=0, nverts=32, box=2.0),
    "c1_sphere64_box1.3": dict(family=gen.SPHERE, n_pairs=4096, seed=1, nverts=64, box=1.3),
    "c2_ellipsoid_4_4096": dict(family=gen.ELLIPSOID, n_pairs=384, seed=2, box=1.5, nmin=4, nmax=4096),
    "c3_noisy_sphere_2000": dict(family=gen.NOISY_SPHERE, n_pairs=64, seed=3, nverts=2000, box=1.4, sigma=0.01),
}
LARGE = {"c3_noisy_sphere_100k": dict(family=gen.NOISY_SPHERE, n_pairs=48, seed=3, nverts=100_000,
                                      box=1.5, sigma=0.01)}


@pytest.mark.parametrize("name", list(WORKLOADS))
@pytest.mark.parametrize("mode", [oracle.GJK, oracle.GJK_EPA])
def test_workload_parity(torch_cuda, name, mode):
    b = gen.make(**WORKLOADS[name])
    ref = oracle_run(mode, b)
    rec, simp, d32, n32 = run_scene(torch_cuda, b, mode)
    check(compare_f64(rec, simp, ref, mode), f"{name} mode{mode}")
    r32 = compare_f32(simp, d32, n32, ref, mode)
    assert r32["exact"] == r32["n"], r32


@pytest.mark.parametrize("mode", [oracle.GJK, oracle.GJK_EPA])
@pytest.mark.parametrize("batch", ["degenerate", "lattice"])
def test_degenerate_parity(torch_cuda, mode, batch):
    b = cases.degenerate_batch(4000) if batch == "degenerate" else cases.lattice_batch(200)
    ref = oracle_run(mode, b)
    rec, simp, d32, n32 = run_scene(torch_cuda, b, mode)
    check(compare_f64(rec, simp, ref, mode), f"{batch} mode{mode}")


# Cluster-culled support sweeps select exactly what the full sweep selects:
# min_verts 1 culls every body, including the tie-heavy degenerate cases.
@pytest.mark.parametrize("name,cmin", [("c1_sphere64_box1.3", 1), ("c2_ellipsoid_4_4096", 1),
                                       ("c2_ellipsoid_4_4096", 512), ("c3_noisy_sphere_2000", 1),
                                       ("c3_noisy_sphere_100k", 1024)])
def test_cluster_parity(torch_cuda, name, cmin):
    b = gen.make(**(WORKLOADS.get(name) or LARGE[name]))
    ref = oracle_run(oracle.GJK_EPA, b)
    rec, simp, d32, n32 = run_scene(torch_cuda, b, oracle.GJK_EPA, cluster_min=cmin)
    check(compare_f64(rec, simp, ref, oracle.GJK_EPA), f"{name} clusters>={cmin}")


@pytest.mark.parametrize("mode", [oracle.GJK, oracle.GJK_EPA])
@pytest.mark.parametrize("batch", ["degenerate", "lattice"])
def test_cluster_degenerate_parity(torch_cuda, mode, batch):
    b = cases.degenerate_batch(4000) if batch == "degenerate" else cases.lattice_batch(200)
    ref = oracle_run(mode, b)
    rec, simp, _, _ = run_scene(torch_cuda, b, mode, cluster_min=1)
    check(compare_f64(rec, simp, ref, mode), f"{batch} clusters mode{mode}")


def test_cluster_layout(torch_cuda):
    """Every body's slots hold a permutation of its vertices, padded with -1,
    and every cluster's sphere contains its vertices."""
    b = gen.make(gen.ELLIPSOID, 200, seed=5, box=1.5, nmin=4, nmax=9000)
    sc = ogjk.DeviceScene(b)
    sc.build_clusters(1)
    assert sc.clusters.n_supers_used > 0
    torch_cuda.cuda.synchronize()
    sup = sc.csuper.cpu().numpy()
    sbounds = sc.csbounds.cpu().numpy().reshape(-1, 8).astype(np.float64)
    slot = sc.cslot.cpu().numpy()
    orig = sc.corig.cpu().numpy()
    cx, cy, cz = (t.cpu().numpy().astype(np.float64) for t in (sc.cx, sc.cy, sc.cz))
    bounds = sc.cbounds.cpu().numpy().reshape(-1, 8).astype(np.float64)
    for i, (o, n) in enumerate(zip(b.off, b.cnt)):
        s0 = slot[i]
        assert s0 % 32 == 0
        pad = -(-n // 32) * 32
        idx = orig[s0:s0 + pad]
        assert sorted(idx[:n].tolist()) == list(range(n)) and (idx[n:] == -1).all()
        pts = b.coords[o + idx[:n]].astype(np.float64)
        assert np.array_equal(np.stack([cx[s0:s0 + n], cy[s0:s0 + n], cz[s0:s0 + n]], 1), pts)
        for c in range(pad // 32):
            q = pts[c * 32:(c + 1) * 32]
            ctr, r, h = bounds[s0 // 32 + c, :3], bounds[s0 // 32 + c, 3], bounds[s0 // 32 + c, 4:7]
            assert (np.linalg.norm(q - ctr, axis=1) <= r).all()
            assert (np.abs(q - ctr) <= h).all()
        nc = pad // 32
        if nc <= 128:
            assert sup[i] == -1
        else:  # superclusters: runs of 32 clusters, bounds containing their vertices
            for k in range(-(-nc // 32)):
                q = pts[k * 1024:(k + 1) * 1024]
                ctr, r, h = sbounds[sup[i] + k, :3], sbounds[sup[i] + k, 3], sbounds[sup[i] + k, 4:7]
                assert (np.linalg.norm(q - ctr, axis=1) <= r).all()
                assert (np.abs(q - ctr) <= h).all()


def _cluster_arrays(torch, batch, env, cmin=1):
    """Builds the scene's clusters of the bodies of at least cmin vertices
    with OGJK_CLUSTER_LDS_MAX=env (None: unset) and returns every array the
    build writes, cut to the filled part."""
    import os
    old = os.environ.pop("OGJK_CLUSTER_LDS_MAX", None)
    try:
        if env is not None:
            os.environ["OGJK_CLUSTER_LDS_MAX"] = env
        sc = ogjk.DeviceScene(batch)
        sc.build_clusters(cmin)
        torch.cuda.synchronize()
    finally:
        os.environ.pop("OGJK_CLUSTER_LDS_MAX", None)
        if old is not None:
            os.environ["OGJK_CLUSTER_LDS_MAX"] = old
    cnt = batch.cnt.astype(np.int64)
    built = cnt >= max(cmin, 1)
    total = int((-(-cnt[built] // 32) * 32).sum())
    slot = sc.cslot.cpu().numpy()
    assert (slot[~built] == -1).all() and (slot[built] >= 0).all()
    ns = int(sc.clusters.n_supers_used)
    bits = lambda t, n: t.cpu().numpy()[:n].view(np.uint32)  # noqa: E731
    return {"x": bits(sc.cx, total), "y": bits(sc.cy, total), "z": bits(sc.cz, total),
            "orig": bits(sc.corig, total), "bounds": bits(sc.cbounds, total // 32 * 8),
            "slot": slot, "super": sc.csuper.cpu().numpy(),
            "sbounds": bits(sc.csbounds, ns * 8), "n_supers": ns}


def test_cluster_build_paths_identical(torch_cuda):
    """Bodies of at most 4,096 vertices are built by one block each (keys, a
    stable sort in LDS, copy and bounds in one kernel); the batched path (keys,
    device segmented sort, scatter, bounds kernel) must give the same arrays
    bit for bit.  Sizes straddle the block classes (256 ... 4096) and the
    batched bodies; repeated points give equal keys, so both sorts must keep
    index order among them."""
    rng = np.random.default_rng(12)
    sizes = [1, 2, 3, 31, 32, 33, 255, 256, 257, 767, 768, 769, 1023, 1024, 1025, 1535, 1536, 1537,
             2047, 2048, 2049, 3071, 3072, 3073, 4095, 4096, 4097, 5000, 9000]
    A, B = [], []
    for n in sizes:
        for side in (A, B):
            p = rng.normal(size=(n, 3)) * rng.uniform(0.2, 2.0, size=3)
            if n > 8:  # repeated points: equal sort keys
                k = max(1, n // 10)
                p[rng.integers(0, n, size=k)] = p[rng.integers(0, n, size=k)]
            side.append(p.astype(np.float32))
    mixed = gen.from_bodies(A, B)
    small = gen.make(gen.ELLIPSOID, 300, seed=7, box=1.5, nmin=4, nmax=4096)
    for name, batch in (("mixed sizes", mixed), ("ellipsoids <= 4096", small)):
        for cmin in (1, 300):  # 300: the smaller bodies are not built
            ref = _cluster_arrays(torch_cuda, batch, "0", cmin)
            new = _cluster_arrays(torch_cuda, batch, None, cmin)
            assert ref["n_supers"] == new["n_supers"], name
            for k in ref:
                assert np.array_equal(ref[k], new[k]), f"{name}, from {cmin}: {k} differs"


def test_cluster_threshold_lowered_after_build(torch_cuda):
    """Clusters built from 512 vertices, then the culling threshold lowered
    to 1: the smaller bodies have no clusters and are swept whole; results
    stay exact."""
    torch = torch_cuda
    b = gen.make(gen.ELLIPSOID, 1500, seed=13, box=1.5, nmin=4, nmax=3000)
    ref = oracle_run(oracle.GJK_EPA, b)
    sc = ogjk.DeviceScene(b)
    sc.build_clusters(512)
    slot = sc.cslot.cpu().numpy()
    assert ((slot == -1) == (b.cnt < 512)).all()
    for cmin in (512, 1):
        sc.set_cluster_min(cmin)
        out = sc.collide(oracle.GJK_EPA, b.n_pairs, f64_record=True)
        torch.cuda.synchronize()
        res = compare_f64(out["f64"].cpu().numpy(), ogjk.simplices_from_device(out["simplices"]),
                          ref, oracle.GJK_EPA)
        check(res, f"clusters from 512, threshold {cmin}")


# Size-bucketed launch tables (device partition, per-bucket lanes, EPA lane
# classes), with and without clusters, against the oracle.
@pytest.mark.parametrize("table", [([64], [8, 64], [32, 64]), ([1], [32, 64], [32, 64]),
                                   ([32, 256, 1024], [8, 16, 32, 64], [16, 32, 64, 64])])
@pytest.mark.parametrize("cmin", [None, 256])
def test_bucketed_parity(torch_cuda, table, cmin):
    b = gen.make(gen.ELLIPSOID, 3000, seed=31, box=1.5, nmin=4, nmax=4096)
    ref = oracle_run(oracle.GJK_EPA, b)
    ogjk.set_buckets(*table)
    try:
        rec, simp, _, _ = run_scene(torch_cuda, b, oracle.GJK_EPA, cluster_min=cmin)
    finally:
        ogjk.set_config(8, 32, False)
    check(compare_f64(rec, simp, ref, oracle.GJK_EPA), f"buckets {table} clusters {cmin}")


# Pairs per workgroup: GJK and EPA blocks of 1, 2 or 4 waves per bucket.
@pytest.mark.parametrize("gw,ew", [([2, 4], [1, 2]), ([4, 1], [2, 1]), ([1, 2], [4, 4])])
def test_block_waves_parity(torch_cuda, gw, ew):
    b = gen.make(gen.ELLIPSOID, 3000, seed=32, box=1.5, nmin=4, nmax=2048)
    ref = oracle_run(oracle.GJK_EPA, b)
    ogjk.set_buckets([96], [8, 32], [32, 64])
    ogjk.set_block_waves(gw, ew)
    try:
        rec, simp, _, _ = run_scene(torch_cuda, b, oracle.GJK_EPA, cluster_min=512)
    finally:
        ogjk.set_config(8, 32, False)
    check(compare_f64(rec, simp, ref, oracle.GJK_EPA), f"block waves gjk {gw} epa {ew}")
    with pytest.raises(ogjk.OgjkError):
        ogjk.set_block_waves([3], [1])


def test_plain_f64_equals_mixed(torch_cuda):
    b = gen.make(gen.SPHERE, 2048, seed=11, nverts=64, box=1.3)
    rec_m, simp_m, _, _ = run_scene(torch_cuda, b, oracle.GJK_EPA)
    rec_p, simp_p, _, _ = run_scene(torch_cuda, b, oracle.GJK_EPA, plain=True)
    assert np.array_equal(rec_m.view(np.uint64), rec_p.view(np.uint64))
    assert np.array_equal(simp_m.view(np.uint8), simp_p.view(np.uint8))


# The plain-FP64 comparison kernels in every launch shape the benchmark
# times: fused (one kernel), split (the reference's GJK kernel then EPA
# kernel, GJK/gpu/openGJK.cu:2968-3004) over SoA and AoS FP64 pools, and one
# thread per pair for GJK.  Each must equal the FP64 oracle bit for bit.
PLAIN_SHAPES = [(1, 16, 16, True), (1, 64, 64, True), (2, 8, 32, "aos"), (2, 8, 32, True),
                (2, 1, 32, True), (2, 16, 64, True), (2, 32, 16, "aos")]


@pytest.mark.parametrize("shape,gl,el,layout", PLAIN_SHAPES)
@pytest.mark.parametrize("mode", [oracle.GJK, oracle.GJK_EPA])
def test_plain_f64_shapes(torch_cuda, shape, gl, el, layout, mode):
    if mode == oracle.GJK and shape == 1:
        pytest.skip("fused applies to GJK+EPA")
    b = gen.make(gen.ELLIPSOID, 1500, seed=41, box=1.5, nmin=4, nmax=700)
    ref = oracle_run(mode, b)
    sc = ogjk.DeviceScene(b, f64="both")
    ogjk.set_config(gl, el, shape)
    try:
        out = sc.collide(mode, b.n_pairs, f64_record=True, plain_f64=layout)
        torch_cuda.cuda.synchronize()
    finally:
        ogjk.set_config(8, 32, 0)
    rec = out["f64"].cpu().numpy()
    simp = ogjk.simplices_from_device(out["simplices"])
    check(compare_f64(rec, simp, ref, mode), f"plain shape {shape} {gl}/{el} {layout} mode{mode}")
    r32 = compare_f32(simp, out["distances"].cpu().numpy(), out["normals"].cpu().numpy(), ref, mode)
    assert r32["exact"] == r32["n"], r32


def test_split_shape_rejects_one_lane_elsewhere(torch_cuda):
    """One GJK lane per pair exists only in the split shape."""
    with pytest.raises(ogjk.OgjkError):
        ogjk.set_config(1, 32, 0)
    ogjk.set_config(8, 32, 0)


def test_known_answers_dropin(torch_cuda):
    b = cases.known_answer_batch()
    simp, dist, nrm = ogjk.collide_host(oracle.GJK_EPA, b)
    assert abs(dist[0] - 3.653650) < 1e-5                      # README.md:111-115
    assert np.allclose(simp["witnesses"][0, 0], [1.025173, 1.490318, 0.255463], atol=1e-6)
    assert np.allclose(simp["witnesses"][0, 1], [-1.025173, -1.490318, -0.255463], atol=1e-6)
    assert abs(-dist[1] - 1.5) < 1e-6                           # README.md:131-137
    assert np.allclose(simp["witnesses"][1, 0], [1.0, 0.5, 0.707107], atol=1e-6)
    assert np.allclose(simp["witnesses"][1, 1], [-0.5, 0.5, 0.707107], atol=1e-6)
    assert np.allclose(nrm[1], [1.0, 0.0, 0.0], atol=1e-6)
    assert -1.2 < dist[2] < -0.8                                # example.cu:450
    assert 0.0 <= dist[3] < 0.01 or dist[3] == 0.0             # example.cu:515
    assert 2.9 < dist[4] < 3.1 and simp["nvrtx"][4] < 4         # example.cu:580
    ref = oracle_run(oracle.GJK_EPA, b)
    r = compare_f32(simp, dist, nrm, ref, oracle.GJK_EPA)
    assert r["exact"] == r["n"], r


def test_host_paths_agree(torch_cuda):
    """API consistency in the manner of examples/main.cpp:378-644."""
    b = gen.make(gen.SPHERE, 1500, seed=4, nverts=64, box=1.3)
    ref = oracle_run(oracle.GJK_EPA, b)
    s1, d1, n1 = ogjk.collide_host(oracle.GJK_EPA, b)
    assert compare_f32(s1, d1, n1, ref, oracle.GJK_EPA)["exact"] == b.n_pairs
    # GJK then EPA as two calls (computeDistances + computeEPA)
    sg, dg, _ = ogjk.collide_host(oracle.GJK, b)
    se, de, ne = ogjk.collide_host(oracle.EPA, b, simplices=sg, distances=dg)
    r = compare_f32(se, de, ne, ref, oracle.GJK_EPA)
    assert r["max_dist_rel"] <= 1e-5, r
    wdiff = np.abs(se["witnesses"] - s1["witnesses"]).max()
    assert wdiff <= 1e-4, wdiff
    # indexed form over a shuffled pool
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(b.cnt))
    inv = np.argsort(perm)
    coords = np.concatenate([b.body(int(j)) for j in perm]).astype(np.float32)
    cnt = b.cnt[perm]
    off = np.zeros(len(cnt), np.int64)
    off[1:] = np.cumsum(cnt[:-1])
    pairs = np.stack([inv[b.pa], inv[b.pb]], 1)
    si, di, ni = ogjk.collide_host_indexed(oracle.GJK_EPA, np.ascontiguousarray(coords), off, cnt, pairs)
    assert np.array_equal(di.view(np.uint32), d1.view(np.uint32))
    assert np.array_equal(ni.view(np.uint32), n1.view(np.uint32))


def test_cluster_threshold_above_every_body(torch_cuda):
    """A build threshold no body reaches builds nothing (no slots, no
    superclusters, every body_slot -1); queries sweep every body whole and
    stay exact."""
    torch = torch_cuda
    b = gen.make(gen.ELLIPSOID, 300, seed=14, box=1.5, nmin=4, nmax=6000)
    ref = oracle_run(oracle.GJK_EPA, b)
    sc = ogjk.DeviceScene(b)
    sc.build_clusters(10**6)
    torch.cuda.synchronize()
    assert (sc.cslot.cpu().numpy() == -1).all() and (sc.csuper.cpu().numpy() == -1).all()
    assert sc.clusters.n_supers_used == 0
    out = sc.collide(oracle.GJK_EPA, b.n_pairs, f64_record=True)
    torch.cuda.synchronize()
    res = compare_f64(out["f64"].cpu().numpy(), ogjk.simplices_from_device(out["simplices"]), ref,
                      oracle.GJK_EPA)
    check(res, "clusters from 10^6 vertices (none built)")


def test_host_staging_reuse(torch_cuda):
    """The host entry points keep their staging buffers between calls: calls
    that grow, shrink and grow again, in every mode, each against the oracle,
    and a call too large for the kept buffers (its own allocations) against
    the device-resident path on the same pairs."""
    torch = torch_cuda
    small = gen.make(gen.SPHERE, 300, seed=61, nverts=32, box=1.2)
    mid = gen.make(gen.ELLIPSOID, 6000, seed=62, nmin=8, nmax=256, box=1.5)
    refs = {}
    for b in (small, mid, small, mid):
        if id(b) not in refs:
            refs[id(b)] = oracle_run(oracle.GJK_EPA, b)
        ref = refs[id(b)]
        s, d, n = ogjk.collide_host(oracle.GJK_EPA, b)
        r = compare_f32(s, d, n, ref, oracle.GJK_EPA)
        assert r["exact"] == r["n"], r
        sg, dg, _ = ogjk.collide_host(oracle.GJK, b)
        se, de, ne = ogjk.collide_host(oracle.EPA, b, simplices=sg, distances=dg)
        sep = ref.distances > 0  # GJK alone gives these pairs' final distances
        assert np.array_equal(dg[sep].view(np.uint32), d[sep].view(np.uint32))
        r = compare_f32(se, de, ne, ref, oracle.GJK_EPA)
        assert r["max_dist_rel"] <= 1e-5, r
    # 100,000 pairs of 128-vertex bodies: 307 MB of coordinates, above the
    # 256 MiB kept per device
    big = gen.make(gen.SPHERE, 100_000, seed=63, nverts=128, box=3.0)
    s, d, n = ogjk.collide_host(oracle.GJK_EPA, big)
    sc = ogjk.DeviceScene(big)
    pairs = torch.from_numpy(np.stack([big.pa, big.pb], 1).astype(np.int32)).cuda()
    dev = sc.collide(oracle.GJK_EPA, big.n_pairs, pairs=pairs)
    torch.cuda.synchronize()
    assert np.array_equal(d.view(np.uint32), dev["distances"].cpu().numpy().view(np.uint32))
    assert np.array_equal(n.view(np.uint32), dev["normals"].cpu().numpy().view(np.uint32))
    assert np.array_equal(s.view(np.uint8), ogjk.simplices_from_device(dev["simplices"]).view(np.uint8))
    assert (d < 0).any() and (d > 0).any()


def test_host_staging_concurrent_callers(torch_cuda):
    """Threads calling the host entry point at once share one device's kept
    staging (the call holds it until its results are copied out): every
    thread's results equal the same call made alone."""
    import threading
    batches = [gen.make(gen.SPHERE, 200 + 700 * k, seed=70 + k, nverts=24 + 16 * k, box=1.4)
               for k in range(4)]
    alone = [ogjk.collide_host(oracle.GJK_EPA, b) for b in batches]
    got = [[None] * 3 for _ in batches]

    def run(k):
        for r in range(3):
            got[k][r] = ogjk.collide_host(oracle.GJK_EPA, batches[k])

    th = [threading.Thread(target=run, args=(k,)) for k in range(len(batches))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for k, ref in enumerate(alone):
        for out in got[k]:
            for a, b in zip(out, ref):
                assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8)), k


def test_invalid_pairs_flagged(torch_cuda):
    b = gen.make(gen.SPHERE, 8, seed=9, nverts=16, box=1.0)
    pairs = np.array([[0, 1], [2, 99], [-1, 3], [4, 5]], np.int32)
    s, d, n = ogjk.collide_host_indexed(oracle.GJK_EPA, b.coords, b.off, b.cnt, pairs)
    assert np.isnan(d[1]) and np.isnan(d[2]) and s["nvrtx"][1] == 0
    assert np.isfinite(d[0]) and np.isfinite(d[3])
