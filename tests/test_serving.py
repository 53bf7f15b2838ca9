"""Serving loop and posed bodies: posed results against the FP64 oracle on
the world vertices, the device entry point against the server, and a
request's results alone and inside a mix of concurrent requests."""
import threading

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


@pytest.fixture(scope="module")
def scene(torch_cuda):
    b = gen.make(gen.ELLIPSOID, 1000, seed=41, box=0.0, nmin=8, nmax=512)
    return b, ogjk.DeviceScene(b)


def oracle_posed(b, pairs, poses):
    pool, off, cnt, pa, pb = oracle.posed_pool(b.coords, b.off, b.cnt, pairs, poses)
    return oracle.run(oracle.GJK_EPA, pool, off, cnt, pa, pb, bits=64, nthreads=8)


def test_posed_server_parity(torch_cuda, scene):
    b, sc = scene
    pairs, poses = gen.posed_requests(len(b.cnt), 1, seed=42, kmin=3000, kmax=3000, tbox=1.5)[0]
    srv = ogjk.Server(sc, max_batch_pairs=8192, slots=2)
    try:
        out, lat, _ = srv.wait(srv.submit(pairs, poses, f64=True))
    finally:
        srv.close()
    ref = oracle_posed(b, pairs, poses)
    res = compare_f64(out["f64"], out["simplices"], ref, oracle.GJK_EPA)
    print(f"\n[posed] {res} latency {lat * 1e3:.3f} ms", flush=True)
    assert (ref.distances < 0).mean() > 0.1  # EPA exercised
    assert res["exact"] == res["n"], res["inexact"]


@pytest.mark.parametrize("lanes", [(16, 32), (32, 64), (64, 64), (64, 32)])
def test_posed_parity_lanes(torch_cuda, scene, lanes):
    """Posed pairs at each launch shape against the oracle: 64-lane groups
    run the two support queries one after the other (wave-uniform poses),
    narrower ones split the group between them."""
    torch = torch_cuda
    b, sc = scene
    pairs, poses = gen.posed_requests(len(b.cnt), 1, seed=46, kmin=2000, kmax=2000, tbox=1.5)[0]
    ogjk.set_config(lanes[0], lanes[1], 0)
    try:
        dev = sc.collide_posed(oracle.GJK_EPA, torch.from_numpy(pairs).cuda(),
                               torch.from_numpy(poses).cuda(), f64_record=True)
        torch.cuda.synchronize()
    finally:
        ogjk.set_config(8, 32, 0)  # the library's default launch shape
    ref = oracle_posed(b, pairs, poses)
    res = compare_f64(dev["f64"].cpu().numpy(), ogjk.simplices_from_device(dev["simplices"]), ref,
                      oracle.GJK_EPA)
    print(f"\n[posed lanes {lanes}] {res}", flush=True)
    assert (ref.distances < 0).mean() > 0.1  # EPA exercised
    assert res["exact"] == res["n"], res["inexact"]


def test_posed_device_equals_server(torch_cuda, scene):
    torch = torch_cuda
    b, sc = scene
    pairs, poses = gen.posed_requests(len(b.cnt), 1, seed=43, kmin=2000, kmax=2000, tbox=1.5)[0]
    dev = sc.collide_posed(oracle.GJK_EPA, torch.from_numpy(pairs).cuda(),
                           torch.from_numpy(poses).cuda(), f64_record=True)
    torch.cuda.synchronize()
    srv = ogjk.Server(sc, max_batch_pairs=4096, slots=2)
    try:
        out, _, _ = srv.wait(srv.submit(pairs, poses, f64=True))
    finally:
        srv.close()
    assert np.array_equal(dev["f64"].cpu().numpy().view(np.uint64), out["f64"].view(np.uint64))
    assert np.array_equal(ogjk.simplices_from_device(dev["simplices"]).view(np.uint8),
                          out["simplices"].view(np.uint8))


def test_alone_vs_mix(torch_cuda, scene):
    b, sc = scene
    reqs = gen.posed_requests(len(b.cnt), 41, seed=44, kmin=1, kmax=3000, tbox=1.5)
    probe = 20
    srv = ogjk.Server(sc, max_batch_pairs=16384, slots=3)
    try:
        alone, _, _ = srv.wait(srv.submit(*reqs[probe], f64=True))
        # the same request among 40 others submitted back to back from two
        # threads, so that batches are packed from whatever has arrived
        tickets = [None] * len(reqs)

        def submit(idx):
            for k in idx:
                tickets[k] = srv.submit(*reqs[k], f64=True)

        th = [threading.Thread(target=submit, args=(list(range(s, len(reqs), 2)),)) for s in (0, 1)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        res = [srv.wait(t) for t in tickets]
        info = srv.info()
    finally:
        srv.close()
    mixed, _, batch = res[probe]
    shared = sum(1 for k, r in enumerate(res) if r[2] == batch and k != probe)
    # at least one request of the mix sat behind others in its launch, so the
    # per-request offsets into the batch buffers are exercised
    firsts = {}
    for k, r in enumerate(res):
        firsts.setdefault(r[2], k)
    assert any(firsts[r[2]] != k for k, r in enumerate(res))
    print(f"\n[mix] launches {info['launches']} for {len(reqs) + 1} requests; probe shared its "
          f"batch with {shared} others", flush=True)
    assert info["launches"] < len(reqs) + 1
    for key in ("f64",):
        assert np.array_equal(alone[key].view(np.uint64), mixed[key].view(np.uint64))
    assert np.array_equal(alone["simplices"].view(np.uint8), mixed["simplices"].view(np.uint8))
    assert np.array_equal(alone["distances"], mixed["distances"])
    assert np.array_equal(alone["normals"], mixed["normals"])
    # and every request of the mix against the oracle, on a sample of pairs,
    # including the float outputs each request receives
    for k in (0, 7, probe, 33):
        pairs, poses = reqs[k]
        m = min(len(pairs), 400)
        ref = oracle_posed(b, pairs[:m], poses[:m])
        out = res[k][0]
        r = compare_f64(out["f64"][:m], out["simplices"][:m], ref, oracle.GJK_EPA)
        assert r["exact"] == r["n"], (k, r["inexact"])
        assert np.array_equal(out["distances"][:m], ref.distances.astype(np.float32))
        assert np.array_equal(out["normals"][:m], ref.normals.astype(np.float32))


def test_destroy_with_outstanding_requests(torch_cuda, scene):
    """ogjk_server_destroy while requests are pending and threads are blocked
    in ogjk_server_wait: every waiter returns (answered, or cancelled when
    its request was never launched), destroy returns, answered requests hold
    the same results as the device entry point, and later submissions to a
    new server work."""
    import time
    torch = torch_cuda
    b, sc = scene
    reqs = gen.posed_requests(len(b.cnt), 60, seed=45, kmin=1500, kmax=1500, tbox=1.5)
    srv = ogjk.Server(sc, max_batch_pairs=2048, slots=2)
    tickets = [srv.submit(p, q) for p, q in reqs]
    got = [None] * len(tickets)

    def waiter(k):
        try:
            out, _, _ = srv.wait(tickets[k])
            got[k] = ("ok", out["distances"].copy())
        except ogjk.OgjkError as e:
            got[k] = ("err", str(e))

    threads = [threading.Thread(target=waiter, args=(k,)) for k in range(len(tickets))]
    for t in threads:
        t.start()
    time.sleep(0.5)  # every waiter is inside its (single) wait call
    done = threading.Event()
    closer = threading.Thread(target=lambda: (srv.close(), done.set()))
    closer.start()
    closer.join(timeout=60)
    assert done.is_set(), "destroy did not return"
    for t in threads:
        t.join(timeout=60)
    assert all(g is not None for g in got)
    ok = [k for k, g in enumerate(got) if g[0] == "ok"]
    cancelled = [k for k, g in enumerate(got) if g[0] == "err"]
    print(f"\n[destroy] {len(ok)} answered, {len(cancelled)} cancelled", flush=True)
    assert all("cancelled" in got[k][1] for k in cancelled)
    assert len(ok) >= 1
    for k in ok[:10]:
        p, q = reqs[k]
        dev = sc.collide_posed(oracle.GJK_EPA, torch.from_numpy(p).cuda(), torch.from_numpy(q).cuda())
        torch.cuda.synchronize()
        assert np.array_equal(got[k][1].view(np.uint32), dev["distances"].cpu().numpy().view(np.uint32))
    srv2 = ogjk.Server(sc, max_batch_pairs=2048, slots=2)
    try:
        out, _, _ = srv2.wait(srv2.submit(*reqs[0]))
        assert np.isfinite(out["distances"]).all()
    finally:
        srv2.close()
