"""EPA edge cases on the GPU against the FP64 oracle, bit for bit:

* the reference's own EPA tests, cases 5-8 of examples/gpu/example.cu:593-892
  (50-point polytopes, 40x40-grid cubes with one rotated 45 degrees about
  every axis, 1,000-point spheres of radius 2 with centres 1 apart, fused and
  as separate GJK and EPA calls), with the reference's pass criteria;
* the EPA exits the reference reaches only rarely: the 64-iteration cap with
  its best-face fallback (GJK/cpu/EPA.c:828-863) and face-slot exhaustion
  (EPA.c:766-775), on pairs of the uneven workload where the oracle's own
  statistics show them.  The vertex cap (EPA.c:597, 131 vertices) cannot be
  reached: each iteration adds one vertex to the 4 of the tetrahedron, and
  there are at most 64 iterations.
"""
import numpy as np
import pytest

import oracle
import ogjk
import cases
from parity_utils import compare_f32, compare_f64, oracle_run

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _scene_run(torch, b, cluster_min=None, plain=False):
    sc = ogjk.DeviceScene(b, f64=plain)
    if cluster_min is not None:
        sc.build_clusters(cluster_min)
    out = sc.collide(oracle.GJK_EPA, b.n_pairs, f64_record=True, plain_f64=plain)
    torch.cuda.synchronize()
    return (out["f64"].cpu().numpy(), ogjk.simplices_from_device(out["simplices"]),
            out["distances"].cpu().numpy(), out["normals"].cpu().numpy())


def _exact(res, name):
    print(f"\n[{name}] {res}", flush=True)
    assert res["exact"] == res["n"], f"{name}: inexact pairs {res['inexact']}"


@pytest.mark.parametrize("cluster_min", [None, 1])
def test_reference_epa_cases(torch_cuda, cluster_min):
    b = cases.reference_epa_cases()
    ref = oracle_run(oracle.GJK_EPA, b)
    rec, simp, d, n = _scene_run(torch_cuda, b, cluster_min)
    _exact(compare_f64(rec, simp, ref, oracle.GJK_EPA), f"example.cu cases 5-7, clusters {cluster_min}")
    assert compare_f32(simp, d, n, ref, oracle.GJK_EPA)["exact"] == b.n_pairs
    # the reference's pass criteria (example.cu:624-640, 723-738, 785-801)
    assert simp["nvrtx"][0] == 4 and d[0] < 0.0                     # case 5
    w = simp["witnesses"]
    assert simp["nvrtx"][1] == 4 and d[1] < 0.0                     # case 6
    assert (np.abs(w[1, 0]) <= 2.0).all() and -1.0 <= w[1, 1, 0] <= 3.0
    assert (np.abs(w[1, 1, 1:]) <= 2.0).all()
    assert simp["nvrtx"][2] == 4 and -3.1 < d[2] < -2.5             # case 7: depth ~3
    assert np.linalg.norm(w[2, 0]) <= 2.1
    assert np.linalg.norm(w[2, 1] - np.float32([1, 0, 0])) <= 2.1


def test_reference_case8_separate_calls(torch_cuda):
    """Case 8: the sphere pair through computeDistances then computeEPA (the
    drop-in host path, two calls)."""
    b = cases.reference_epa_cases()
    ref = oracle_run(oracle.GJK_EPA, b)
    sg, dg, _ = ogjk.collide_host(oracle.GJK, b)
    se, de, ne = ogjk.collide_host(oracle.EPA, b, simplices=sg, distances=dg)
    r = compare_f32(se, de, ne, ref, oracle.GJK_EPA)  # all three pairs penetrate
    assert r["exact"] == r["n"], r
    assert -3.1 < de[2] < -2.5


@pytest.fixture(scope="module")
def cap_batch():
    b = cases.epa_cap_pairs()
    ref = oracle.run(oracle.GJK_EPA, b.coords, b.off, b.cnt, b.pa, b.pb, bits=64, nthreads=8,
                     stats=True)
    # the oracle reaches the cap on every pair, and runs out of face slots on most
    assert b.n_pairs >= 10 and (ref.stats["epa_iters"] == 64).all()
    assert (ref.stats["face_overflow"] > 0).sum() >= b.n_pairs // 2
    return b, ref


@pytest.mark.parametrize("path", ["pipeline", "clusters", "plain_f64", "lanes_16_16"])
def test_epa_cap_and_face_overflow(torch_cuda, cap_batch, path):
    b, ref = cap_batch
    if path == "lanes_16_16":
        ogjk.set_config(16, 16, 0)
    else:
        ogjk.set_config(64, 64, 0)
    try:
        rec, simp, d, n = _scene_run(torch_cuda, b, cluster_min=1 if path == "clusters" else None,
                                     plain=path == "plain_f64")
    finally:
        ogjk.set_config(8, 32, 0)
    _exact(compare_f64(rec, simp, ref, oracle.GJK_EPA), f"EPA cap pairs, {path}")
