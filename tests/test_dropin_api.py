"""The reference-named C++ API (include/opengjk/gpu.h: GJK::GPU::*, the flat
openGJK.h functions, the *_device and *_indexed_device forms) run on the GPU
by a program built against it (tests/dropin_api_check.hip, built by
__graft_entry__.build()), as a program written for the reference calls it
(examples/main.cpp:378-644).  Every call's outputs are compared with the
FP64 oracle rounded once to the FP32 ABI:

  * GJK calls: distance, nvrtx, vrtx_idx, simplex vertices and witnesses
    equal the oracle's GJK bit for bit;
  * GJK+EPA calls: distance, witnesses, normal equal the oracle's GJK+EPA;
  * EPA calls on the simplices/distances a GJK call returned: penetrating
    pairs equal the oracle's GJK+EPA (the simplex vertices are rebuilt exactly
    from their indices); separated pairs keep the GJK distance and witnesses,
    and their normal is normalize(w1 - w0) of the FP32 witnesses in FP64
    (EPA.c:350-373 on the given simplex), also bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import ogjk
from ogjk import gen
import cases
from parity_utils import oracle_run

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "bin", "dropin_api_check")

GJK_CALLS = ["gjk_ns_1", "gjk_ns_2", "gjk_flat", "gjk_device", "gjk_indexed",
             "gjk_indexed_device_host_built", "gjk_indexed_device_device_built"]
GJK_EPA_CALLS = ["gjkepa_ns", "gjkepa_flat", "gjkepa_indexed"]
EPA_AFTER_GJK_CALLS = ["epa_ns", "epa_ci_ns", "epa_ci_readme", "epa_flat", "epa_indexed",
                       "epa_device", "epa_indexed_device_allocated",
                       "epa_indexed_device_device_built"]


def _batch():
    """Known answers, tie-heavy and degenerate pairs, and 64-vertex spheres
    (~60% penetrating)."""
    parts = [cases.known_answer_batch(), cases.degenerate_batch(600, seed=21),
             gen.make(gen.SPHERE, 1200, seed=22, nverts=64, box=1.3)]
    A, B = [], []
    for p in parts:
        for i in range(p.n_pairs):
            A.append(p.body(int(p.pa[i])))
            B.append(p.body(int(p.pb[i])))
    return gen.from_bodies(A, B)


@pytest.fixture(scope="module")
def run_program(tmp_path_factory):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(BIN):
        raise RuntimeError(f"{BIN} missing: run __graft_entry__.build()")
    b = _batch()
    d = tmp_path_factory.mktemp("dropin")
    with open(d / "input.bin", "wb") as f:
        np.array([b.n_pairs], np.int32).tofile(f)
        b.cnt.astype(np.int32).tofile(f)
        b.coords.astype(np.float32).tofile(f)
    print(f"\n[dropin] {b.n_pairs} pairs through {BIN}", flush=True)
    r = subprocess.run([BIN, str(d / "input.bin"), str(d)], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return b, d


def _read(d, name, n):
    raw = np.fromfile(d / f"{name}.bin", dtype=np.uint8)
    s = raw[:108 * n].view(ogjk.SIMPLEX_DTYPE)
    dist = raw[108 * n:112 * n].view(np.float32)
    nrm = raw[112 * n:124 * n].view(np.float32).reshape(n, 3)
    return s, dist, nrm


def _f32(x):
    return np.asarray(x, np.float64).astype(np.float32)


def _eq(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return (a == b) | (np.isnan(a) & np.isnan(b))


@pytest.fixture(scope="module")
def refs(run_program):
    b, _ = run_program
    return oracle_run(oracle.GJK, b), oracle_run(oracle.GJK_EPA, b)


@pytest.mark.parametrize("name", GJK_CALLS)
def test_gjk_calls(run_program, refs, name):
    b, d = run_program
    ref = refs[0]
    s, dist, _ = _read(d, name, b.n_pairs)
    ok = _eq(dist, _f32(ref.distances))
    ok &= s["nvrtx"] == ref.simplices["nvrtx"]
    ok &= (s["vrtx_idx"] == ref.simplices["vrtx_idx"]).all((1, 2))
    ok &= _eq(s["vrtx"], _f32(ref.simplices["vrtx"])).all((1, 2))
    ok &= _eq(s["witnesses"], _f32(ref.simplices["witnesses"])).all((1, 2))
    assert ok.all(), f"{name}: pairs differ from the oracle: {np.where(~ok)[0][:10].tolist()}"


@pytest.mark.parametrize("name", GJK_EPA_CALLS)
def test_gjk_epa_calls(run_program, refs, name):
    b, d = run_program
    ref = refs[1]
    s, dist, nrm = _read(d, name, b.n_pairs)
    ok = _eq(dist, _f32(ref.distances))
    ok &= _eq(s["witnesses"], _f32(ref.simplices["witnesses"])).all((1, 2))
    ok &= _eq(nrm, _f32(ref.normals)).all(1)
    assert ok.all(), f"{name}: pairs differ from the oracle: {np.where(~ok)[0][:10].tolist()}"
    assert (dist < 0).sum() > 500  # the batch does exercise EPA


@pytest.mark.parametrize("name", EPA_AFTER_GJK_CALLS)
def test_epa_after_gjk_calls(run_program, refs, name):
    b, d = run_program
    g, ge = refs
    s, dist, nrm = _read(d, name, b.n_pairs)
    sep = _f32(g.distances).astype(np.float64) > 2.220446049250313e-16
    pen = ~sep & np.isfinite(g.distances)
    # penetrating: the oracle's GJK+EPA, bit for bit
    okp = _eq(dist, _f32(ge.distances)) & _eq(nrm, _f32(ge.normals)).all(1)
    okp &= _eq(s["witnesses"], _f32(ge.simplices["witnesses"])).all((1, 2))
    bad = np.where(pen & ~okp)[0]
    assert len(bad) == 0, f"{name}: penetrating pairs differ: {bad[:10].tolist()}"
    # separated: GJK's distance and witnesses kept, normal from the FP32 witnesses
    w0 = _f32(g.simplices["witnesses"][:, 0]).astype(np.float64)
    w1 = _f32(g.simplices["witnesses"][:, 1]).astype(np.float64)
    dv = w1 - w0
    ln = np.sqrt((dv[:, 0] * dv[:, 0] + dv[:, 1] * dv[:, 1]) + dv[:, 2] * dv[:, 2])
    with np.errstate(invalid="ignore", divide="ignore"):
        n_ref = np.where((ln > 2.220446049250313e-16)[:, None], dv / ln[:, None], [1.0, 0.0, 0.0])
    oks = _eq(dist, _f32(g.distances)) & _eq(nrm, _f32(n_ref)).all(1)
    oks &= _eq(s["witnesses"], _f32(g.simplices["witnesses"])).all((1, 2))
    bad = np.where(sep & ~oks)[0]
    assert len(bad) == 0, f"{name}: separated pairs differ: {bad[:10].tolist()}"
    assert pen.sum() > 500 and sep.sum() > 300
