"""ogjk_collide_scene_noalloc (the allocation-free variant of
ogjk_collide_scene, whose workspace is a static device array) against
ogjk_collide_scene, through the C ABI: two calls in a row, the second with a
smaller input, each bit-identical to the allocating entry point and to the
FP64 oracle.  Also: a pair count above OGJK_NOALLOC_MAX_PAIRS is rejected
on the host before anything is launched."""
import ctypes as C

import numpy as np
import pytest

import oracle
import ogjk
from ogjk import gen
from parity_utils import compare_f64, oracle_run

pytestmark = pytest.mark.gpu

NOALLOC_MAX_PAIRS = 65536  # OGJK_NOALLOC_MAX_PAIRS in include/opengjk/ogjk.h


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _call(fn, sc, n, out, stream):
    f = out["f64"].data_ptr()
    st = fn(ogjk.GJK_EPA, C.byref(sc.scene), None, n, out["simplices"].data_ptr(),
            out["distances"].data_ptr(), out["normals"].data_ptr(), f, stream)
    assert st == 0, ogjk.lib().ogjk_last_error().decode()


def test_noalloc_matches_collide_scene(torch_cuda):
    torch = torch_cuda
    L = ogjk.lib()
    ogjk.set_config(8, 32, False)
    # first call: 6000 pairs of 64-vertex spheres (~60% penetrate); second:
    # 1500 pairs of an uneven batch, so stale counters or jobs of the first
    # call would show up as wrong results
    big = gen.make(gen.SPHERE, 6000, seed=31, nverts=64, box=1.29)
    small = gen.make(gen.ELLIPSOID, 1500, seed=32, box=1.5, nmin=4, nmax=600)
    for step, b in enumerate((big, small)):
        print(f"\n[noalloc] call {step}: {b.n_pairs} pairs", flush=True)
        sc = ogjk.DeviceScene(b)
        st = torch.cuda.current_stream().cuda_stream
        ref_out = sc.outputs(b.n_pairs, f64=True)
        new_out = sc.outputs(b.n_pairs, f64=True)
        _call(L.ogjk_collide_scene, sc, b.n_pairs, ref_out, st)
        _call(L.ogjk_collide_scene_noalloc, sc, b.n_pairs, new_out, st)
        torch.cuda.synchronize()
        for k in ("simplices", "distances", "normals", "f64"):
            a, r = new_out[k].cpu().numpy(), ref_out[k].cpu().numpy()
            assert np.array_equal(a.view(np.uint8), r.view(np.uint8)), f"call {step}: {k} differs"
        ref = oracle_run(oracle.GJK_EPA, b)
        res = compare_f64(new_out["f64"].cpu().numpy(),
                          ogjk.simplices_from_device(new_out["simplices"]), ref, oracle.GJK_EPA)
        assert res["exact"] == res["n"], res


def test_noalloc_rejects_oversize(torch_cuda):
    L = ogjk.lib()
    b = gen.make(gen.SPHERE, 4, seed=33, nverts=8, box=1.0)
    sc = ogjk.DeviceScene(b)
    out = sc.outputs(4, f64=True)
    # argument checks run on the host before any launch: no kernel touches
    # these buffers with the oversize count
    st = L.ogjk_collide_scene_noalloc(ogjk.GJK_EPA, C.byref(sc.scene), None,
                                      NOALLOC_MAX_PAIRS + 1, out["simplices"].data_ptr(),
                                      out["distances"].data_ptr(), out["normals"].data_ptr(),
                                      out["f64"].data_ptr(), None)
    assert st == 1  # OGJK_ERR_INVALID
    assert b"exceed" in L.ogjk_last_error()
