"""Shared parity checks: GPU results (through the C ABI) vs the FP64 oracle.

Tolerances (stated once, used by every parity test):
  distance / penetration depth: |x - x_ref| / max(|x_ref|, FLOOR) <= 1e-5
  witnesses, contact normal:    Euclidean |v - v_ref| <= 1e-4
with FLOOR = 1e-3 (scene scale ~1: centres within a few units, radii 0.3-2).
The GPU path is designed to reproduce the FP64 oracle exactly, so the tests
also report (and bound) the fraction of pairs that match bit for bit.
"""
from __future__ import annotations

import numpy as np

import oracle

DIST_RTOL = 1e-5
VEC_ATOL = 1e-4
FLOOR = 1e-3


def same(a, b):
    """Elementwise equality that treats NaN == NaN."""
    a = np.asarray(a)
    b = np.asarray(b)
    return (a == b) | (np.isnan(a) & np.isnan(b))


def oracle_run(mode, batch, nthreads=8, simplices=None, distances=None):
    return oracle.run(mode, batch.coords, batch.off, batch.cnt, batch.pa, batch.pb, bits=64,
                      nthreads=nthreads, simplices=simplices, distances=distances)


def compare_f64(rec, simp32, ref, mode):
    """rec: (n,24) FP64 records from the GPU; simp32: gkSimplex records (FP32);
    ref: oracle.Result (FP64).  Returns a dict of counts and worst errors."""
    n = len(rec)
    d, dr = rec[:, 0], ref.distances
    w0, w1 = rec[:, 1:4], rec[:, 4:7]
    w0r, w1r = ref.simplices["witnesses"][:, 0], ref.simplices["witnesses"][:, 1]
    out = {"n": n}
    exact = same(d, dr) & same(w0, w0r).all(1) & same(w1, w1r).all(1)
    exact &= simp32["nvrtx"] == ref.simplices["nvrtx"]
    exact &= (simp32["vrtx_idx"] == ref.simplices["vrtx_idx"]).all((1, 2))
    exact &= same(rec[:, 10:22].reshape(n, 4, 3), ref.simplices["vrtx"]).all((1, 2))
    if mode & oracle.EPA:
        nr = rec[:, 7:10]
        exact &= same(nr, ref.normals).all(1)
        nerr = np.linalg.norm(nr - ref.normals, axis=1)
    else:
        nerr = np.zeros(n)
    derr = np.abs(d - dr) / np.maximum(np.abs(dr), FLOOR)
    werr = np.maximum(np.linalg.norm(w0 - w0r, axis=1), np.linalg.norm(w1 - w1r, axis=1))
    both_nan = np.isnan(d) & np.isnan(dr)
    derr = np.where(both_nan, 0.0, derr)
    werr = np.where(np.isnan(w0r).any(1) & np.isnan(w0).any(1), 0.0, werr)
    nerr = np.where(np.isnan(nerr) & np.isnan(ref.normals).any(1) if mode & oracle.EPA else False,
                    0.0, nerr)
    ok = (derr <= DIST_RTOL) & (werr <= VEC_ATOL) & (nerr <= VEC_ATOL)
    out.update(exact=int(exact.sum()), within_tol=int(ok.sum()), max_dist_rel=float(np.nanmax(derr, initial=0)),
               max_wit=float(np.nanmax(werr, initial=0)), max_nrm=float(np.nanmax(nerr, initial=0)),
               bad=np.where(~ok)[0][:10].tolist(), inexact=np.where(~exact)[0][:10].tolist())
    return out


def compare_f32(simp32, dist32, nrm32, ref, mode):
    """FP32 ABI outputs vs the FP64 oracle rounded once to float."""
    n = len(dist32)
    dr = ref.distances.astype(np.float32)
    exact = same(dist32, dr)
    exact &= same(simp32["witnesses"], ref.simplices["witnesses"].astype(np.float32)).all((1, 2))
    exact &= simp32["nvrtx"] == ref.simplices["nvrtx"]
    if mode & oracle.EPA:
        exact &= same(nrm32, ref.normals.astype(np.float32)).all(1)
    derr = np.abs(dist32.astype(np.float64) - ref.distances) / np.maximum(np.abs(ref.distances), FLOOR)
    derr = np.where(np.isnan(dist32) & np.isnan(ref.distances), 0.0, derr)
    return {"n": n, "exact": int(exact.sum()), "max_dist_rel": float(np.nanmax(derr, initial=0)),
            "inexact": np.where(~exact)[0][:10].tolist()}
