#!/usr/bin/env python3
"""Benchmark: batched GJK+EPA pairs/s on one MI355X.

Workload (default): 100,000 pairs of 64-vertex polytopes, points uniform on
spheres of radius U[0.5,1.5] with centres in U[-1.29,1.29]^3 (about 60% of
pairs penetrate), GJK then EPA for every pair, FP32 vertex storage with the
FP32 support sweep + exact FP64 re-rank, FP64 sub-distance/EPA arithmetic.

A step is one GJK+EPA pass over the whole batch (GJK kernel, device worklist
of penetrating pairs, persistent EPA kernel), inputs resident in HBM as an
SoA scene.  K steps are timed between device synchronisations; per-step
times come from HIP events recorded on the launch stream.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "opengjk-gpu_amd")]

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "sphere64": dict(family=0, n_pairs=100_000, seed=1, nverts=64, box=1.29,
                     desc="100000 pairs x 64-vertex sphere polytopes, radius U[0.5,1.5], "
                          "centres U[-1.29,1.29]^3 (~60% penetrating), GJK then EPA"),
    "ellipsoid_uneven": dict(family=2, n_pairs=1_000_000, seed=2, box=1.5, nmin=4, nmax=4096,
                             desc="1000000 pairs, vertex counts log-uniform in [4,4096], "
                                  "ellipsoids axes U[0.3,2] rotated, centres U[-1.5,1.5]^3, GJK then EPA"),
}
# lanes per pair (GJK, EPA) used when --lanes is not given
DEFAULT_LANES = {"sphere64": "8,32", "ellipsoid_uneven": "64,64"}
# rocprofv3 summary holding the measured HBM traffic of one step of the workload
PROFILE = {"sphere64": "gjk_epa_sphere64_100k.md", "ellipsoid_uneven": "gjk_epa_ellipsoid_uneven_1m.md"}


def measured_traffic(workload: str):
    """HBM bytes per step from the committed rocprofv3 summary (FETCH_SIZE and
    WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", PROFILE.get(workload, ""))
    if not os.path.isfile(path):
        return None
    for line in open(path):
        if line.strip().startswith("traffic_bytes_per_step:"):
            return float(line.split(":", 1)[1].strip())
    return None


def bytes_per_launch(batch, mode_epa=True) -> float:
    """Algorithmic HBM bytes of one launch: every vertex read once (12 B FP32
    SoA), the two body-table records per pair (off 8 + cnt 4 + r1 4 B each),
    and the results written through the gkSimplex ABI (108 B simplex + 4 B
    distance + 12 B normal)."""
    verts = int(batch.cnt[batch.pa].astype(np.int64).sum() + batch.cnt[batch.pb].astype(np.int64).sum())
    n = batch.n_pairs
    return 12.0 * verts + 32.0 * n + (124.0 if mode_epa else 112.0) * n


def cpu_baseline(batch, gpu_out, budget_s: float = 20.0):
    """CPU baseline leg: the FP64 oracle (the reference CPU algorithm restated)
    on this box's cores, over a bounded prefix of the same workload.  The GPU
    results for the same pairs are checked against it here as well."""
    import oracle
    from ogjk import gen
    cores = min(len(os.sched_getaffinity(0)), 16)
    n = batch.n_pairs
    # size the sample from a quick probe so the leg stays within budget_s
    probe = min(n, 2000)
    sub = gen.Batch(batch.coords, batch.off, batch.cnt, batch.pa[:probe], batch.pb[:probe])
    t = time.perf_counter()
    oracle.run(3, sub.coords, sub.off, sub.cnt, sub.pa, sub.pb, bits=64, nthreads=1)
    per_pair_1c = (time.perf_counter() - t) / probe
    m = int(min(n, max(probe, budget_s * 0.5 * cores / max(per_pair_1c, 1e-9))))
    sub = gen.Batch(batch.coords, batch.off, batch.cnt, batch.pa[:m], batch.pb[:m])
    t = time.perf_counter()
    ref = oracle.run(3, sub.coords, sub.off, sub.cnt, sub.pa, sub.pb, bits=64, nthreads=cores)
    el = time.perf_counter() - t
    # parity of the timed GPU results on the same pairs (FP64 records)
    rec = gpu_out["f64"][:m].cpu().numpy()
    d, dr = rec[:, 0], ref.distances
    rel = np.abs(d - dr) / np.maximum(np.abs(dr), 1e-3)
    w = np.maximum(np.linalg.norm(rec[:, 1:4] - ref.simplices["witnesses"][:, 0], axis=1),
                   np.linalg.norm(rec[:, 4:7] - ref.simplices["witnesses"][:, 1], axis=1))
    nn = np.linalg.norm(rec[:, 7:10] - ref.normals, axis=1)
    exact = (d == dr) & (rec[:, 1:4] == ref.simplices["witnesses"][:, 0]).all(1) & \
            (rec[:, 4:7] == ref.simplices["witnesses"][:, 1]).all(1) & (rec[:, 7:10] == ref.normals).all(1)
    parity = {"pairs_checked": m, "max_rel_err_distance": float(np.nanmax(rel)),
              "max_abs_err_witness": float(np.nanmax(w)), "max_abs_err_normal": float(np.nanmax(nn)),
              "bit_exact_fraction": float(exact.mean()),
              "overlap_fraction": float((dr <= 2.220446049250313e-16).mean())}
    base = {"value": m / el, "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": f"first {m} pairs of the workload, FP64 oracle (restated reference CPU "
                      f"openGJK+EPA), {cores} threads, {el:.2f}s; single-core "
                      f"{1.0 / per_pair_1c:.0f} pairs/s"}
    return base, parity


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sphere64", choices=sorted(WORKLOADS))
    ap.add_argument("--pairs", type=int, default=0, help="override the pair count")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-FP64 comparison")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 256 MiB elementwise copy (reads and writes 256 MiB) so a "
                         "profiler's FETCH_SIZE/WRITE_SIZE can be scaled against known bytes")
    ap.add_argument("--lanes", default=None, help="GJK,EPA lanes per pair (default per workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent and would shard by contiguous ranges; see DESIGN.md)",
              file=sys.stderr)
        return 2

    import torch
    import ogjk
    from ogjk import gen

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    lanes = args.lanes or DEFAULT_LANES[args.workload]
    lg, le = (int(x) for x in lanes.split(","))
    ogjk.set_config(lg, le, False)
    if args.calibrate:
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        for _ in range(3):
            y = x * 1.0
        torch.cuda.synchronize()
        del x, y
    wl = dict(WORKLOADS[args.workload])
    desc = wl.pop("desc")
    if args.pairs:
        wl["n_pairs"] = args.pairs
    t0 = time.perf_counter()
    batch = gen.make(nthreads=min(len(os.sched_getaffinity(0)), 16), **wl)
    print(f"# generated {batch.n_pairs} pairs, {batch.coords.shape[0]} vertices "
          f"in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    n = batch.n_pairs
    scene = ogjk.DeviceScene(batch, f64=not args.no_plain)
    out = scene.outputs(n, f64=True)
    stream = torch.cuda.current_stream()

    def step(o, plain=False, rec=False):
        scene.collide(3, n, out=o, plain_f64=plain, f64_record=rec)

    # verification record for the parity leg (untimed)
    step(out, rec=True)
    torch.cuda.synchronize()
    timed_out = scene.outputs(n, f64=False)

    def timed(plain: bool):
        for _ in range(args.warmup):
            step(timed_out, plain)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        t = time.perf_counter()
        for a, b in evs:
            a.record(stream)
            step(timed_out, plain)
            b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        ms = np.array([a.elapsed_time(b) for a, b in evs])
        return wall, ms

    wall, ms = timed(False)
    print(f"# mixed: {wall / args.steps * 1e3:.3f} ms/step wall, kernel median {np.median(ms):.3f} ms",
          file=sys.stderr, flush=True)
    plain = pipe64 = None
    if not args.no_plain:
        # the same split pipeline over FP64 vertex storage (isolates the FP32 sweep)
        wall_q, ms_q = timed(True)
        pipe64 = {"pairs_per_s": n * args.steps / wall_q, "ms_per_step": wall_q / args.steps * 1e3}
        print(f"# pipeline over fp64 storage: {pipe64['ms_per_step']:.3f} ms/step", file=sys.stderr,
              flush=True)
        # plain FP64 per-pair kernel: one group per pair runs GJK then EPA in one
        # non-persistent launch over FP64 SoA vertices, no worklist; the faster
        # of 32 and 64 lanes per pair is the denominator
        plain = {"variants": {}}
        for el in (32, 64):
            ogjk.set_config(lg, el, True)
            wall_p, ms_p = timed(True)
            plain["variants"][f"{el}_lanes"] = {"pairs_per_s": n * args.steps / wall_p,
                                                "ms_per_step": wall_p / args.steps * 1e3}
            print(f"# plain fp64 per-pair kernel, {el} lanes: {wall_p / args.steps * 1e3:.3f} ms/step",
                  file=sys.stderr, flush=True)
        # the plain kernel's results equal the mixed pipeline's, pair by pair
        chk = scene.outputs(n, f64=True)
        scene.collide(3, n, out=chk, plain_f64=True, f64_record=True)
        torch.cuda.synchronize()
        plain["identical_to_mixed_fraction"] = float((chk["f64"] == out["f64"]).all(dim=1).float().mean())
        del chk
        ogjk.set_config(lg, le, False)
        best = max(plain["variants"].values(), key=lambda v: v["pairs_per_s"])
        plain.update(best)

    value = n * args.steps / wall
    algo = bytes_per_launch(batch)
    k_ms = float(np.mean(ms))
    achieved = algo / (k_ms * 1e-3) / 1e9
    line = {
        "metric": "GJK+EPA pairs/s vs plain FP64 GPU path at 1e-5 rel err; p50/p99 query latency",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": desc, "pairs": n, "vertices": int(batch.coords.shape[0]),
                   "lanes_per_pair": {"gjk": lg, "epa": le},
                   "layout": "SoA fp32 scene; GJK kernel -> device worklist -> persistent EPA kernel"},
        "step_ms_p50": float(np.percentile(ms, 50)), "step_ms_p99": float(np.percentile(ms, 99)),
        "plain_f64": plain,
        "speedup_vs_plain_f64": (value / plain["pairs_per_s"]) if plain else None,
        "pipeline_f64_storage": pipe64,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload),
                     "kernel": "k_gjk + k_epa_list (one step)",
                     "kernel_ms_mean": k_ms, "algorithmic_bytes_per_launch": algo},
    }
    if not args.no_cpu:
        base, parity = cpu_baseline(batch, out)
        line["cpu_baseline"] = base
        line["parity"] = parity
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
