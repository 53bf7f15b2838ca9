#!/usr/bin/env python3
"""Benchmark: batched GJK+EPA pairs/s on one MI355X.

Workloads (--workload):
  sphere64 (default): 100,000 pairs of 64-vertex polytopes, points uniform on
      spheres of radius U[0.5,1.5], centres U[-1.29,1.29]^3 (about 60% of pairs
      penetrate).
  ellipsoid_uneven: 1,000,000 pairs, vertex counts log-uniform in [4, 4096],
      points on rotated ellipsoids (axes U[0.3,2]), centres U[-1.5,1.5]^3.
  noisy_sphere_large: 10,000 pairs of 100,000-vertex polytopes, unit-sphere
      points perturbed by N(0, 0.01^2) per coordinate, centres U[-1.5,1.5]^3
      (about 50% penetrate).
  serving, broadphase, scene_step, hulls: see the functions of those names.
Every pair runs GJK then EPA: FP32 vertex storage, FP32 support sweep with an
exact FP64 re-rank, FP64 sub-distance and EPA arithmetic; the larger
workloads sweep large bodies through their vertex clusters.

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
EPS64 = 2.220446049250313e-16

# gen: generator arguments; lanes: (GJK, EPA) lanes per pair; cluster_min:
# bodies with at least this many vertices are swept through their clusters
# (None: no clusters); roofline: "compulsory" (every vertex once) or
# "stream" (every body re-streamed once per oracle sweep); profile: the
# rocprofv3 summary holding the measured HBM traffic of one step.
WORKLOADS = {
    "sphere64": dict(
        gen=dict(family=0, n_pairs=100_000, seed=1, nverts=64, box=1.29),
        desc="100000 pairs x 64-vertex sphere polytopes, radius U[0.5,1.5], centres "
             "U[-1.29,1.29]^3 (~60% penetrating), GJK then EPA",
        lanes=(8, 32), cluster_min=None, roofline="compulsory",
        profile="gjk_epa_sphere64_100k.md"),
    "ellipsoid_uneven": dict(
        gen=dict(family=2, n_pairs=1_000_000, seed=2, box=1.5, nmin=4, nmax=4096),
        desc="1000000 pairs, vertex counts log-uniform in [4,4096], ellipsoids axes U[0.3,2] "
             "rotated, centres U[-1.5,1.5]^3, GJK then EPA",
        lanes=(64, 64), cluster_min=512, roofline="compulsory",
        profile="gjk_epa_ellipsoid_uneven_1m.md", tuning="ellipsoid_uneven_1m.json"),
    "noisy_sphere_large": dict(
        gen=dict(family=3, n_pairs=10_000, seed=3, nverts=100_000, box=1.5, sigma=0.01),
        desc="10000 pairs x 100000-vertex polytopes, unit-sphere points + N(0,0.01^2), centres "
             "U[-1.5,1.5]^3 (~50% penetrating), GJK then EPA",
        lanes=(64, 64), cluster_min=1024, roofline="stream",
        profile="gjk_epa_noisy_sphere_100k.md"),
}


# rocprofv3 summaries of the legs that are not in WORKLOADS
LEG_PROFILES = {"broadphase": "broadphase_1m.md", "scene_step": "scene_step_20k.md",
                "hulls": "hulls_100k.md"}


def measured_traffic(workload: str):
    """HBM bytes per step from the committed rocprofv3 summary (FETCH_SIZE and
    WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    name = WORKLOADS[workload]["profile"] if workload in WORKLOADS else LEG_PROFILES.get(workload)
    if name is None:
        return None
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.isfile(path):
        return None
    for line in open(path):
        if line.strip().startswith("traffic_bytes_per_step:"):
            return float(line.split(":", 1)[1].strip())
    return None


def compulsory_bytes(batch) -> float:
    """Every vertex read once (12 B FP32 SoA), the two body-table records per
    pair (off 8 + cnt 4 + r1 4 B each), and the results written through the
    gkSimplex ABI (108 B simplex + 4 B distance + 12 B normal)."""
    verts = int(batch.cnt[batch.pa].astype(np.int64).sum() + batch.cnt[batch.pb].astype(np.int64).sum())
    n = batch.n_pairs
    return 12.0 * verts + 32.0 * n + 124.0 * n


def cpu_baseline(batch, gpu_out, budget_s: float = 20.0):
    """CPU baseline leg: the FP64 oracle (the reference CPU algorithm restated)
    on this box's cores, over a bounded prefix of the same workload.  The GPU
    results for the same pairs are checked against it here as well.  Also
    returns the oracle's stream bytes per pair (12 B x vertices x sweeps)."""
    import oracle
    from ogjk import gen
    cores = min(len(os.sched_getaffinity(0)), 16)
    n = batch.n_pairs

    def prefix(m):
        return gen.Batch(batch.coords, batch.off, batch.cnt, batch.pa[:m], batch.pb[:m])

    # single-core probe, doubled until it takes 0.3 s, sizes the sample
    probe, el1 = 8, 0.0
    while True:
        probe = min(n, probe * 2)
        sub = prefix(probe)
        t = time.perf_counter()
        oracle.run(3, sub.coords, sub.off, sub.cnt, sub.pa, sub.pb, bits=64, nthreads=1)
        el1 = time.perf_counter() - t
        if el1 > 0.3 or probe == n:
            break
    per_pair_1c = el1 / probe
    m = int(min(n, max(probe, budget_s * 0.5 * cores / max(per_pair_1c, 1e-9))))
    sub = prefix(m)
    t = time.perf_counter()
    ref = oracle.run(3, sub.coords, sub.off, sub.cnt, sub.pa, sub.pb, bits=64, nthreads=cores,
                     stats=True)
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
              "overlap_fraction": float((dr <= EPS64).mean())}
    st = ref.stats
    na = batch.cnt[batch.pa[:m]].astype(np.float64)
    nb = batch.cnt[batch.pb[:m]].astype(np.float64)
    stream_per_pair = float((12.0 * (na * st["sweeps_a"] + nb * st["sweeps_b"])).mean())
    base = {"value": m / el, "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": f"first {m} pairs of the workload, FP64 oracle (restated reference CPU "
                      f"openGJK+EPA), {cores} threads, {el:.2f}s; single-core "
                      f"{1.0 / per_pair_1c:.1f} pairs/s"}
    return base, parity, stream_per_pair


def serving(args) -> int:
    """Serving leg: a resident scene of 20,000 bodies (8-512 vertices,
    log-uniform, ellipsoids in their local frames); requests arrive by a
    Poisson process at --qps, each with K ~ U{1..10000} pairs (i != j
    uniform) and a fresh pose per pair side (uniform unit quaternion,
    translation U[-10,10]^3); GJK+EPA through the serving loop.  Latency is
    submit to results in host memory."""
    import threading
    import torch
    import ogjk
    from ogjk import gen
    ogjk.set_config(32, 64, False)  # hulls of up to 512 vertices: measured best
    b = gen.make(gen.ELLIPSOID, 10_000, seed=4, box=0.0, nmin=8, nmax=512,
                 nthreads=min(len(os.sched_getaffinity(0)), 16))
    scene = ogjk.DeviceScene(b)
    nb = len(b.cnt)
    n_req = args.requests
    reqs = gen.posed_requests(nb, n_req, seed=4)
    total_pairs = int(sum(len(p) for p, _ in reqs))
    print(f"# scene {nb} bodies, {b.coords.shape[0]} vertices; {n_req} requests, {total_pairs} pairs",
          file=sys.stderr, flush=True)

    # kernel roofline: one device launch of 10,000 posed pairs, HIP events
    kp, kq = reqs[0]
    while len(kp) < 10_000:
        extra = gen.posed_requests(nb, 1, seed=len(kp) + 99)[0]
        kp, kq = np.concatenate([kp, extra[0]]), np.concatenate([kq, extra[1]])
    kp, kq = torch.from_numpy(kp[:10_000]).cuda(), torch.from_numpy(kq[:10_000]).cuda()
    kout = scene.collide_posed(3, kp, kq)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(20):
        scene.collide_posed(3, kp, kq, out=kout)
    ev[1].record()
    torch.cuda.synchronize()
    k_ms = ev[0].elapsed_time(ev[1]) / 20
    cnt = b.cnt.astype(np.int64)
    kpn = kp.cpu().numpy()
    k_bytes = float(12 * (cnt[kpn[:, 0]] + cnt[kpn[:, 1]]).sum() + 10_000 * (8 + 56 + 32 + 124))

    def run_at(qps, rng_seed):
        srv = ogjk.Server(scene, max_batch_pairs=65536, slots=args.slots)
        rng = np.random.default_rng(rng_seed)
        gaps = rng.exponential(1.0 / qps, size=n_req)
        arrive = np.cumsum(gaps)
        tickets = [None] * n_req
        done_ev = threading.Event()

        def submitter():
            t0 = time.perf_counter()
            for k in range(n_req):
                while time.perf_counter() - t0 < arrive[k]:
                    pass
                tickets[k] = srv.submit(*reqs[k])
            done_ev.set()

        th = threading.Thread(target=submitter)
        t_start = time.perf_counter()
        th.start()
        lats = np.zeros(n_req)
        for k in range(n_req):
            while tickets[k] is None:
                time.sleep(0)
            _, lat, _ = srv.wait(tickets[k])
            lats[k] = lat
        wall = time.perf_counter() - t_start
        th.join()
        info = srv.info()
        srv.close()
        return {"qps": qps, "p50_ms": float(np.percentile(lats, 50) * 1e3),
                "p99_ms": float(np.percentile(lats, 99) * 1e3),
                "max_ms": float(lats.max() * 1e3), "pairs_per_s": total_pairs / wall,
                "launches": info["launches"], "slots": info["slots"], "wall_s": wall}

    sweep = []
    for qps in sorted(set([args.qps] + [250, 500, 2000, 4000, 6000])):
        r = run_at(qps, 17 + qps)
        sweep.append(r)
        print(f"# qps {qps}: p50 {r['p50_ms']:.3f} ms p99 {r['p99_ms']:.3f} ms, "
              f"{r['pairs_per_s'] / 1e6:.2f} M pairs/s, {r['launches']} launches", file=sys.stderr,
              flush=True)
    main_r = next(r for r in sweep if r["qps"] == args.qps)
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    line = {
        "metric": "GJK+EPA pairs/s vs plain FP64 GPU path at 1e-5 rel err; p50/p99 query latency",
        "value": main_r["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1, "steps": n_req,
        "warmup": 0, "ms_per_step": main_r["wall_s"] / n_req * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "serving: 20000-body scene (8-512 vertices log-uniform), Poisson "
                               f"requests at {args.qps} QPS, K ~ U{{1..10000}} pairs, fresh pose per "
                               "pair side, t ~ U[-10,10]^3", "requests": n_req,
                   "pairs": total_pairs, "lanes_per_pair": {"gjk": 32, "epa": 64},
                   "server": {"slots": main_r["slots"], "max_batch_pairs": 65536}},
        "latency_ms": {"p50": main_r["p50_ms"], "p99": main_r["p99_ms"], "max": main_r["max_ms"]},
        "qps_sweep": sweep,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_gjk + k_epa_list over 10000 posed pairs (one launch)",
                     "kernel_ms_mean": k_ms, "algorithmic_bytes_per_launch": k_bytes,
                     "algorithmic_bytes": "compulsory: vertices once + pair, poses, record, results"},
    }
    if not args.no_cpu:
        base, parity = serving_cpu_baseline(scene, b, reqs)
        line["cpu_baseline"] = base
        line["parity"] = parity
    print(json.dumps(line), flush=True)
    return 0


def serving_cpu_baseline(scene, b, reqs, max_pairs: int = 200_000):
    """The FP64 oracle on the world vertices of the first requests' pairs (up
    to max_pairs), timed on this box's cores; the server's FP64 records for
    the same requests are compared with it."""
    import oracle
    import ogjk
    cores = min(len(os.sched_getaffinity(0)), 16)
    take, n = [], 0
    for p, q in reqs:
        if n >= max_pairs:
            break
        take.append((p, q))
        n += len(p)
    pairs = np.concatenate([p for p, _ in take])
    poses = np.concatenate([q for _, q in take])
    pool, off, cnt, pa, pb = oracle.posed_pool(b.coords, b.off, b.cnt, pairs, poses)
    t = time.perf_counter()
    ref = oracle.run(3, pool, off, cnt, pa, pb, bits=64, nthreads=cores)
    el = time.perf_counter() - t
    srv = ogjk.Server(scene, max_batch_pairs=65536, slots=2)
    try:
        outs = [srv.wait(srv.submit(p, q, f64=True))[0] for p, q in take]
    finally:
        srv.close()
    rec = np.concatenate([o["f64"] for o in outs])
    d, dr = rec[:, 0], ref.distances
    exact = (d == dr) & (rec[:, 1:4] == ref.simplices["witnesses"][:, 0]).all(1) & \
            (rec[:, 4:7] == ref.simplices["witnesses"][:, 1]).all(1) & (rec[:, 7:10] == ref.normals).all(1)
    rel = np.abs(d - dr) / np.maximum(np.abs(dr), 1e-3)
    parity = {"pairs_checked": int(len(d)), "max_rel_err_distance": float(np.nanmax(rel)),
              "bit_exact_fraction": float(exact.mean()),
              "overlap_fraction": float((dr <= EPS64).mean())}
    base = {"value": len(d) / el, "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": f"pairs of the first {len(take)} requests ({len(d)} pairs), FP64 oracle on "
                      f"their FP64 world vertices, {cores} threads, {el:.2f}s (transform untimed)"}
    return base, parity


def _event_steps(torch, fn, steps: int, warmup: int):
    """Warm up, then run `steps` calls of fn between HIP events on the
    current stream.  Returns (wall seconds, per-step event ms)."""
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    t = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    return wall, np.array([a.elapsed_time(b) for a, b in evs])


def broadphase(args) -> int:
    """Broad-phase leg: 1,000,000 spheres (centres U[-100,100]^3, radii
    U[0.2,1]) on a 100^3 grid of 2.0 cells; a step is one ogjk_broad_phase
    call (cell keys, radix sort, start table, count, scan, emit, total read
    back).  Throughput in objects/s."""
    import torch
    import oracle
    import ogjk
    from ogjk import gen
    n = args.pairs or 1_000_000
    box = 100.0 * (n / 1_000_000) ** (1.0 / 3.0)
    g = max(1, int(box))
    s = gen.spheres(n, seed=12, box=box, rmin=0.2, rmax=1.0)
    grid = ogjk.Grid(2.0 * box / g, box, g)
    d = torch.from_numpy(s).cuda()
    bp = ogjk.BroadPhase(n, grid, max_pairs=8 * n)
    pairs, found = bp(d)
    torch.cuda.synchronize()
    got = pairs.cpu().numpy()
    wall, ms = _event_steps(torch, lambda: bp(d), args.steps, args.warmup)
    ms_step = wall / args.steps * 1e3
    # algorithmic bytes: every sphere read once (16 B), every pair written once (8 B)
    alg = 16.0 * n + 8.0 * found
    achieved = alg / (np.mean(ms) * 1e-3) / 1e9
    line = {
        "metric": "broad-phase objects/s (uniform grid -> gkCollisionPair list)",
        "value": n / (wall / args.steps), "unit": "objects/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"broad phase: {n} spheres, centres U[-{box:.1f},{box:.1f}]^3, radii "
                               f"U[0.2,1], grid {g}^3 cells of {grid.cell_size:.3f}",
                   "objects": n, "pairs_found": found},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("broadphase"),
                     "kernel": "whole broad phase (10 launches + total readback)",
                     "kernel_ms_mean": float(np.mean(ms)), "algorithmic_bytes_per_launch": alg,
                     "algorithmic_bytes": "16 B per sphere read + 8 B per pair written"},
    }
    if not args.no_cpu:
        t = time.perf_counter()
        ref, rfound = oracle.broad_phase(s, grid.cell_size, grid.boundary, grid.grid_size)
        el = time.perf_counter() - t
        line["cpu_baseline"] = {
            "value": n / el, "unit": "objects/s", "cores": 1, "kind": "port",
            "sample": f"the whole workload ({n} spheres), C oracle of the reference grid, one "
                      f"thread, {el:.2f}s"}
        line["parity"] = {"pairs_found": found, "oracle_found": rfound,
                          "identical_list": bool(rfound == found and np.array_equal(got, ref))}
    print(json.dumps(line), flush=True)
    return 0


def scene_step(args) -> int:
    """Scene-step leg: the serving scene's 20,000 bodies (8-512 vertices,
    log-uniform ellipsoids in local frames), one pose per body (uniform unit
    quaternion, translation U[-40,40]^3), bounding spheres through the device
    broad phase (grid of at most 64^3 cells, cells one largest diameter
    wide), then GJK+EPA on every pair found, written as 40-byte contact
    records.  A step is both phases."""
    import torch
    import oracle
    import ogjk
    from ogjk import gen
    ogjk.set_config(32, 64, False)  # broad-phase pairs are close: measured best
    b = gen.make(gen.ELLIPSOID, 10_000, seed=4, box=0.0, nmin=8, nmax=512,
                 nthreads=min(len(os.sched_getaffinity(0)), 16))
    nb = len(b.cnt)
    radii = gen.body_radii(b.coords, b.off, b.cnt)
    tbox = 40.0
    poses = gen.body_poses(nb, seed=7, tbox=tbox)
    grid = ogjk.Grid.for_spheres(float(radii.max()), tbox, 64)
    sc = ogjk.DeviceScene(b)
    step = ogjk.SceneStep(sc, radii, grid, max_pairs=20 * nb, contacts=True)
    dp = torch.from_numpy(poses).cuda()
    pairs, found, out = step(dp)
    torch.cuda.synchronize()
    got = pairs.cpu().numpy()
    n = len(got)
    wall, ms = _event_steps(torch, lambda: step(dp), args.steps, args.warmup)
    # narrow phase alone on the same pairs, for the kernel roofline
    ws = step.workspace
    nwall, nms = _event_steps(torch, lambda: sc.collide_posed_bodies_contacts(
        3, pairs, dp, out=out, workspace=ws), args.steps, args.warmup)
    cnt = b.cnt.astype(np.int64)
    # vertices once per pair side, pair 8 B, two poses 56 B, body table 32 B, contact record 40 B
    alg = float(12 * (cnt[got[:, 0]] + cnt[got[:, 1]]).sum() + n * (8 + 56 + 32 + 40))
    achieved = alg / (np.mean(nms) * 1e-3) / 1e9
    line = {
        "metric": "scene step pairs/s (broad phase + posed GJK+EPA on the pairs found)",
        "value": n / (wall / args.steps), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"scene step: {nb} bodies (8-512 vertices log-uniform), one pose "
                               f"per body, t ~ U[-{tbox:.0f},{tbox:.0f}]^3, {grid}",
                   "bodies": nb, "pairs": n, "lanes_per_pair": {"gjk": 32, "epa": 64}},
        "narrow_phase_ms": float(np.mean(nms)),
        "broad_phase_ms": float(np.mean(ms) - np.mean(nms)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("scene_step"),
                     "kernel": f"k_gjk + k_epa_list over the {n} pairs of a step",
                     "kernel_ms_mean": float(np.mean(nms)), "algorithmic_bytes_per_launch": alg,
                     "algorithmic_bytes": "compulsory: vertices once + pair, poses, body "
                                          "table, 40-byte contact record"},
    }
    if not args.no_cpu:
        cores = min(len(os.sched_getaffinity(0)), 16)
        spheres = np.concatenate([poses[:, 4:7], radii[:, None]], 1)
        t = time.perf_counter()
        ref_pairs, rfound = oracle.broad_phase(spheres, grid.cell_size, grid.boundary,
                                               grid.grid_size)
        side = np.concatenate([poses[ref_pairs[:, 0]], poses[ref_pairs[:, 1]]], 1)
        t_bp = time.perf_counter() - t
        pool, off, cnt2, pa, pb = oracle.posed_pool(b.coords, b.off, b.cnt, ref_pairs, side)
        t = time.perf_counter()
        ref = oracle.run(3, pool, off, cnt2, pa, pb, bits=64, nthreads=cores)
        t_np = time.perf_counter() - t
        el = t_bp + t_np
        step_out = step.scene.collide_posed_bodies(3, pairs, dp, f64_record=True)
        torch.cuda.synchronize()
        rec = step_out["f64"][:n].cpu().numpy()
        exact = ((rec[:, 0] == ref.distances) &
                 (rec[:, 1:4] == ref.simplices["witnesses"][:, 0]).all(1) &
                 (rec[:, 4:7] == ref.simplices["witnesses"][:, 1]).all(1) &
                 (rec[:, 7:10] == ref.normals).all(1)) if len(ref_pairs) == n else np.zeros(1)
        line["cpu_baseline"] = {
            "value": rfound / el, "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": f"one whole step: C oracle broad phase (1 thread, {t_bp:.2f}s) + FP64 "
                      f"oracle GJK+EPA on the world vertices of the {rfound} pairs ({cores} "
                      f"threads, {t_np:.2f}s; world transform untimed)"}
        line["parity"] = {"pairs_found": found, "oracle_found": rfound,
                          "identical_pair_list": bool(rfound == found and
                                                      np.array_equal(got, ref_pairs)),
                          "bit_exact_fraction": float(exact.mean()),
                          "overlap_fraction": float((ref.distances <= EPS64).mean())}
    print(json.dumps(line), flush=True)
    return 0


def hulls(args) -> int:
    """Mesh-ingestion leg: 100,000 point clouds of 8-32 points uniform on the
    unit sphere (the reference's random hull variants, MIN_HULL_VERTS ..
    MAX_HULL_VERTS); a step ingests them all as the reference does each shape
    (convex hull, centre of mass, centring, hull of the centred points and its
    inertia factors).  Throughput in clouds/s."""
    import torch
    import oracle
    import ogjk
    from concurrent.futures import ThreadPoolExecutor
    n = args.pairs or 100_000
    rng = np.random.default_rng(5)
    cnt = rng.integers(8, 33, n)
    u = rng.normal(size=(int(cnt.sum()), 3))
    u = (u / np.linalg.norm(u, axis=1, keepdims=True)).astype(np.float32)
    off = np.zeros(n, np.int64)
    off[1:] = np.cumsum(cnt)[:-1]
    pts = torch.from_numpy(u).cuda()
    doff = torch.from_numpy(off).cuda()
    dcnt = torch.from_numpy(cnt.astype(np.int32)).cuda()
    r = ogjk.ingest_hulls(pts, doff, dcnt)
    torch.cuda.synchronize()
    wall, ms = _event_steps(torch, lambda: ogjk.ingest_hulls(pts, doff, dcnt), args.steps,
                            args.warmup)
    # the hull kernel alone, for the roofline: points read once, faces written once
    hw, hms = _event_steps(torch, lambda: ogjk.build_hulls(pts, doff, dcnt), args.steps,
                           args.warmup)
    nf = r["nfaces"].cpu().numpy()
    alg = 12.0 * len(u) + 12.0 * float(nf.clip(min=0).sum()) + 16.0 * n
    achieved = alg / (np.mean(hms) * 1e-3) / 1e9
    line = {
        "metric": "mesh ingestion clouds/s (convex hull, centring, inertia factors)",
        "value": n / (wall / args.steps), "unit": "clouds/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"hull ingestion: {n} clouds of 8-32 points uniform on the unit "
                               "sphere", "clouds": n, "points": int(len(u)),
                   "faces": int(nf.clip(min=0).sum())},
        "hull_kernel_ms": float(np.mean(hms)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("hulls"),
                     "kernel": "k_hull (one hull build of every cloud)",
                     "kernel_ms_mean": float(np.mean(hms)), "algorithmic_bytes_per_launch": alg,
                     "algorithmic_bytes": "12 B per point read, 12 B per face written, 16 B "
                                          "per cloud of offsets and counts"},
    }
    if not args.no_cpu:
        cores = min(len(os.sched_getaffinity(0)), 16)
        t = time.perf_counter()
        cen_o, faces_o, fo_o, nf_o, k_o = oracle.ingest_batch(u, off, cnt, nthreads=cores)
        el = time.perf_counter() - t
        m = n
        faces = r["faces"].cpu().numpy()
        fo = r["face_off"].cpu().numpy()
        k = r["inertia_k"].cpu().numpy()
        cen = r["points"].cpu().numpy()
        same = [nf[i] == nf_o[i] and
                np.array_equal(faces[fo[i]:fo[i] + nf[i]], faces_o[fo_o[i]:fo_o[i] + nf_o[i]])
                for i in range(n)]
        same = np.array(same) & (k.view(np.uint32) == k_o.view(np.uint32)).all(1)
        cen_same = bool(np.array_equal(cen.view(np.uint32), cen_o.view(np.uint32)))
        line["cpu_baseline"] = {
            "value": m / el, "unit": "clouds/s", "cores": cores, "kind": "port",
            "sample": f"all {m} clouds, C oracle of the reference mesh builder (the same "
                      f"sequence), {cores} threads, {el:.2f}s"}
        line["parity"] = {"clouds_checked": m, "bit_exact_fraction": float(np.mean(same)),
                          "centred_points_identical": cen_same}
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sphere64", choices=sorted(WORKLOADS) + ["serving", "broadphase", "scene_step", "hulls"])
    ap.add_argument("--qps", type=int, default=1000, help="serving: request rate")
    ap.add_argument("--requests", type=int, default=2000, help="serving: requests per rate")
    ap.add_argument("--slots", type=int, default=4,
                    help="serving: launches in flight (at most the process's hardware queues, "
                         "4 by default: more slots share queues and block behind long launches)")
    ap.add_argument("--pairs", type=int, default=0, help="override the pair count")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-FP64 comparison")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 256 MiB elementwise copy (reads and writes 256 MiB) so a "
                         "profiler's FETCH_SIZE/WRITE_SIZE can be scaled against known bytes")
    ap.add_argument("--lanes", default=None, help="GJK,EPA lanes per pair (default per workload)")
    ap.add_argument("--cluster-min", type=int, default=None,
                    help="cull bodies with at least this many vertices (0: no clusters; "
                         "default per workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent and would shard by contiguous, cost-balanced ranges of the "
              "pair list, one process per GPU)",
              file=sys.stderr)
        return 2

    import torch
    import ogjk
    from ogjk import gen

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    if args.calibrate:
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        for _ in range(3):
            y = x * 1.0
        torch.cuda.synchronize()
        del x, y
    if args.workload == "serving":
        return serving(args)
    if args.workload == "broadphase":
        return broadphase(args)
    if args.workload == "scene_step":
        return scene_step(args)
    if args.workload == "hulls":
        return hulls(args)
    from ogjk import autotune
    spec = WORKLOADS[args.workload]
    lg, le = (int(x) for x in args.lanes.split(",")) if args.lanes else spec["lanes"]
    cmin = spec["cluster_min"] if args.cluster_min is None else (args.cluster_min or None)
    # a size-bucket table written by the autotuner, unless lanes are forced
    table = None
    tpath = os.path.join(ROOT, "opengjk-gpu_amd", "tuning", spec.get("tuning") or "-")
    if not args.lanes and os.path.isfile(tpath):
        table = autotune.load(tpath)
        if args.cluster_min is None:
            cmin = table["cluster_min"]
    ogjk.set_config(lg, le, False)
    wl = dict(spec["gen"])
    if args.pairs:
        wl["n_pairs"] = args.pairs
    t0 = time.perf_counter()
    batch = gen.make(nthreads=min(len(os.sched_getaffinity(0)), 16), **wl)
    print(f"# generated {batch.n_pairs} pairs, {batch.coords.shape[0]} vertices "
          f"in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    n = batch.n_pairs
    scene = ogjk.DeviceScene(batch, f64=not args.no_plain)
    build_s = None
    if cmin:
        t0 = time.perf_counter()
        scene.build_clusters(cmin)
        torch.cuda.synchronize()
        build_s = time.perf_counter() - t0
        print(f"# clusters built in {build_s:.2f}s", file=sys.stderr, flush=True)
    if table is not None:
        autotune.apply(table)
    out = scene.outputs(n, f64=True)
    stream = torch.cuda.current_stream()

    def step(o, plain=False, rec=False):
        scene.collide(3, n, out=o, plain_f64=plain, f64_record=rec)

    # verification record for the parity leg (untimed)
    step(out, rec=True)
    torch.cuda.synchronize()
    timed_out = scene.outputs(n, f64=False)

    def timed(plain: bool, steps=None):
        """Warm up, then time `steps` steps (None: as many as take about 4 s,
        at least 2 and at most --steps)."""
        for _ in range(args.warmup):
            step(timed_out, plain)
        torch.cuda.synchronize()
        if steps is None:
            t = time.perf_counter()
            step(timed_out, plain)
            torch.cuda.synchronize()
            one = time.perf_counter() - t
            steps = int(max(2, min(args.steps, 4.0 / max(one, 1e-6))))
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        t = time.perf_counter()
        for a, b in evs:
            a.record(stream)
            step(timed_out, plain)
            b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        ms = np.array([a.elapsed_time(b) for a, b in evs])
        return wall, ms, steps

    wall, ms, _ = timed(False, args.steps)
    print(f"# mixed: {wall / args.steps * 1e3:.3f} ms/step wall, kernel median {np.median(ms):.3f} ms",
          file=sys.stderr, flush=True)
    # the same step captured once as a HIP graph (counter reset, partition,
    # GJK per bucket, persistent EPA per class) and replayed: no host work per
    # step; its outputs must equal the eager step's bit for bit
    gout = scene.outputs(n, f64=False)
    graph, _ = scene.capture(3, n, out=gout)
    for _ in range(args.warmup):
        graph.launch()
    torch.cuda.synchronize()
    gev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t = time.perf_counter()
    for a, b in gev:
        a.record(stream)
        graph.launch()
        b.record(stream)
    torch.cuda.synchronize()
    g_wall = time.perf_counter() - t
    g_ms = np.array([a.elapsed_time(b) for a, b in gev])
    graph_info = {"ms_per_step": g_wall / args.steps * 1e3,
                  "pairs_per_s": n * args.steps / g_wall,
                  "step_ms_p50": float(np.percentile(g_ms, 50)),
                  "identical_to_eager": bool(all(torch.equal(gout[k], timed_out[k])
                                                 for k in ("simplices", "distances", "normals")))}
    graph.close()
    print(f"# graph replay: {graph_info['ms_per_step']:.3f} ms/step, identical to eager: "
          f"{graph_info['identical_to_eager']}", file=sys.stderr, flush=True)
    plain = pipe64 = None
    if not args.no_plain:
        # comparison legs run fewer steps when a step is long (about 4 s each)
        # the same split pipeline over FP64 vertex storage, no clusters
        # (isolates the FP32 sweep and the culling)
        wall_q, _, csteps = timed(True)
        pipe64 = {"pairs_per_s": n * csteps / wall_q, "ms_per_step": wall_q / csteps * 1e3,
                  "steps": csteps}
        print(f"# pipeline over fp64 storage: {pipe64['ms_per_step']:.3f} ms/step", file=sys.stderr,
              flush=True)
        # plain FP64 per-pair kernel: one group per pair runs GJK then EPA in one
        # non-persistent launch over FP64 SoA vertices, no worklist, no
        # clusters; the faster of 32 and 64 lanes per pair is the denominator
        plain = {"variants": {}}
        for el in (32, 64):
            ogjk.set_config(lg, el, True)
            wall_p, _, psteps = timed(True)
            plain["variants"][f"{el}_lanes"] = {"pairs_per_s": n * psteps / wall_p,
                                                "ms_per_step": wall_p / psteps * 1e3,
                                                "steps": psteps}
            print(f"# plain fp64 per-pair kernel, {el} lanes: {wall_p / psteps * 1e3:.3f} ms/step",
                  file=sys.stderr, flush=True)
        # the plain kernel's results equal the mixed pipeline's, pair by pair
        chk = scene.outputs(n, f64=True)
        scene.collide(3, n, out=chk, plain_f64=True, f64_record=True)
        torch.cuda.synchronize()
        plain["identical_to_mixed_fraction"] = float((chk["f64"] == out["f64"]).all(dim=1).float().mean())
        del chk
        ogjk.set_config(lg, le, False)
        if table is not None:
            autotune.apply(table)
        best = max(plain["variants"].values(), key=lambda v: v["pairs_per_s"])
        plain.update(best)

    value = n * args.steps / wall
    k_ms = float(np.mean(ms))
    base = parity = stream_pp = None
    if not args.no_cpu:
        base, parity, stream_pp = cpu_baseline(batch, out)
    if spec["roofline"] == "stream" and stream_pp is not None:
        algo, algo_kind = stream_pp * n, ("stream: 12 B x body vertices x the FP64 oracle's sweeps "
                                          "of that body, mean over the CPU sample x pairs")
    else:
        algo, algo_kind = compulsory_bytes(batch), "compulsory: every vertex once + records"
    achieved = algo / (k_ms * 1e-3) / 1e9
    comp = compulsory_bytes(batch) / (k_ms * 1e-3) / 1e9
    line = {
        "metric": "GJK+EPA pairs/s vs plain FP64 GPU path at 1e-5 rel err; p50/p99 query latency",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": spec["desc"], "pairs": n, "vertices": int(batch.coords.shape[0]),
                   "lanes_per_pair": ({"buckets_max_verts": table["max_verts"],
                                       "gjk": table["gjk_lanes"], "epa": table["epa_lanes"],
                                       "table": "opengjk-gpu_amd/tuning/" + spec["tuning"]}
                                      if table else {"gjk": lg, "epa": le}),
                   "clusters": {"min_verts": cmin, "build_s": build_s} if cmin else None,
                   "layout": "SoA fp32 scene; GJK kernel -> device worklist -> persistent EPA kernel"},
        "step_ms_p50": float(np.percentile(ms, 50)), "step_ms_p99": float(np.percentile(ms, 99)),
        "graph": graph_info,
        "plain_f64": plain,
        "speedup_vs_plain_f64": (value / plain["pairs_per_s"]) if plain else None,
        "pipeline_f64_storage": pipe64,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload),
                     "kernel": "k_gjk + k_epa_list (one step)",
                     "kernel_ms_mean": k_ms, "algorithmic_bytes_per_launch": algo,
                     "algorithmic_bytes": algo_kind,
                     # every vertex read once: a floor on HBM traffic, so this
                     # fraction stays <= 1 where culling beats the stream figure
                     "achieved_compulsory": comp, "frac_compulsory": comp / HBM_PEAK_GBS},
    }
    if base is not None:
        line["cpu_baseline"] = base
        line["parity"] = parity
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
