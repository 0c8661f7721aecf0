"""Throughput benchmark of the ray-casting simulate() path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 1024] [--sensors 1000]
                    [--cpu-seconds 12] [--no-cpu]

Workload (default): fBm terrain, 1024 x 1024 grid at 0.1 m (2,093,058 triangles),
1,000 Ouster-style spherical sensors (128 x 1024 rays, +-22.5 deg, 0.1-100 m) 1.5 m above
the ground with random yaw, all seven attributes (hits, ranges, points, normals, face,
geom and object ids): 131,072,000 rays per step.  A step is one simulate() call over all
poses; inputs (map, sensor, poses, output buffers) are resident in HBM before timing.

Prints one JSON line: Mrays/s, the roofline of the traversal kernel (41 algorithmic bytes
per ray written to HBM, against 8 TB/s), the CPU baseline (the CPU oracle on the host
cores, on a bounded sample of the same sensors) and the accuracy of the sampled rays
against it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the default workload's traversal kernel
TRAFFIC_PROFILE = os.path.join("profiles", "k_simulate_terrain_1000x128x1024.md")


def profiled(default_workload: bool, key: str, cast, path: str = TRAFFIC_PROFILE):
    """A figure measured with rocprofv3 for a default workload, recorded in its profile
    summary on a `key: value` line (`traffic_bytes_per_launch`: HBM-side bytes per launch;
    `l2_hit_rate`: TCC hits / (hits + misses)), or None."""
    if not default_workload:
        return None
    try:
        with open(os.path.join(ROOT, path)) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return cast(line.split(":", 1)[1].strip())
    except (OSError, ValueError):
        pass
    return None


def roofline(kernel: str, alg_bytes: float, kernel_ms: float, profile: str | None = None, note: str = "") -> dict:
    """HBM roofline object of one kernel: algorithmic bytes per launch over its measured
    launch time against 8 TB/s, and the HBM-side traffic per launch of its profile summary."""
    gbs = alg_bytes / (kernel_ms * 1e-3) / 1e9
    traffic = profiled(profile is not None, "traffic_bytes_per_launch", int, profile) if profile else None
    out = {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": kernel,
           "kernel_avg_ms": round(kernel_ms, 4), "alg_bytes_per_launch": int(alg_bytes)}
    if traffic:
        out["traffic_source"] = profile
        out["traffic_gbs"] = round(traffic / (kernel_ms * 1e-3) / 1e9, 1)
        out["l2_hit_rate"] = profiled(True, "l2_hit_rate", float, profile)
    if note:
        out["note"] = note
    return out


BATCH_PROFILE = os.path.join("profiles", "k_batch_objects_2000_mixed.md")
ICP_PROFILE = os.path.join("profiles", "k_simulate_o1dn_objects_2000x32768.md")


BYTES_PER_RAY_ALL = 1 + 4 + 12 + 12 + 4 + 4 + 4  # hits, ranges, points, normals, 3 ids


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--sensors", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched / pose-gradient / ICP workloads")
    ap.add_argument("--extra-only", choices=["batched", "pose_gradient", "icp", "next_rows", "math_layer"],
                    help="only run one secondary workload with default launch parameters (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (scale factor for HBM counters)")
    ap.add_argument("--visit-stats", type=int, metavar="POSES", default=0,
                    help="only count node visits, leaves and triangle tests per ray of the default workload's "
                         "kernel over its first POSES sensor poses (a counting build of the same traversal)")
    ap.add_argument("--compare-layouts", nargs="*", metavar="WORKLOAD",
                    help="only time simulate() on the binary and the 4-wide quantised BVH of the same map "
                         "(workloads: terrain particles icp room; default all) and check the outputs are identical")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(V, F, model, Tbm, seconds, gpu_numpy_for_pose, flag_poses: int = 100):
    """Times the oracle on the host cores, one sensor pose at a time, for about `seconds`,
    and compares the GPU's results of the poses it casts (first pass) with the oracle's."""
    import oracle
    from oracle import diff as D

    t0 = time.time()
    om = oracle.OracleMap(V, F)
    log(f"cpu: oracle BVH built in {time.time() - t0:.1f} s")
    threads = os.cpu_count() or 1
    rays = 0
    elapsed = 0.0
    poses = 0
    per_pose = []
    acc = None
    reflagged = 0
    oracle.simulate(om, model, Tbm[:1], threads=threads, flags=False)  # warm-up, not timed
    # the reference's benchmark loop (velodyne_benchmark.hpp:92-114): run for at least
    # `seconds`, report the running mean; poses are cast in order and, once all are done,
    # again from the first (the accuracy check covers the first pass)
    while elapsed < seconds:
        p = poses % len(Tbm)
        t = time.perf_counter()
        o = oracle.simulate(om, model, Tbm[p:p + 1], threads=threads, flags=False)
        dt = time.perf_counter() - t
        elapsed += dt
        per_pose.append(dt)
        if poses < len(Tbm):
            g = gpu_numpy_for_pose(p)
            r = D.compare(g, {**o, "flags": np.zeros(len(o["ranges"]), np.uint8)}, model.range.max)
            if r["mismatching_rays"]:
                # classify the differences: the same pose again with the ambiguity flags (untimed)
                r = D.compare(g, oracle.simulate(om, model, Tbm[p:p + 1], threads=threads), model.range.max)
                reflagged += 1
            acc = D.merge(acc, {**r, "flagged": {}})
        rays += len(o["ranges"])
        poses += 1
    log(f"cpu: {poses} pose casts, {rays} rays in {elapsed:.2f} s on {threads} threads")
    # the oracle's ambiguity flags over the first poses (untimed): rays within 1e-6 of an edge,
    # near misses, near ties and range limits, where an FP32 intersector may decide otherwise
    k = min(flag_poses, len(Tbm))
    fo = oracle.simulate(om, model, Tbm[:k], threads=threads, attrs=("hits",))["flags"]
    flagged = {"poses": k, "rays": int(len(fo)), "any": int((fo != 0).sum()),
               "edge": int((fo & oracle.FLAG_EDGE != 0).sum()), "near_miss": int((fo & oracle.FLAG_NEAR_MISS != 0).sum()),
               "tie": int((fo & oracle.FLAG_TIE != 0).sum()), "range_limit": int((fo & oracle.FLAG_RANGE != 0).sum())}
    med = float(np.median(per_pose))
    base = {"value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "median": model.size() / med / 1e6, "cpu": cpu_model(),
            "sample": f"{poses} casts of the {len(Tbm)} sensor poses in order ({rays} rays, {elapsed:.1f} s), all "
                      f"attributes, oracle (scalar C++ port of the Embree simulate loop: exact FP64 ray/triangle "
                      f"predicate, FP64 box tests, binned-SAH binary BVH; no SIMD packets) on {threads} host threads "
                      f"after one untimed pose, BVH build excluded; value = rays / total time, median = per-pose "
                      f"median; accuracy over the first {min(poses, len(Tbm))} poses"}
    accuracy = {"rays": acc["rays"], "poses": min(poses, len(Tbm)), "mismatch": acc["mismatch"],
                "mismatching_rays": acc["mismatching_rays"], "mismatching_unflagged": acc["mismatching_unflagged"],
                "poses_reflagged": reflagged,
                "range_max_err": acc["diff"]["ranges"]["max_diff"], "point_max_err": acc["diff"]["points"]["max_diff"],
                "normal_max_err": acc["diff"]["normals"]["max_diff"], "diff": acc["diff"],
                "ambiguous": flagged["any"], "flagged": flagged,
                "note": "diff: apps/rmagine_simulation_diff_calc/Main.cpp:83-147 per attribute (a = GPU, b = oracle); "
                        "ambiguous: rays the oracle flags over the first poses"}
    return base, accuracy


def measure_hbm_gbs(nbytes: int = 2 << 30, reps: int = 5) -> float:
    """HBM bandwidth of this GPU as a plain device copy achieves it: bytes read + written
    per second of the best of `reps` timed copies of `nbytes`."""
    import torch
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    src.fill_(1)
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = None
    for _ in range(reps):
        a, b = _events()
        a.record()
        dst.copy_(src)
        b.record()
        b.synchronize()
        ms = a.elapsed_time(b)
        best = ms if best is None else min(best, ms)
    del src, dst
    torch.cuda.empty_cache()
    return 2 * nbytes / (best * 1e-3) / 1e9


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def extra_batched(cpu_seconds: float, do_cpu: bool, tune: bool = True) -> dict:
    """Heterogeneous batch: 2,000 sensors (40 % VLP-16, 30 % pinhole 640x480, 30 % O1Dn with
    1k-50k rays) on a 500k-triangle scene, hits + ranges, one persistent launch, tuned; the
    split into two launches (image sensors, then ray lists) is timed beside it."""
    import torch

    from rmagine_amd import autotune, scenes
    from rmagine_amd.simulation import BatchSimulatorHip, HipMap, simulator_for

    V, F, sensors = scenes.mixed_sensor_workload()
    hm = HipMap(V, F)
    entries = [{"model": m, "Tbm": torch.as_tensor(P, device="cuda"), "attrs": ("hits", "ranges")} for m, P in sensors]
    batch = BatchSimulatorHip(hm, entries)
    rays = batch.total_rays
    mix = {"vlp16": 0, "pinhole": 0, "o1dn": 0}
    for m, _ in sensors:
        mix[{0: "vlp16", 1: "pinhole", 2: "o1dn"}[m.kind]] += 1
    key = autotune.cache_key(len(F), mix)
    if tune:
        tuned = autotune.tune_batch(batch, key=key, chunks=(64, 256, 1024), waves=(16, 32), tile_rows=(2, 4, 8),
                                    reps=2, use_cache=False, layouts=(False, True))
    else:
        tuned = {"chunk_rays": 256, "waves_per_cu": 32, "tile_rows": 4, "lds_stack": 10, "lds_stack_ms": {},
                 "rays_sorted": True, "ray_order_ms": None, "wide": False, "layout_ms": None}
    ms = autotune.time_batch(batch, reps=5)
    batch.set_params(launches=2)
    split_ms = autotune.time_batch(batch, reps=5)
    batch.set_params(launches=1)
    batch.simulate()  # the parity check below reads the one-launch results
    layout = "4-wide" if tuned.get("wide") else "binary"
    log(f"batched: {rays} rays, {ms:.2f} ms -> {rays / ms / 1e3:.1f} Mrays/s in one persistent launch "
        f"(two launches, images then ray lists: {split_ms:.2f} ms) (tuned {tuned['chunk_rays']}/"
        f"{tuned['waves_per_cu']}/{tuned['tile_rows']}, LDS stack {tuned['lds_stack']} {tuned['lds_stack_ms']}, "
        f"rays {'sorted' if tuned['rays_sorted'] else 'list order'} {tuned['ray_order_ms']}, {layout} nodes; "
        f"layouts {tuned.get('layout_ms')})")
    # the same sensors with one rmh_simulate launch each
    sims = [simulator_for(m, hm) for m, _ in sensors]
    bundles = [None] * len(sims)
    for i, (e, sim) in enumerate(zip(entries, sims)):
        bundles[i] = sim.simulate(e["Tbm"], attrs=("hits", "ranges"))
    torch.cuda.synchronize()
    a, b = _events()
    a.record()
    for i, (e, sim) in enumerate(zip(entries, sims)):
        sim.simulate(e["Tbm"], bundles[i])
    b.record()
    b.synchronize()
    plain_ms = a.elapsed_time(b)
    # parity of the batched results with the per-sensor launches (both pinned to the oracle)
    mism = 0
    for bb, pb in zip(batch.bundles, bundles):
        mism += int((bb.ranges != pb.ranges).sum().item()) + int((bb.hits != pb.hits).sum().item())
    o1dn_rays = sum(m.size() for m, _ in sensors if m.kind == 2)
    alg = 5 * rays + 12 * o1dn_rays + 32 * len(sensors)
    out = {"workload": f"{len(F)} tris (random boxes + icospheres, 200 m cube), {len(sensors)} sensors "
                       f"(800 VLP-16, 600 pinhole 640x480 fov 90, 600 O1Dn 1k-50k rays), hits+ranges",
           "roofline": roofline("k_batch (one persistent launch, tuned)", alg, ms, BATCH_PROFILE,
                                "algorithmic bytes: hits + ranges (5 B/ray), O1Dn directions (12 B/ray), poses"),
           "rays": rays, "launch": "one persistent launch", "batched_ms": round(ms, 3),
           "batched_mrays_s": round(rays / ms / 1e3, 1), "two_launches_ms": round(split_ms, 3),
           "per_sensor_launches_ms": round(plain_ms, 3), "per_sensor_mrays_s": round(rays / plain_ms / 1e3, 1),
           "tuned": {**{k: tuned[k] for k in ("chunk_rays", "waves_per_cu", "tile_rows", "lds_stack")},
                     "lds_stack_ms": {str(k): round(v, 3) for k, v in tuned["lds_stack_ms"].items()},
                     "ray_order": "sorted" if tuned["rays_sorted"] else "list", "ray_order_ms": tuned["ray_order_ms"],
                     "layout": "4-wide" if tuned.get("wide") else "binary", "layout_ms": tuned.get("layout_ms")},
           "batched_vs_per_sensor_mismatches": mism}
    if do_cpu:
        import oracle
        om = oracle.OracleMap(V, F)
        rng = np.random.default_rng(0)
        idx = rng.permutation(len(sensors))
        crays, cel = 0, 0.0
        for i in idx:
            m, P = sensors[i]
            t = time.perf_counter()
            oracle.simulate(om, m, P, attrs=("hits", "ranges"), threads=os.cpu_count() or 1, flags=False)
            cel += time.perf_counter() - t
            crays += m.size()
            if cel > cpu_seconds:
                break
        cpu = crays / cel / 1e6
        out["cpu_oracle_mrays_s"] = round(cpu, 2)
        out["cpu_sample"] = f"{crays} rays of randomly drawn sensors of this batch, {os.cpu_count()} threads"
        out["batched_over_cpu"] = round(out["batched_mrays_s"] / cpu, 1)
    return out


def extra_pose_grad() -> dict:
    """10k VLP-16 particles on the ~50k-triangle room: loss + 6-DoF gradient per particle,
    one batched launch vs one launch per particle."""
    import torch

    from rmagine_amd import adjoint as A
    from rmagine_amd import scenes
    from rmagine_amd.simulation import HipMap, simulator_for

    V, F, T_true, P = scenes.fine_room_workload()
    hm = HipMap(V, F)
    sim = simulator_for(scenes.vlp16_workload_model(), hm)
    res = sim.simulate(torch.as_tensor(T_true, device="cuda"), attrs=("hits", "ranges"))
    torch.cuda.synchronize()
    rng = np.random.default_rng(3)
    target = (res.ranges + torch.as_tensor(rng.normal(scale=0.01, size=res.ranges.shape[0]).astype(np.float32),
                                           device="cuda")).contiguous()
    mask = res.hits.contiguous()
    Pd = torch.as_tensor(P, device="cuda")
    n = Pd.shape[0]
    out = torch.empty((n, 8), dtype=torch.float64, device="cuda")
    ws = torch.empty(A.workspace_bytes(sim, n), dtype=torch.uint8, device="cuda")
    for _ in range(3):  # untimed warm-up, as for the headline workload
        A.pose_grad(sim, Pd, target, mask, out=out, workspace=ws)
    torch.cuda.synchronize()
    a, b = _events()
    reps = 5
    a.record()
    for _ in range(reps):
        A.pose_grad(sim, Pd, target, mask, out=out, workspace=ws)
    b.record()
    b.synchronize()
    batched = a.elapsed_time(b) / reps
    outs = [torch.empty((1, 8), dtype=torch.float64, device="cuda") for _ in range(n)]
    ws1 = torch.empty(A.workspace_bytes(sim, 1), dtype=torch.uint8, device="cuda")
    for i in range(min(n, 100)):
        A.pose_grad(sim, Pd[i:i + 1], target, mask, out=outs[i], workspace=ws1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    a.record()
    for i in range(n):
        A.pose_grad(sim, Pd[i:i + 1], target, mask, out=outs[i], workspace=ws1)
    b.record()
    b.synchronize()
    plain = (time.perf_counter() - t) * 1e3
    same = bool(torch.equal(torch.cat(outs), out))
    alg = 5 * 14400 + 32 * n + 64 * n
    return {"workload": f"{len(F)} tris room, {n} VLP-16 particles (sigma 0.2 m, 3 deg), Huber(0.05) on "
                        "true-pose ranges + N(0, 0.01 m)",
            "roofline": roofline("k_pose_grad + k_pose_grad_reduce (one evaluation)", alg, batched, None,
                                 "algorithmic bytes: the shared target scan (4 B range + 1 B mask per ray, read once), "
                                 "poses in, 8 doubles per pose out; the kernel is bound by traversal latency"),
            "rays_per_eval": n * 14400, "batched_ms_per_eval": round(batched, 3),
            "batched_evals_s": round(1e3 / batched, 2), "plain_ms_per_eval": round(plain, 3),
            "plain_evals_s": round(1e3 / plain, 3), "batched_over_plain": round(plain / batched, 1),
            "batched_equals_plain": same}


def extra_icp() -> dict:
    """2,000 pose hypotheses x 32,768-ray O1Dn scans on the 500k-triangle scene: 10 RCC
    iterations (simulate, 5 x (p2l statistics FP64, Umeyama FP64, pose update), pose update)
    recorded once into a HIP graph and replayed."""
    import torch

    from rmagine_amd import icp, scenes
    from rmagine_amd import types as T
    from rmagine_amd.simulation import HipMap, simulator_for

    V, F = scenes.objects_mesh()
    hm = HipMap(V, F)
    rng = np.random.default_rng(4)
    model = T.O1DnModel(32768, 1, T.Interval(0.1, 100.0), (0.0, 0.0, 0.0), scenes.random_unit_dirs(32768, rng))
    sim = simulator_for(model, hm)
    T_true = T.make_transforms([[0.0, 0.0, 0.0]], [[0.0, 0.0, 0.0, 1.0]])
    data = sim.simulate(torch.as_tensor(T_true, device="cuda"), attrs=("hits", "points"))
    n = 2000
    hyp = np.repeat(T_true, n, 0).copy()
    hyp[:, 4:7] += rng.normal(scale=0.2, size=(n, 3)).astype(np.float32)
    for i in range(n):
        a = 0.5 * rng.normal(scale=np.deg2rad(2.0))
        hyp[i, 0:4] = [0.0, 0.0, np.sin(a), np.cos(a)]
    # ray-cast correspondences between different poses land on different objects of this
    # sparse scene unless they are gated: 0.5 m point-to-plane gate
    rcc = icp.BatchedRCC(hm, model, data.points, data.hits, torch.as_tensor(hyp, device="cuda"), max_dist=0.5)
    s = torch.cuda.Stream()
    graph = rcc.capture(10, s)
    torch.cuda.synchronize()
    a, b = _events()
    a.record(s)
    graph.launch(s)
    b.record(s)
    b.synchronize()
    ms = a.elapsed_time(b)
    err0 = np.linalg.norm(hyp[:, 4:7], axis=1)
    err1 = np.linalg.norm(rcc.Tbm.cpu().numpy()[:, 4:7], axis=1)
    grad_norm = float(np.median(np.linalg.norm(rcc.gradient().cpu().numpy(), axis=1)))
    # the iteration's kernels timed one by one (eager, same buffers): the simulate and one
    # point-to-plane statistics pass
    sim_ms, st_ms = [], []
    for _ in range(3):
        a, b = _events()
        a.record()
        rcc.sim.simulate(rcc.Tbm, rcc.model_scan)
        b.record()
        b.synchronize()
        sim_ms.append(a.elapsed_time(b))
        a, b = _events()
        a.record()
        rcc._statistics(None)
        b.record()
        b.synchronize()
        st_ms.append(a.elapsed_time(b))
    sim_ms, st_ms = min(sim_ms), min(st_ms)
    rays_it = n * 32768
    # FP64 Umeyama of the last statistics against the FP64 CPU oracle (LAPACK SVD)
    from oracle import stats as S
    from rmagine_amd import math as M
    out64 = torch.empty((n, 7), dtype=torch.float64, device="cuda")
    M.umeyama_transform(rcc.stats, out64=out64)
    st = M.stats_to_numpy(rcc.stats)
    g64 = out64.cpu().numpy()
    umerr = 0.0
    for i in range(0, n, 10):
        ref = S.umeyama({k: st[k][i] for k in ("dataset_mean", "model_mean", "covariance", "n_meas")})
        sgn = 1.0 if np.dot(ref[:4], g64[i, :4]) >= 0 else -1.0
        umerr = max(umerr, float(np.abs(np.r_[g64[i, :4] * sgn, g64[i, 4:]] - ref).max()))
    return {"workload": f"{len(F)} tris, {n} hypotheses x 32768 O1Dn rays, 10 x (simulate + 5 x (p2l FP64 "
                        "stats + FP64 Umeyama + update)), one HIP graph",
            "graph_ms_10_iterations": round(ms, 3), "ms_per_iteration": round(ms / 10, 3),
            "roofline": roofline("k_simulate<O1DN> (hits, points, normals)", 25 * rays_it + 12 * 32768 + 32 * n, sim_ms,
                                 ICP_PROFILE, "algorithmic bytes: hits + points + normals written (25 B/ray), the "
                                              "shared direction list, poses"),
            "roofline_statistics": roofline("k_stats<uint32_t> (p2l, FP64)", 25 * rays_it + 13 * 32768 + 160 * n, st_ms,
                                            None, "algorithmic bytes: model points + normals + mask (25 B per "
                                                  "correspondence), the shared dataset scan, pre-transforms and "
                                                  "statistics"),
            "simulate_ms": round(sim_ms, 3), "statistics_ms": round(st_ms, 4),
            "gauss_newton_gradient_norm_median": grad_norm,
            "simulated_mrays_s": round(10 * n * 32768 / ms / 1e3, 1),
            "median_translation_error_m": {"before": round(float(np.median(err0)), 4),
                                           "after": round(float(np.median(err1)), 6)},
            "umeyama_fp64_max_abs_err_vs_cpu_oracle": umerr}


def _timed_ms(fn, reps: int = 5) -> float:
    """Mean launch time of fn() (one untimed call first), HIP events on the current stream."""
    import torch
    fn()
    torch.cuda.synchronize()
    a, b = _events()
    ms = []
    for _ in range(reps):
        a.record()
        fn()
        b.record()
        b.synchronize()
        ms.append(a.elapsed_time(b))
    return float(np.mean(ms))


def ouster_ondn_model():
    """128 x 1024 Ouster-style ray list with per-ray origins: the directions of the default
    spherical sensor, each beam starting 15.8 mm from the sensor centre along its azimuth (the
    beam-origin offset of Ouster sensor metadata)."""
    from rmagine_amd import types as T
    rows, cols = 128, 1024
    phi = (np.float32(-22.5 * T.DEG_TO_RAD) + np.arange(rows, dtype=np.float64) * np.float64(T.f32(45.0 / 127.0 * T.DEG_TO_RAD)))
    th = (np.float32(-np.pi) + np.arange(cols, dtype=np.float64) * np.float64(T.f32(2.0 * np.pi / 1024.0)))
    P, Th = np.meshgrid(phi, th, indexing="ij")
    dirs = np.stack([np.cos(P) * np.cos(Th), np.cos(P) * np.sin(Th), np.sin(P)], -1).reshape(-1, 3).astype(np.float32)
    origs = (0.0158 * np.stack([np.cos(Th), np.sin(Th), np.zeros_like(Th)], -1)).reshape(-1, 3).astype(np.float32)
    return T.OnDnModel(cols, rows, T.Interval(0.1, 100.0), origs, dirs)


def extra_next_rows(do_cpu: bool, cpu_seconds: float = 3.0) -> dict:
    """The parts either side of the simulate path, each timed with HIP events (mean of 5
    launches after an untimed one) on the terrain of the default workload (1024^2 grid at
    0.1 m, 2,093,058 triangles, 1,000 sensor poses):
    * range noise (rmh_noise_apply, Gaussian / relative Gaussian / uniform dust) over
      131,072,000 ranges, the size of one default step's output: 8 algorithmic bytes per range
      (read and write), against 8 TB/s;
    * OnDn simulation (per-ray origins): 1,000 poses x the 128 x 1024 Ouster-style ray list,
      all seven attributes: 41 B written + 24 B of direction and origin read per ray; timed
      with the sensor's defaults (for a grid of this size: its 4 x 16 tiles, classic walk)
      and after autotune.tune_sensor timed every ray order and walk;
    * closest-point queries (rmh_closest_points): 4,194,304 points within 2 m of the surface,
      unbounded search: 12 B read + 40 B written per query; in random order (which the call
      sorts by the Morton code of their cells and walks in that order) and again sorted by
      the caller by the Morton code of their 0.5 m cells;
    * BVH refit (rmh_map_refit) after moving every vertex, synchronous, one launch per tree
      depth (the schedule made at the first, untimed refit): per triangle 48 B of triangle and
      16 B of normal written, 36 B of vertices read; per node 64 B written;
    * map file loading on the host (rmh_scene_load): the terrain as binary PLY.
    With do_cpu, the oracle checks samples of the GPU results (OnDn poses, closest-point
    queries, noise prefixes) and the OnDn poses it casts give its throughput."""
    import tempfile

    import torch

    from rmagine_amd import noise as NZ
    from rmagine_amd import scenes
    from rmagine_amd.simulation import HipMap, OnDnSimulatorHip, load_map_file

    dev = torch.device("cuda:0")
    V, F, Tbm = scenes.terrain_workload(n_grid=1024, n_sensors=1000)
    hm = HipMap(V, F, device=dev)
    info = hm.info()
    out = {"workload": f"terrain 1024x1024 fBm at 0.1 m ({len(F)} tris, {info['n_nodes']} nodes), "
                       f"{len(Tbm)} sensor poses"}

    # range noise
    n = 131_072_000
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    ranges = torch.rand(n, generator=gen, device=dev) * 99.9 + 0.1
    kinds = (("gaussian", NZ.GaussianNoiseHip, (0.0, 0.01), 0), ("rel_gaussian", NZ.RelGaussianNoiseHip, (0.0, 0.002, 1.0), 1),
             ("uniform_dust", NZ.UniformDustNoiseHip, (0.001, 0.5), 2))
    noise = {"elements": n}
    for name, cls, prm, kind in kinds:
        nz = cls(*prm)
        ms = _timed_ms(lambda: nz.apply(ranges))
        noise[name] = roofline(f"k_noise<{name}>", 8 * n, ms, note="algorithmic bytes: one range read and written")
        noise[name]["gelem_s"] = round(n / ms / 1e6, 2)
        if do_cpu:
            from oracle import noise as ON
            k = 1 << 20
            x = ranges[:k].clone()
            host = x.cpu().numpy()
            cls(*prm).apply(x)  # call count 0
            t = time.perf_counter()
            want = ON.apply(host, kind, *(list(prm) + [0.0] * (3 - len(prm))), max_range=10000.0, seed=42, epoch=0)
            cel = time.perf_counter() - t
            got = x.cpu().numpy()
            if kind == 2:
                bad = int((got != want).sum())
            else:
                bad = int((np.abs(got.astype(np.float64) - want) > np.spacing(np.abs(want))).sum())
            noise[name]["cpu_oracle_gelem_s"] = round(k / cel / 1e9, 4)
            noise[name]["oracle_check"] = {"elements": k, "outside_1_ulp" if kind != 2 else "differing": bad}
    del ranges
    torch.cuda.empty_cache()
    out["range_noise"] = noise

    # OnDn simulation
    model = ouster_ondn_model()
    sim = OnDnSimulatorHip(hm, device=dev)
    sim.setModel(model)
    Tbm_d = torch.as_tensor(Tbm, device=dev)
    res = sim.simulate(Tbm_d)
    rays = model.size() * len(Tbm)
    ms_default = _timed_ms(lambda: sim.simulate(Tbm_d, res))
    from rmagine_amd import autotune
    tuned = autotune.tune_sensor(sim, Tbm_d, res)
    ms = _timed_ms(lambda: sim.simulate(Tbm_d, res))
    ondn = {"rays": rays, "ms": round(ms, 3), "mrays_s": round(rays / ms / 1e3, 1),
            "walk": "speculative" if tuned["speculative"] else "classic",
            "ray_order": tuned["ray_order"],
            "untuned_ms": round(ms_default, 3), "tuner_trials_ms": tuned["trials"],
            "roofline": roofline("k_simulate<OnDn> (traversal + attribute writers)", (BYTES_PER_RAY_ALL + 24) * rays, ms,
                                 note="algorithmic bytes: 41 B of outputs written, 24 B of direction and origin read per ray")}
    if do_cpu:
        import oracle
        from oracle import diff as D
        om = oracle.OracleMap(V, F)
        per = model.size()
        acc, crays, cel, p = None, 0, 0.0, 0
        while cel < cpu_seconds and p < len(Tbm):
            t = time.perf_counter()
            o = oracle.simulate(om, model, Tbm[p:p + 1], threads=os.cpu_count() or 1, flags=False)
            cel += time.perf_counter() - t
            sl = slice(p * per, (p + 1) * per)
            g = {"hits": res.hits[sl].cpu().numpy(), "ranges": res.ranges[sl].cpu().numpy(),
                 "points": res.points[sl].cpu().numpy(), "normals": res.normals[sl].cpu().numpy(),
                 "face_ids": res.face_ids[sl].cpu().numpy().view(np.uint32),
                 "geom_ids": res.geom_ids[sl].cpu().numpy().view(np.uint32),
                 "object_ids": res.object_ids[sl].cpu().numpy().view(np.uint32)}
            r = D.compare(g, {**o, "flags": np.zeros(len(o["ranges"]), np.uint8)}, model.range.max)
            acc = D.merge(acc, {**r, "flagged": {}})
            crays += per
            p += 1
        ondn["cpu_oracle_mrays_s"] = round(crays / cel / 1e6, 2)
        ondn["oracle_check"] = {"poses": p, "rays": crays, "mismatching_rays": acc["mismatching_rays"],
                                "range_max_err": acc["diff"]["ranges"]["max_diff"]}
        del om
    del res
    torch.cuda.empty_cache()
    out["ondn_simulate"] = ondn

    # closest points
    rng = np.random.default_rng(11)
    nq = 1 << 22
    lo, hi = V.min(0), V.max(0)
    q = rng.uniform(lo, hi, size=(nq, 3)).astype(np.float32)
    q[:, 2] = scenes.terrain_height_at(scenes.terrain_heights(1024, seed=scenes.SEED_BASE + 1), 0.1,
                                       q[:, 0], q[:, 1]) + rng.uniform(-2.0, 2.0, nq).astype(np.float32)
    qd = torch.as_tensor(q, device=dev)
    cp = hm.closest_points(qd)
    ms = _timed_ms(lambda: hm.closest_points(qd))
    closest = {"queries": nq, "ms": round(ms, 3), "mqueries_s": round(nq / ms / 1e3, 1),
               "roofline": roofline("k_closest", 52 * nq, ms, note="algorithmic bytes: 12 B query read, 40 B written")}
    # the same queries in Morton order of their cells (sorted by the caller, outside the timed
    # region): neighbouring lanes then walk neighbouring parts of the tree, as the points of
    # one scan do; results are the same per query
    cell = np.floor((q - lo) / 0.5).astype(np.int64)
    key = np.zeros(nq, np.int64)
    for b in range(16):
        for k in range(3):
            key |= ((cell[:, k] >> b) & 1) << (3 * b + k)
    perm = np.argsort(key, kind="stable")
    qs = torch.as_tensor(np.ascontiguousarray(q[perm]), device=dev)
    cps = hm.closest_points(qs)
    ms_s = _timed_ms(lambda: hm.closest_points(qs))
    pd = torch.as_tensor(perm, device=dev)
    closest["morton_sorted_queries"] = {
        "ms": round(ms_s, 3), "mqueries_s": round(nq / ms_s / 1e3, 1),
        "results_equal": bool(torch.equal(cps["dist"], cp["dist"][pd]) and torch.equal(cps["face_ids"], cp["face_ids"][pd]))}
    del qs, cps, pd
    if do_cpu:
        from oracle import closest as OC
        k = 2
        t = time.perf_counter()
        d, _, f = OC.closest_points(V, F, q[:k])
        cel = time.perf_counter() - t
        gd = cp["dist"][:k].cpu().numpy()
        gf = cp["face_ids"][:k].cpu().numpy().astype(np.int64)
        closest["oracle_check"] = {"queries": k, "dist_max_rel_err": float(np.max(np.abs(gd - d) / np.maximum(d, 1e-30))),
                                   "face_ids_equal": int((gf == f).sum()), "oracle_s_per_query": round(cel / k, 3),
                                   "note": "oracle: brute force over every triangle (a check, not a baseline)"}
    del cp, qd
    out["closest_points"] = closest

    # refit
    Vm = V.copy()
    Vm[:, 2] += (0.05 * np.sin(V[:, 0] * 0.7) * np.cos(V[:, 1] * 0.3)).astype(np.float32)
    Vs = [torch.as_tensor(Vm, device=dev), torch.as_tensor(V, device=dev)]
    it = [0]

    def refit():
        hm.refit(Vs[it[0] & 1])
        it[0] += 1
    ms = _timed_ms(refit, reps=6)  # ends on the moved vertices
    alg = (48 + 16 + 36) * len(F) + 64 * info["n_nodes"]
    refit_out = {"ms": round(ms, 3), "roofline": roofline("k_refit_tris + k_refit_level per depth (rmh_map_refit)", alg, ms,
                                                          note="algorithmic bytes: triangles, normals and vertices of "
                                                               "every face, 64 B per node")}
    # the refitted map casts exactly what a map built from the moved vertices casts
    fresh = HipMap(Vm, F, device=dev)
    sims = []
    for m_ in (hm, fresh):
        s_ = OnDnSimulatorHip(m_, device=dev)
        s_.setModel(model)
        sims.append(s_.simulate(Tbm_d[:16], attrs=("hits", "ranges", "face_ids")))
    refit_out["vs_rebuilt_map_mismatches"] = int(sum(int((getattr(sims[0], k) != getattr(sims[1], k)).sum().item())
                                                     for k in ("hits", "ranges", "face_ids")))
    refit_out["rays_compared"] = 16 * model.size()
    del fresh, sims
    out["refit"] = refit_out

    # map file loading (host)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "terrain.ply")
        head = (f"ply\nformat binary_little_endian 1.0\nelement vertex {len(V)}\nproperty float x\nproperty float y\n"
                f"property float z\nelement face {len(F)}\nproperty list uchar int vertex_indices\nend_header\n")
        fb = np.empty(len(F), dtype=[("n", "u1"), ("i", "<i4", 3)])
        fb["n"] = 3
        fb["i"] = F.astype(np.int32)
        with open(path, "wb") as fh:
            fh.write(head.encode())
            fh.write(np.ascontiguousarray(V, "<f4").tobytes())
            fh.write(fb.tobytes())
        size = os.path.getsize(path)
        t = time.perf_counter()
        s_ = load_map_file(path)
        sec = time.perf_counter() - t
        same = bool(np.array_equal(s_["vertices"], V) and np.array_equal(s_["faces"], F.astype(np.uint32)))
    out["map_loading"] = {"format": "binary PLY", "bytes": size, "s": round(sec, 3), "mb_s": round(size / sec / 1e6, 1),
                          "mtris_s": round(len(F) / sec / 1e6, 2), "arrays_equal_to_the_mesh": same}
    return out


def extra_math_layer(do_cpu: bool) -> dict:
    """The batched pose operations and moments of the math layer at sizes that fill the GPU,
    each timed with HIP events (mean of 5 launches after an untimed one), with its algorithmic
    bytes against 8 TB/s:
    * rmh_transform_mul (multNxN), 16,777,216 transform pairs: 64 B read, 32 B written each;
    * rmh_transform_inv, 16,777,216 transforms: 32 B read, 32 B written;
    * rmh_transform_points, one transform applied to 67,108,864 points (mult1xN): 12 B read,
      12 B written per point;
    * rmh_transform_pack_quat, 16,777,216 (quaternion, translation) pairs: 28 B read, 28 B
      written (the stamps are kept);
    * rmh_vec3_moments (segmented FP64 sum, mean and covariance), 67,108,864 point pairs with
      a mask in 2,048 segments: 25 B read per pair.
    With do_cpu, the oracle's moments of the first 8 segments are compared with the GPU's."""
    import torch

    from rmagine_amd import math as M

    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(9)
    nt, npts = 1 << 24, 1 << 26

    def transforms(n):
        T = torch.rand((n, 8), generator=gen, device=dev) * 2.0 - 1.0
        T[:, :4] /= T[:, :4].norm(dim=1, keepdim=True)
        T[:, 7] = 0.0
        return T

    A, B = transforms(nt), transforms(nt)
    C = torch.empty_like(A)
    out = {}

    def leg(name, fn, nbytes, unit_count, unit):
        ms = _timed_ms(fn)
        out[name] = {"ms": round(ms, 4), unit: unit_count,
                     "roofline": roofline(name, nbytes, ms, note=f"algorithmic bytes: {nbytes / unit_count:.0f} B per {unit[:-1]}")}

    leg("k_tf_mul (rmh_transform_mul)", lambda: M.transform_mul(A, B, out=C), 96 * nt, nt, "transforms")
    leg("k_tf_inv (rmh_transform_inv)", lambda: M.transform_inv(A, out=C), 64 * nt, nt, "transforms")
    q = A[:, :4].contiguous()
    t = A[:, 4:7].contiguous()
    leg("k_tf_pack_quat (rmh_transform_pack_quat)", lambda: M.pack(q, t, out=C), 56 * nt, nt, "transforms")
    del B, q, t
    P = torch.rand((npts, 3), generator=gen, device=dev) * 100.0 - 50.0
    Q = torch.empty_like(P)
    leg("k_tf_points (rmh_transform_points)", lambda: M.transform_points(A[:1], P, out=Q), 24 * npts, npts, "points")
    del A, C
    mask = (torch.rand(npts, generator=gen, device=dev) > 0.1).to(torch.uint8)
    seg = npts // 2048
    res = {}

    def moments():
        res.update(M.vec3_moments(P, Q, mask, segment_size=seg, n_segments=2048))
    leg("k_moments (rmh_vec3_moments)", moments, 25 * npts, npts, "pairs")
    if do_cpu:
        from oracle import stats as OS
        k = 8
        o = OS.moments(P[:k * seg].cpu().numpy(), Q[:k * seg].cpu().numpy(), mask[:k * seg].cpu().numpy(),
                       segment_size=seg, n_segments=k)
        g = {key: res[key][:k].cpu().numpy() for key in ("sum", "mean", "cov", "count")}
        rel = max(float(np.max(np.abs(g[key] - o[key]) / np.maximum(np.abs(o[key]), 1.0))) for key in ("sum", "mean", "cov"))
        out["k_moments (rmh_vec3_moments)"]["oracle_check"] = {"segments": k, "max_rel_err": rel,
                                                                "counts_equal": bool(np.array_equal(g["count"], o["count"]))}
    del P, Q, mask
    torch.cuda.empty_cache()
    return out


def compare_layouts(names) -> None:
    """Binary against 4-wide quantised node layout on the same map: ms per simulate() (min of
    5 HIP-event timings after 2 warm-ups) and a bit-for-bit comparison of every output."""
    import torch

    from rmagine_amd import scenes
    from rmagine_amd import types as T
    from rmagine_amd.simulation import HipMap, simulator_for

    def bits(x):
        return x.view(torch.int32) if x.dtype == torch.float32 else x

    def run(name, V, F, model, P, attrs):
        P = torch.as_tensor(P, device="cuda")
        out, res = {}, {}
        for wide in (False, True):
            hm = HipMap(V, F, wide=wide)
            sim = simulator_for(model, hm)
            r = sim.simulate(P, attrs=attrs)
            sim.simulate(P, r)
            torch.cuda.synchronize()
            ms = []
            for _ in range(5):
                a, b = _events()
                a.record()
                sim.simulate(P, r)
                b.record()
                b.synchronize()
                ms.append(a.elapsed_time(b))
            inf = hm.info()
            out["wide" if wide else "binary"] = {"ms": round(min(ms), 3), "build_ms": round(inf["build_ms"], 2),
                                                 "wide_nodes": inf["n_wide_nodes"], "wide_levels": inf["wide_levels"],
                                                 "binary_depth": inf["max_depth"]}
            res[wide] = r
            del sim, hm
        same = all(torch.equal(bits(getattr(res[False], n)), bits(getattr(res[True], n))) for n in attrs)
        rays = P.shape[0] * model.size()
        print(json.dumps({"workload": name, "rays": rays, "identical": same,
                          "mrays_s": {k: round(rays / v["ms"] / 1e3, 1) for k, v in out.items()}, **out}), flush=True)
        del res
        torch.cuda.empty_cache()

    names = names or ["terrain", "particles", "icp", "room"]
    for name in names:
        if name == "terrain":
            V, F, P = scenes.terrain_workload()
            run(name, V, F, scenes.ouster_workload_model(), P, T.ATTRIBUTES)
        elif name == "particles":
            V, F, _, P = scenes.fine_room_workload()
            run(name, V, F, scenes.vlp16_workload_model(), P, ("hits", "ranges"))
        elif name == "icp":
            V, F = scenes.objects_mesh()
            rng = np.random.default_rng(4)
            model = T.O1DnModel(32768, 1, T.Interval(0.1, 100.0), (0.0, 0.0, 0.0), scenes.random_unit_dirs(32768, rng))
            hyp = np.repeat(T.make_transforms([[0.0, 0.0, 0.0]], [[0.0, 0.0, 0.0, 1.0]]), 2000, 0).copy()
            hyp[:, 4:7] += rng.normal(scale=0.2, size=(2000, 3)).astype(np.float32)
            run(name, V, F, model, hyp, ("hits", "ranges", "points", "normals"))
        elif name == "room":
            V, F, P = scenes.room_workload()
            run(name, V, F, scenes.vlp16_workload_model(), P, ("hits", "ranges"))
        else:
            raise SystemExit(f"unknown workload {name!r}")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    if args.copy_calibration:
        n = 1 << 30
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "launches": 5}))
        return
    if args.compare_layouts is not None:
        compare_layouts(args.compare_layouts)
        return
    if args.visit_stats:
        from rmagine_amd import scenes
        from rmagine_amd.simulation import HipMap, SphereSimulatorHip
        V, F, Tbm = scenes.terrain_workload(n_grid=args.grid, n_sensors=args.sensors)
        sim = SphereSimulatorHip(HipMap(V, F))
        sim.setModel(scenes.ouster_workload_model())
        st = sim.visit_stats(torch.as_tensor(Tbm[:args.visit_stats], device="cuda"))
        st["active_lanes_per_node_iteration"] = ((st["vector_node_visits"] + st["scalar_node_visits"]) /
                                                 max(1, st["wave_iterations"]))
        st["uniform_share_of_wave_iterations"] = st["uniform_wave_iterations"] / max(1, st["wave_iterations"])
        st["vector_node_bytes_per_ray"] = 56 * st["vector_node_visits_per_ray"]  # 3 x 16 B boxes + 8 B refs
        st["triangle_bytes_per_ray"] = 48 * st["triangle_tests_per_ray"]
        print(json.dumps({"visit_stats": st, "workload": f"terrain {args.grid}^2, first {args.visit_stats} of "
                                                        f"{len(Tbm)} Ouster-style poses"}))
        return

    if args.extra_only:
        fn = {"batched": lambda: extra_batched(0.0, False, tune=False), "pose_gradient": extra_pose_grad,
              "icp": extra_icp, "next_rows": lambda: extra_next_rows(not args.no_cpu),
              "math_layer": lambda: extra_math_layer(not args.no_cpu)}[args.extra_only]
        print(json.dumps({args.extra_only: fn()}))
        return

    from rmagine_amd import scenes
    from rmagine_amd.simulation import HipMap, SphereSimulatorHip

    dev = torch.device("cuda:0")
    t0 = time.time()
    V, F, Tbm = scenes.terrain_workload(n_grid=args.grid, n_sensors=args.sensors)
    log(f"workload: {len(F)} tris, {len(Tbm)} poses ({time.time() - t0:.1f} s)")
    model = scenes.ouster_workload_model()
    hm = HipMap(V, F, device=dev)
    info = hm.info()
    log(f"map: {info}")
    sim = SphereSimulatorHip(hm, device=dev)
    sim.setModel(model)
    Tbm_d = torch.as_tensor(Tbm, device=dev)
    res = sim.simulate(Tbm_d)  # allocates all 7 attributes (resident in HBM)
    torch.cuda.synchronize()
    rays_per_step = model.size() * len(Tbm)

    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        sim.simulate(Tbm_d, res, stream=stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        sim.simulate(Tbm_d, res, stream=stream)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    kernel_ms = [a.elapsed_time(b) for a, b in ev]
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    value = rays_per_step * args.steps / wall / 1e6
    alg_bytes = BYTES_PER_RAY_ALL * rays_per_step
    achieved = alg_bytes / avg_kernel_s / 1e9
    log(f"gpu: {args.steps} steps in {wall:.3f} s, kernel avg {avg_kernel_s * 1e3:.3f} ms "
        f"(min {min(kernel_ms):.3f}, max {max(kernel_ms):.3f}), {value:.1f} Mrays/s")

    per_pose = model.size()

    def gpu_numpy_for_pose(p):
        sl = slice(p * per_pose, (p + 1) * per_pose)
        return {"hits": res.hits[sl].cpu().numpy(), "ranges": res.ranges[sl].cpu().numpy(),
                "points": res.points[sl].cpu().numpy(), "normals": res.normals[sl].cpu().numpy(),
                "face_ids": res.face_ids[sl].cpu().numpy().view(np.uint32),
                "geom_ids": res.geom_ids[sl].cpu().numpy().view(np.uint32),
                "object_ids": res.object_ids[sl].cpu().numpy().view(np.uint32)}

    cpu = None
    accuracy = None
    if not args.no_cpu:
        cpu, accuracy = cpu_baseline(V, F, model, Tbm, args.cpu_seconds, gpu_numpy_for_pose)

    default_workload = args.grid == 1024 and args.sensors == 1000
    traffic = profiled(default_workload, "traffic_bytes_per_launch", int)
    l2_hit = profiled(default_workload, "l2_hit_rate", float)
    hbm_measured = measure_hbm_gbs()
    log(f"hbm: plain device copy {hbm_measured:.0f} GB/s (read + write)")
    out = {
        "metric": "Mrays/s at <=1e-5 m range err vs Embree semantics (simulate, all attributes)",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"terrain {args.grid}x{args.grid} fBm at 0.1 m ({len(F)} tris), {len(Tbm)} Ouster-style "
                        f"spherical sensors 128x1024 (+-22.5 deg, 0.1-100 m), all 7 attributes",
            "rays_per_step": rays_per_step,
            "bvh": {k: info[k] for k in ("n_nodes", "max_depth", "build_ms", "sah_cost")},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "traffic_source": TRAFFIC_PROFILE if traffic else None,
            # the profiled HBM-side bytes over this run's kernel time (fetches include lines
            # the Infinity Cache serves, so this bounds the HBM rate from above)
            "traffic_gbs": round(traffic / avg_kernel_s / 1e9, 1) if traffic else None,
            "traffic_frac": round(traffic / avg_kernel_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
            "l2_hit_rate": l2_hit,
            # what does bind the kernel (same profile): the texture data path that returns
            # vector-load data is busy this share of the kernel, and the VALU issues this
            # many instructions per SIMD and cycle (a wave64 FP32 fma takes 4 cycles: 0.25)
            "td_busy": profiled(default_workload, "td_busy", float),
            "valu_instr_per_simd_cycle": profiled(default_workload, "valu_instr_per_simd_cycle", float),
            "kernel": "k_simulate<spherical> (traversal + attribute writers)",
            "kernel_avg_ms": round(avg_kernel_s * 1e3, 4),
            "alg_bytes_per_launch": alg_bytes,
            "copy_gbs_measured": round(hbm_measured, 1),
            "frac_of_measured_copy": round(achieved / hbm_measured, 5),
        },
        "cpu_baseline": cpu,
        "accuracy_vs_oracle": accuracy,
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    if not args.no_extra:
        extra = {}
        del res
        torch.cuda.empty_cache()
        for name, fn in (("batched_heterogeneous", lambda: extra_batched(8.0, not args.no_cpu)),
                         ("pose_gradient", extra_pose_grad), ("icp_graph", extra_icp),
                         ("next_rows", lambda: extra_next_rows(not args.no_cpu)),
                         ("math_layer", lambda: extra_math_layer(not args.no_cpu))):
            t0 = time.time()
            extra[name] = fn()
            log(f"{name}: {extra[name]} ({time.time() - t0:.1f} s)")
            torch.cuda.empty_cache()
        out["extra"] = extra
    print(json.dumps(out))


if __name__ == "__main__":
    main()
