"""Autotuner for the batched persistent launch (rmh_batch_*).

The persistent kernel has three launch parameters: rays handed out per work grab
(`chunk_rays`, a multiple of 64), resident waves per compute unit (`waves_per_cu`, which
sets the persistent grid) and the shape of the 64-ray tile a wave traces (`tile_rows` rows
x 64/tile_rows columns of a spherical or pinhole image).  Small chunks balance uneven
sensors better, large ones cost fewer atomics; more waves hide more latency until LDS and
registers run out; the tile shape decides how coherent a wave's rays are (on a spinning
lidar wide tiles win because rays of one row end at similar ranges).  The tuner times
every candidate on the caller's actual batch with HIP events and keeps the fastest:
first chunk x waves at the default tile, then the tile rows at the best pair, then the
depth of the traversal stack kept in LDS (`lds_stack`: 8, 10 or 16 rows of 8 B per ray; a
shallower one leaves more LDS for resident waves, a deeper one spills less to private
memory), then the ray order of the O1Dn/OnDn sensors (sorted by direction, or the list
order the caller gave), then (when asked with launches=(1, 2)) one persistent launch for
every entry or two (image sensors, then ray lists, each kernel with its own walk), then
(when asked with layouts=(False, True)) the BVH node layout of the batch's map: binary or
4-wide with quantised boxes.  Stack depth, ray order, launch split and layout change speed,
never results.  Results are cached as JSON keyed by device architecture, mesh
size (power of two bucket of the face count) and sensor mix, in $RMH_TUNE_CACHE or
~/.cache/rmagine_amd/autotune.json.

`tune_sensor` does the same for one O1Dn/OnDn simulator: it times the ray order (sorted,
list, or the model's grid in image tiles) and the walk (speculative or classic) of its
sensor on the caller's poses and keeps the fastest pair (a scanner's coherent ray grid walks fastest the classic way, often in its
own order; scattered rays such as ICP scans speculatively and sorted).

The reference has no counterpart (its launch shape is fixed by optixLaunch).
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path

import torch

DEFAULT_CHUNKS = (64, 128, 256, 512, 1024)
DEFAULT_WAVES = (8, 16, 24, 32)
DEFAULT_TILE_ROWS = (1, 2, 4, 8, 16)
DEFAULT_LDS_STACKS = (8, 10, 16)
DEFAULT_RAY_ORDERS = (True, False)  # True = direction-sorted


def cache_path() -> Path:
    p = os.environ.get("RMH_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "rmagine_amd" / "autotune.json"


def cache_key(n_faces: int, model_mix: dict, device=None) -> str:
    props = torch.cuda.get_device_properties(device or torch.cuda.current_device())
    arch = getattr(props, "gcnArchName", props.name)
    bucket = int(round(math.log2(max(1, n_faces))))
    mix = ",".join(f"{k}:{model_mix[k]}" for k in sorted(model_mix))
    return f"{arch}|faces~2^{bucket}|{mix}"


def _load(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def _store(path: Path, data: dict) -> None:
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home is not an error; the tuned values are still applied


def time_batch(batch, reps: int = 3, stream=None) -> float:
    """Median milliseconds of batch.simulate() over `reps` launches (after one warm-up)."""
    stream = stream or torch.cuda.current_stream()
    batch.simulate(stream=stream)
    ms = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        batch.simulate(stream=stream)
        b.record(stream)
        b.synchronize()
        ms.append(a.elapsed_time(b))
    ms.sort()
    return ms[len(ms) // 2]


def tune_batch(batch, key: str | None = None, chunks=DEFAULT_CHUNKS, waves=DEFAULT_WAVES,
               tile_rows=DEFAULT_TILE_ROWS, reps: int = 3, use_cache: bool = True, stream=None,
               layouts=(False,), lds_stacks=DEFAULT_LDS_STACKS, ray_orders=DEFAULT_RAY_ORDERS,
               launches=(1,)) -> dict:
    """Picks (chunk_rays, waves_per_cu, tile_rows, lds_stack, ray order) for `batch` and, if
    `layouts` holds both False and True, the node layout of its map (True = 4-wide
    quantised); applies the choice and returns the record.  Re-capture graphs of the batch
    afterwards."""
    path = cache_path()
    hm = batch.map
    if use_cache and key:
        rec = _load(path).get(key)
        if rec:
            batch.set_params(chunk_rays=rec["chunk_rays"], waves_per_cu=rec["waves_per_cu"],
                             tile_rows=rec.get("tile_rows", 4), lds_stack=rec.get("lds_stack", 0),
                             launches=rec.get("launches", 0))
            if "rays_sorted" in rec:
                batch.set_ray_order(rec["rays_sorted"])
            if "wide" in rec and hm.wide != rec["wide"]:
                hm.set_layout(rec["wide"])
            return dict(rec, cached=True)
    trials = []
    rows0 = batch.params()["tile_rows"]
    for c in chunks:
        for w in waves:
            batch.set_params(chunk_rays=c, waves_per_cu=w, tile_rows=rows0)
            trials.append((time_batch(batch, reps, stream), c, w, rows0))
    ms, c, w, r = min(trials)
    for rr in tile_rows:
        if rr == rows0:
            continue
        batch.set_params(chunk_rays=c, waves_per_cu=w, tile_rows=rr)
        trials.append((time_batch(batch, reps, stream), c, w, rr))
    trials.sort()
    ms, c, w, r = trials[0]
    batch.set_params(chunk_rays=c, waves_per_cu=w, tile_rows=r)
    rec = {"chunk_rays": c, "waves_per_cu": w, "tile_rows": r, "ms": ms,
           "trials": [{"chunk_rays": tc, "waves_per_cu": tw, "tile_rows": tr, "ms": tm} for tm, tc, tw, tr in trials]}
    keys = ["chunk_rays", "waves_per_cu", "tile_rows", "ms"]
    lds0 = batch.params()["lds_stack"]
    stack_ms = {lds0: ms}
    for d in lds_stacks:
        if d not in stack_ms:
            batch.set_params(lds_stack=d)
            stack_ms[d] = time_batch(batch, reps, stream)
    best_d = min(stack_ms, key=stack_ms.get)
    batch.set_params(lds_stack=best_d)
    ms = stack_ms[best_d]
    rec.update(lds_stack=best_d, ms=ms, lds_stack_ms=stack_ms)
    keys.append("lds_stack")
    order0 = batch.ray_order_sorted()
    order_ms = {order0: ms}
    for o in ray_orders:
        if bool(o) not in order_ms:
            batch.set_ray_order(bool(o))
            if batch.ray_order_sorted() == bool(o):  # image-only batches have one order
                order_ms[bool(o)] = time_batch(batch, reps, stream)
    best_o = min(order_ms, key=order_ms.get)
    if batch.ray_order_sorted() != best_o:
        batch.set_ray_order(best_o)
    ms = order_ms[best_o]
    rec.update(rays_sorted=best_o, ms=ms, ray_order_ms={("sorted" if k else "list"): v for k, v in order_ms.items()})
    keys.append("rays_sorted")
    launch0 = batch.params()["launches"]
    launch_ms = {launch0: ms}
    for n in launches:
        if n not in launch_ms:
            batch.set_params(launches=n)
            launch_ms[n] = time_batch(batch, reps, stream)
    best_l = min(launch_ms, key=launch_ms.get)
    batch.set_params(launches=best_l)
    ms = launch_ms[best_l]
    rec.update(launches=best_l, ms=ms, launches_ms=launch_ms)
    keys.append("launches")
    if len(set(layouts)) > 1:
        start = hm.wide
        timed = {start: ms}
        for wide in layouts:
            if wide in timed:
                continue
            if hm.set_layout(wide) != wide:  # this tree has no 4-wide form
                continue
            timed[wide] = time_batch(batch, reps, stream)
        best = min(timed, key=timed.get)
        if hm.wide != best:
            hm.set_layout(best)
        rec.update(wide=best, ms=timed[best], layout_ms={("wide" if k else "binary"): v for k, v in timed.items()})
        keys.append("wide")
    if key:
        data = _load(path)
        data[key] = {k: rec[k] for k in keys}
        _store(path, data)
    return dict(rec, cached=False)


def tune_sensor(sim, Tbm, bundle=None, reps: int = 3, stream=None) -> dict:
    """Times every (ray order, walk) pair of an O1Dn/OnDn simulator's sensor on `Tbm` with
    HIP events (median of `reps` launches after one warm-up): orders "sorted", "list" and
    "tiles" (the model's W x H grid in 4 x 16 tiles), walks speculative and classic.  Applies
    the fastest and returns {"ray_order", "rays_sorted", "speculative", "ms", "trials"}.  Image
    simulators have one choice and return it untimed.  Results never change; no launch of the
    sensor may be in flight."""
    from . import types as T
    if not isinstance(sim.model, (T.O1DnModel, T.OnDnModel)):
        return {"ray_order": "list", "rays_sorted": False, "speculative": False, "ms": None, "trials": {}}
    stream = stream or torch.cuda.current_stream()
    if bundle is None:
        bundle = sim.simulate(Tbm, stream=stream)
    trials = {}
    for order in ("sorted", "list", "tiles"):
        sim.set_ray_order(order)
        for spec in (True, False):
            sim.set_walk(spec)
            sim.simulate(Tbm, bundle, stream=stream)
            ms = []
            for _ in range(reps):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                sim.simulate(Tbm, bundle, stream=stream)
                b.record(stream)
                b.synchronize()
                ms.append(a.elapsed_time(b))
            ms.sort()
            trials[(order, spec)] = ms[len(ms) // 2]
    (best_order, best_spec), best = min(trials.items(), key=lambda kv: kv[1])
    sim.set_ray_order(best_order)
    sim.set_walk(best_spec)
    return {"ray_order": best_order, "rays_sorted": best_order == "sorted", "speculative": best_spec, "ms": best,
            "trials": {f"{k[0]}/{'speculative' if k[1] else 'classic'}": round(v, 3) for k, v in trials.items()}}
