"""Batch autotuner (rmagine_amd.autotune): it applies the fastest launch shape, LDS stack
depth, ray order and node layout it timed, records them, re-applies them from its cache, and never changes results
(every output byte equal before and after tuning)."""
import json

import numpy as np
import pytest
import torch

from rmagine_amd import autotune, scenes
from rmagine_amd import types as T
from rmagine_amd.simulation import BatchSimulatorHip, HipMap, simulator_for

pytestmark = pytest.mark.gpu


def _outputs(batch):
    out = [b.to_numpy() for b in batch.simulate()]
    torch.cuda.synchronize()
    return out


def test_tune_batch_shape_and_layout(tmp_path, monkeypatch):
    monkeypatch.setenv("RMH_TUNE_CACHE", str(tmp_path / "tune.json"))
    V, F, P = scenes.room_workload(n_sensors=32)
    hm = HipMap(V, F)
    rng = np.random.default_rng(2)
    d = rng.normal(size=(2048, 3))
    o1 = T.O1DnModel(2048, 1, T.Interval(0.1, 60.0), (0.0, 0.0, 0.05),
                     (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32))
    entries = [{"model": m, "Tbm": torch.as_tensor(P[i * 16:(i + 1) * 16], device="cuda"), "attrs": T.ATTRIBUTES}
               for i, m in enumerate((scenes.vlp16_workload_model(), o1))]
    batch = BatchSimulatorHip(hm, entries)
    before = _outputs(batch)
    rec = autotune.tune_batch(batch, key="room-test", chunks=(64, 256), waves=(8, 16), tile_rows=(2, 4),
                              reps=1, layouts=(False, True), launches=(1, 2))
    assert not rec["cached"] and rec["wide"] in (False, True)
    assert set(rec["layout_ms"]) == {"binary", "wide"}
    assert rec["ms"] == min(rec["layout_ms"].values())
    p = batch.params()
    assert (p["chunk_rays"], p["waves_per_cu"], p["tile_rows"]) == (rec["chunk_rays"], rec["waves_per_cu"], rec["tile_rows"])
    assert hm.wide == rec["wide"]
    assert len(rec["trials"]) == 2 * 2 + 1
    assert rec["lds_stack"] in (8, 10, 16) and set(rec["lds_stack_ms"]) == {8, 10, 16}
    assert p["lds_stack"] == rec["lds_stack"]
    assert set(rec["ray_order_ms"]) == {"sorted", "list"}
    assert batch.ray_order_sorted() == rec["rays_sorted"]
    assert rec["launches"] in (1, 2) and set(rec["launches_ms"]) == {1, 2} and p["launches"] == rec["launches"]
    after = _outputs(batch)
    for a, b in zip(before, after):
        for k in a:
            np.testing.assert_array_equal(a[k].view(np.uint8), b[k].view(np.uint8), err_msg=k)
    stored = json.loads((tmp_path / "tune.json").read_text())["room-test"]
    assert stored["wide"] == rec["wide"] and stored["chunk_rays"] == rec["chunk_rays"]
    assert stored["lds_stack"] == rec["lds_stack"] and stored["rays_sorted"] == rec["rays_sorted"]
    assert stored["launches"] == rec["launches"]
    # a fresh batch on a map in the other layout picks the cached choice up
    hm.set_layout(not rec["wide"])
    batch.set_params(chunk_rays=1024, waves_per_cu=32, tile_rows=8, lds_stack=16 if rec["lds_stack"] != 16 else 8)
    batch.set_ray_order(not rec["rays_sorted"])
    batch.set_params(launches=3 - rec["launches"])
    again = autotune.tune_batch(batch, key="room-test", layouts=(False, True))
    assert again["cached"] and hm.wide == rec["wide"]
    assert batch.params()["chunk_rays"] == rec["chunk_rays"] and batch.params()["tile_rows"] == rec["tile_rows"]
    assert batch.params()["lds_stack"] == rec["lds_stack"] and batch.ray_order_sorted() == rec["rays_sorted"]
    assert batch.params()["launches"] == rec["launches"]
    for a, b in zip(before, _outputs(batch)):
        for k in a:
            np.testing.assert_array_equal(a[k].view(np.uint8), b[k].view(np.uint8), err_msg=k)


def test_batch_lds_stack_depths_and_validation():
    """Every LDS stack depth gives the same bytes; unsupported depths are refused."""
    V, F, P = scenes.room_workload(n_sensors=8)
    hm = HipMap(V, F)
    entries = [{"model": scenes.vlp16_workload_model(), "Tbm": torch.as_tensor(P, device="cuda"),
                "attrs": T.ATTRIBUTES}]
    batch = BatchSimulatorHip(hm, entries, lds_stack=8)
    assert batch.params()["lds_stack"] == 8
    ref = _outputs(batch)
    for d in (10, 16):
        batch.set_params(lds_stack=d)
        assert batch.params()["lds_stack"] == d
        for a, b in zip(ref, _outputs(batch)):
            for k in a:
                np.testing.assert_array_equal(a[k].view(np.uint8), b[k].view(np.uint8), err_msg=f"{k} lds={d}")
    for bad in (4, 12, 24, 32):
        with pytest.raises(Exception):
            batch.set_params(lds_stack=bad)
        assert batch.params()["lds_stack"] == 16


def test_batch_one_or_two_launches():
    """One persistent launch for every entry (the default) and the split into an image
    kernel and a ray-list kernel write the same bytes; other launch counts are refused."""
    V, F, P = scenes.room_workload(n_sensors=15)
    hm = HipMap(V, F)
    rng = np.random.default_rng(8)
    d = rng.normal(size=(1500, 3))
    o1 = T.O1DnModel(1500, 1, T.Interval(0.1, 60.0), (0.0, 0.0, 0.05),
                     (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32))
    pin = T.PinholeModel(96, 64, T.Interval(0.05, 80.0), (50.0, 50.0), (47.5, 31.5))
    # a scanner-grid list (its sensor walks the classic way in its own simulate, the batch
    # speculatively: the bytes must agree) beside scattered lists
    ph, th = np.meshgrid(np.radians(np.linspace(-15.0, 15.0, 16)), np.radians(np.arange(90) * 4.0), indexing="ij")
    gd = np.stack([np.cos(ph) * np.cos(th), np.cos(ph) * np.sin(th), np.sin(ph)], -1).reshape(-1, 3).astype(np.float32)
    grid = T.O1DnModel(90, 16, T.Interval(0.1, 60.0), (0.0, 0.0, 0.05), gd)
    models = (o1, scenes.vlp16_workload_model(), pin, o1, grid)  # list entries before image entries
    entries = [{"model": m, "Tbm": torch.as_tensor(P[3 * i:3 * i + 3], device="cuda"), "attrs": T.ATTRIBUTES}
               for i, m in enumerate(models)]
    batch = BatchSimulatorHip(hm, entries)
    assert batch.params()["launches"] == 1
    one = _outputs(batch)
    batch.set_params(launches=2)
    assert batch.params()["launches"] == 2
    for a, b in zip(one, _outputs(batch)):
        for k in a:
            np.testing.assert_array_equal(a[k].view(np.uint8), b[k].view(np.uint8), err_msg=k)
    for i, m in enumerate(models):  # and each entry equals its own simulate()
        sim = simulator_for(m, hm)
        ref = sim.simulate(torch.as_tensor(P[3 * i:3 * i + 3], device="cuda"), attrs=T.ATTRIBUTES).to_numpy()
        for k in ref:
            np.testing.assert_array_equal(one[i][k].view(np.uint8), ref[k].view(np.uint8), err_msg=f"entry {i} {k}")
    with pytest.raises(Exception):
        batch.set_params(launches=3)
    assert batch.params()["launches"] == 2


def test_batch_refuses_entries_beyond_32_bit_tile_indices():
    """The batched kernel enumerates an entry's 64-ray tiles with 32-bit indices: an entry of
    2^32 tiles or more is refused with RMH_ERR_UNSUPPORTED when the batch is built, before any
    launch; one tile pose fewer is accepted.  Only the device table is built (no outputs are
    allocated and nothing is launched), so the pose pointer is never read."""
    import ctypes

    from rmagine_amd import _lib
    from rmagine_amd.simulation import _SensorHandle

    V, F, _ = scenes.room_workload(n_sensors=1)
    hm = HipMap(V, F)
    sh = _SensorHandle(scenes.ouster_workload_model(), torch.device("cuda"))
    Tbm = torch.zeros(1, 8, dtype=torch.float32, device="cuda")
    tiles = (128 // 4) * (1024 // 16)  # tile_rows 4: 4 x 16 rays per tile
    prm = T.rmh_batch_params(0, 0, 4, 0, 0)

    def create(n_poses):
        table = (T.rmh_batch_entry * 1)()
        table[0] = T.rmh_batch_entry(sh.handle.value, T.transform_to_c(T.identity_transforms(1)[0]),
                                     int(Tbm.data_ptr()), n_poses, 0, T.rmh_outputs())
        h = ctypes.c_void_p()
        rc = _lib.lib().rmh_batch_create(hm.handle, table, 1, ctypes.byref(prm), ctypes.byref(h))
        return rc, h

    rc, h = create((1 << 32) // tiles)
    assert rc == 5 and not h.value  # RMH_ERR_UNSUPPORTED, nothing returned
    assert b"2^32" in _lib.lib().rmh_last_error()
    rc, h = create((1 << 32) // tiles - 1)
    assert rc == 0 and h.value
    rays, chunks = ctypes.c_uint64(), ctypes.c_uint32()
    _lib.call("rmh_batch_get_info", h, ctypes.byref(rays), ctypes.byref(chunks), None)
    assert rays.value == ((1 << 32) // tiles - 1) * 128 * 1024
    assert chunks.value == rays.value // 256
    _lib.call("rmh_batch_destroy", h)
