"""Ouster metadata loader on the reference's own OS0-128 metadata (dat/ouster_config,
copied under tests/golden/ouster): sizes from lidar_mode and the beam table, and the
geometry of every direction against the JSON's angles (elevation = beam altitude,
azimuth = column angle + beam azimuth offset), which a swapped axis, sign or index fails."""
import json
from pathlib import Path

import numpy as np
import pytest

from rmagine_amd import _lib
from rmagine_amd.ouster import o1dn_from_ouster_meta_file

META = Path(__file__).resolve().parent / "golden" / "ouster" / "os0-128x1024x20-meta.json"


def test_os0_128_directions():
    m = o1dn_from_ouster_meta_file(META)
    meta = json.loads(META.read_text())
    alt = np.array(meta["beam_altitude_angles"], np.float64)
    az = np.array(meta["beam_azimuth_angles"], np.float64)
    assert (m.width, m.height) == (1024, 128) and len(alt) == 128
    assert (m.range.min, m.range.max) == (np.float32(0.1), np.float32(80.0)) or (m.range.min, m.range.max) == (0.1, 80.0)
    d = np.asarray(m.dirs, np.float64).reshape(128, 1024, 3)
    np.testing.assert_allclose(np.linalg.norm(d, axis=2), 1.0, atol=3e-7)
    elev = np.degrees(np.arcsin(np.clip(d[..., 2], -1, 1)))
    np.testing.assert_allclose(elev, np.repeat(alt[:, None], 1024, axis=1), atol=1e-4)
    azim = np.degrees(np.arctan2(d[..., 1], d[..., 0]))
    want = (np.arange(1024)[None, :] * 360.0 / 1024 + az[:, None] + 180.0) % 360.0 - 180.0
    diff = (azim - want + 180.0) % 360.0 - 180.0
    assert np.abs(diff).max() < 1e-3


def test_bad_metadata(tmp_path):
    p = tmp_path / "m.json"
    p.write_text('{"lidar_mode": "1024x10", "beam_altitude_angles": [1, 2]}')
    with pytest.raises(_lib.RmhError):
        o1dn_from_ouster_meta_file(p)
    p.write_text('{"lidar_mode": 5}')
    with pytest.raises(_lib.RmhError):
        o1dn_from_ouster_meta_file(p)
    p.write_text('not json')
    with pytest.raises(_lib.RmhError):
        o1dn_from_ouster_meta_file(p)
