"""The C-ABI library loads and exports every entry point include/rmh.h declares (CPU)."""
import ctypes
import re
from pathlib import Path

from rmagine_amd import _lib
from rmagine_amd import types as T

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "rmh.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rmh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "rmh_simulate" in names and "rmh_map_build" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) <= set(_lib.exported_symbols())


def test_pod_sizes():
    assert ctypes.sizeof(T.rmh_transform) == 32
    assert ctypes.sizeof(T.rmh_spherical_model) == 32
    assert ctypes.sizeof(T.rmh_pinhole_model) == 32
    assert ctypes.sizeof(T.rmh_outputs) == 7 * 8


def test_error_path_without_device():
    # argument validation happens before any device call
    lib = _lib.lib()
    assert lib.rmh_simulate(None, None, None, None, 1, None, None) == 3  # no map
    assert b"no map" in lib.rmh_last_error()
    assert lib.rmh_sensor_create(7, ctypes.byref(T.rmh_spherical_model()), ctypes.byref(ctypes.c_void_p())) != 0
