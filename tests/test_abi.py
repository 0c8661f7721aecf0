"""The C-ABI library loads and exports every entry point include/rmh.h declares (CPU)."""
import ctypes
import re

import pytest
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


def test_map_size_limit_without_device():
    """Maps whose BVH could hold 2^26 or more interior nodes (more than 2^26 faces) are
    refused before any device work: the ray kernels address nodes with 32-bit byte offsets
    (index << 6).  The pointers are never dereferenced: the size check comes first."""
    lib = _lib.lib()
    out = ctypes.c_void_p()
    fake = ctypes.c_void_p(0x1000)
    rc = lib.rmh_map_build(fake, 3, fake, (1 << 26) + 1, None, 0, None, None, ctypes.byref(out))
    assert rc == 5  # RMH_ERR_UNSUPPORTED
    assert b"2^26" in lib.rmh_last_error()
    assert out.value is None
    rc = lib.rmh_map_build(fake, 3, fake, 0xFFFFFFFF, None, 0, None, None, ctypes.byref(out))
    assert rc == 5


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror in rmagine_amd.types has the size and field offsets that a C
    compiler gives the same struct from include/rmh.h."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {name: cls for name, cls in vars(T).items()
               if name.startswith("rmh_") and isinstance(cls, type) and issubclass(cls, ctypes.Structure)}
    assert "rmh_batch_params" in structs and len(structs) >= 19
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rmh.h"', 'int main(void) {']
    for name, cls in sorted(structs.items()):
        lines.append(f'  printf("{name} (sizeof) %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = Path(__file__).resolve().parents[1] / "include"
    subprocess.run([cc, "-std=c99", f"-I{inc}", str(src), "-o", str(exe)], check=True, capture_output=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        parts = line.split()
        got[tuple(parts[:-1])] = int(parts[-1])
    for name, cls in structs.items():
        assert got[(name, "(sizeof)")] == ctypes.sizeof(cls), name
        for field, _ in cls._fields_:
            assert got[(name, field)] == getattr(cls, field).offset, f"{name}.{field}"
