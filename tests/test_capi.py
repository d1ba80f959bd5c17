"""The C-ABI library loads and exports every symbol include/microcosmos_hip.h
declares; the host binding and the header agree.  No device calls."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from microcosmos_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def _header_functions():
    txt = (ROOT / "include" / "microcosmos_hip.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mc_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _header_functions()
    for n in ("mc_plan_create", "mc_step", "mc_rollout", "mc_gaussian_filter", "mc_steric_field",
              "mc_ib_lbm_step", "mc_rod_solve", "mc_last_error", "mc_version"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for n in _header_functions():
        assert hasattr(lib, n), n
    assert set(_header_functions()) == set(_lib.SIGNATURES)


def test_version_and_error_without_device():
    lib = _lib.load()
    assert b"gfx950" in lib.mc_version()
    assert lib.mc_last_error() == b""
    # a NULL plan pointer is rejected before any device work
    st = lib.mc_plan_create(None, 8, 8, 8, 8, 0, None, 0, None, 0, 4)
    assert st == 5
    assert b"NULL" in lib.mc_last_error()


def test_struct_layouts_match_header(tmp_path):
    """sizeof / offsetof of every ABI struct, from gcc on the header, equal the
    ctypes mirrors used by the host binding."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"mc_params": _lib.McParams, "mc_state": _lib.McState,
               "mc_edge_params": _lib.McEdgeParams, "mc_actuation": _lib.McActuation}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "microcosmos_hip.h"',
             "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True,
                                                         text=True, check=True).stdout.split("\n") if l)
    for cname, cls in structs.items():
        assert int(out[cname]) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "_LIB_PATH", tmp_path / "nope.so")
    with pytest.raises(RuntimeError, match="no fallback"):
        _lib.load()
