"""The C-ABI library loads and exports every symbol include/clover_hip.h
declares (no compute calls: this runs without a GPU)."""
from __future__ import annotations

import ctypes
import subprocess

import pytest

from cloverleaf3d_ref_amd import _lib

import kernel_specs as ks


def test_header_declares_every_reference_kernel():
    names = set(_lib.header_functions())
    for k in ks.SPECS:
        assert f"clover_hip_{k}" in names, k
    for f in ("clover_hip_init", "clover_hip_finalize", "clover_hip_fetch", "clover_chunk_create",
              "clover_chunk_run", "clover_chunk_field_summary"):
        assert f in names


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.fail(f"{_lib.LIB_PATH} not built (run __graft_entry__.build())")
    L = _lib.lib()
    missing = [n for n in _lib.header_functions() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(_lib.header_functions()) <= exported


def test_library_carries_gfx950_code():
    out = subprocess.run(["readelf", "-S", str(_lib.LIB_PATH)], capture_output=True, text=True, check=True).stdout
    assert ".hip_fatbin" in out


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "nope.so")
    with pytest.raises(_lib.CloverError):
        _lib.lib()


def test_deck_struct_matches_header():
    # the ctypes mirror lists the same members, in order, as the C struct
    import re
    text = _lib.HEADER.read_text()
    body = text[text.index("typedef struct {\n    int x_cells"):text.index("} clover_deck;")]
    members = []
    for decl in re.findall(r"^\s*(?:int|double)\s+([^;]+);", body, flags=re.M):
        for m in decl.split(","):
            members.append(m.strip().split("[")[0])
    assert members == [f[0] for f in _lib.CloverDeck._fields_]
    assert ctypes.sizeof(_lib.CloverDeck) > 0
