"""The C-ABI library loads and exports every symbol include/clover_hip.h
declares (no compute calls: this runs without a GPU)."""
from __future__ import annotations

import ctypes
import pathlib
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


# The reference's kernel modules and the subroutines the drivers call
# (ideal_gas_kernel.f90:23, viscosity_kernel.f90:28, calc_dt_kernel.f90:24,
# PdV_kernel.f90:28, revert_kernel.f90:25, accelerate_kernel.f90:23,
# flux_calc_kernel.f90:22, advec_cell_kernel.f90:26, advec_mom_kernel.f90:25,
# reset_field_kernel.f90:23, update_halo_kernel.f90:28,
# update_tile_halo_kernel.f90:25, field_summary_kernel.f90:23).
REFERENCE_MODULES = {
    "ideal_gas_kernel_module": ["ideal_gas_kernel"],
    "viscosity_kernel_module": ["viscosity_kernel"],
    "calc_dt_kernel_module": ["calc_dt_kernel"],
    "pdv_kernel_module": ["pdv_kernel"],
    "revert_kernel_module": ["revert_kernel"],
    "accelerate_kernel_module": ["accelerate_kernel"],
    "flux_calc_kernel_module": ["flux_calc_kernel"],
    "advec_cell_kernel_module": ["advec_cell_kernel"],
    "advec_mom_kernel_mod": ["advec_mom_kernel"],
    "reset_field_kernel_module": ["reset_field_kernel"],
    "update_halo_kernel_module": ["update_halo_kernel"],
    "update_tile_halo_kernel_module": [f"update_tile_halo_{f}_kernel" for f in
                                       ("left", "right", "top", "bottom", "front", "back")],
    "field_summary_kernel_module": ["field_summary_kernel"],
}


def test_fortran_shims_cover_every_reference_module():
    import re
    src = (_lib.PKG / "fortran" / "clover_hip_shims.f90").read_text().lower()
    modules = {}
    for m in re.finditer(r"^module (\w+)\s*$(.*?)^end module \1", src, flags=re.M | re.S):
        modules[m.group(1)] = re.findall(r"^\s*subroutine (\w+)\(", m.group(2), flags=re.M)
    for mod, subs in REFERENCE_MODULES.items():
        assert mod in modules, mod
        assert sorted(subs) == sorted(modules[mod]), mod
    # every C name the shims bind is declared in the header and exported
    bound = set(re.findall(r"name='(clover_hip_\w+)'", src))
    assert {f"clover_hip_{k.lower()}" for k in ks.SPECS} <= bound
    declared = {n.lower() for n in _lib.header_functions()}
    assert bound <= declared


def test_fortran_shims_compile():
    import shutil
    flang = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    if not pathlib.Path(flang).exists():
        pytest.skip("no flang in this image")
    subprocess.run(["make", "-C", str(_lib.PKG / "fortran")], check=True, capture_output=True)
    assert (_lib.PKG / "_build" / "fortran" / "clover_hip_shims.o").exists()


def test_fortran_shim_library_for_the_boundary_tests():
    """oracle/_fshim/libfshim.so (the shims behind C-callable wrappers, built
    by __graft_entry__.build() when flang is present) exports a wrapper per
    kernel and resolves the product library by its soname."""
    import shutil
    flang = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    if not pathlib.Path(flang).exists():
        pytest.skip("no flang in this image")
    so = _lib.PKG.parent / "oracle" / "_fshim" / "libfshim.so"
    assert so.exists(), "build it: make -C oracle fshim"
    out = subprocess.run(["nm", "-D", "--defined-only", str(so)], capture_output=True, text=True, check=True).stdout
    wrappers = {line.split()[-1] for line in out.splitlines() if " T ref_" in line}
    assert len(wrappers) == 18
    dyn = subprocess.run(["readelf", "-d", str(so)], capture_output=True, text=True, check=True).stdout
    assert "[libclover_hip.so]" in dyn
    dyn = subprocess.run(["readelf", "-d", str(_lib.LIB_PATH)], capture_output=True, text=True, check=True).stdout
    assert "Library soname: [libclover_hip.so]" in dyn


def test_noalloc_field_summary_rejects_oversized_tile():
    # the size check comes before any device call, so this runs without a GPU
    L = _lib.lib()
    fn = L.clover_hip_field_summary_kernel_noalloc
    fn.restype = ctypes.c_int
    L.clover_hip_last_error.restype = ctypes.c_char_p
    dummy = ctypes.c_void_p(256)
    rc = fn(1, 1300, 1, 1300, 1, 1300, *([dummy] * 8), ctypes.c_void_p(None))
    assert rc == -2
    assert b"exceeds the maximum" in L.clover_hip_last_error()
    rc = fn(1, 8, 1, 8, 1, 8, *([dummy] * 7), ctypes.c_void_p(None), ctypes.c_void_p(None))
    assert rc == -2 and b"null" in L.clover_hip_last_error()
