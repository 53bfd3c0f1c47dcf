"""Loads the in-tree HIP library (libclover_hip.so) and declares its C ABI.

There is no fallback: if the library is missing or fails to load, every
entry point raises.  Build it with ``__graft_entry__.build()`` or
``make -C cloverleaf3d_ref_amd/csrc``.
"""
from __future__ import annotations

import ctypes
import pathlib
import re
from typing import List

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "libclover_hip.so"
HEADER = PKG.parent / "include" / "clover_hip.h"

MAX_STATES = 16


class CloverDeck(ctypes.Structure):
    """Mirror of `clover_deck` (include/clover_hip.h)."""
    _fields_ = [
        ("x_cells", ctypes.c_int), ("y_cells", ctypes.c_int), ("z_cells", ctypes.c_int),
        ("xmin", ctypes.c_double), ("ymin", ctypes.c_double), ("zmin", ctypes.c_double),
        ("xmax", ctypes.c_double), ("ymax", ctypes.c_double), ("zmax", ctypes.c_double),
        ("dtinit", ctypes.c_double), ("dtmax", ctypes.c_double), ("dtrise", ctypes.c_double),
        ("dtmin", ctypes.c_double),
        ("dtc_safe", ctypes.c_double), ("dtu_safe", ctypes.c_double), ("dtv_safe", ctypes.c_double),
        ("dtw_safe", ctypes.c_double), ("dtdiv_safe", ctypes.c_double),
        ("end_time", ctypes.c_double),
        ("end_step", ctypes.c_int),
        ("summary_frequency", ctypes.c_int),
        ("tiles_per_chunk", ctypes.c_int),
        ("tile_dims", ctypes.c_int),
        ("number_of_states", ctypes.c_int),
    ] + [(f"state_{n}", ctypes.c_double * MAX_STATES) for n in
         ("density", "energy", "xvel", "yvel", "zvel", "xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "radius")] + [
        ("state_geometry", ctypes.c_int * MAX_STATES),
    ]


class CloverSummary(ctypes.Structure):
    _fields_ = [("step", ctypes.c_int), ("time", ctypes.c_double), ("vol", ctypes.c_double),
                ("mass", ctypes.c_double), ("press", ctypes.c_double), ("ie", ctypes.c_double),
                ("ke", ctypes.c_double)]


class CloverError(RuntimeError):
    pass


_lib = None


def header_functions() -> List[str]:
    """Names of every function the public header declares."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
    return sorted(set(re.findall(r"\b(clover_[a-z0-9_A-Z]+)\s*\(", text)))


# launch-shape kernel indices (CLOVER_SHAPE_* of clover_hip.h)
SHAPE_KERNELS = ["visc_dt", "pdv_predict", "accel_pdv", "cell_x", "cell_y", "cell_z",
                 "mom_x", "mom_y", "mom_z"]


def lib() -> ctypes.CDLL:
    """The loaded library (raises CloverError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise CloverError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(str(LIB_PATH))
    vp = ctypes.c_void_p
    L.clover_hip_last_error.restype = ctypes.c_char_p
    L.clover_chunk_profile_name.restype = ctypes.c_char_p
    L.clover_chunk_device_bytes.restype = ctypes.c_longlong
    L.clover_chunk_device_bytes.argtypes = [vp]
    L.clover_chunk_create.argtypes = [ctypes.POINTER(CloverDeck), ctypes.c_int, ctypes.POINTER(vp)]
    L.clover_chunk_create_ex.argtypes = [ctypes.POINTER(CloverDeck), ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.clover_chunk_destroy.argtypes = [vp]
    L.clover_chunk_step.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.clover_chunk_field_summary.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.clover_chunk_visit_prepare.argtypes = [vp]
    L.clover_chunk_run.argtypes = [vp, ctypes.c_int, ctypes.POINTER(CloverSummary), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int)]
    L.clover_chunk_state.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]
    L.clover_chunk_tiles.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.clover_chunk_tile_info.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.clover_chunk_fetch_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    L.clover_chunk_set_profiling.argtypes = [vp, ctypes.c_int]
    L.clover_chunk_profile.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                       ctypes.c_int]
    L.clover_ensemble_create.argtypes = [ctypes.POINTER(CloverDeck), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(vp)]
    L.clover_ensemble_destroy.argtypes = [vp]
    L.clover_ensemble_run.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.clover_ensemble_summaries.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.clover_ensemble_state.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    L.clover_ensemble_fetch_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    L.clover_ensemble_device_bytes.restype = ctypes.c_longlong
    L.clover_ensemble_device_bytes.argtypes = [vp]
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.clover_chunk_autotune.argtypes = [vp, dp, dp]
    L.clover_chunk_dt_control.argtypes = [vp, ip, ip, ip, dp, dp]
    L.clover_chunk_get_shape.argtypes = [vp, ctypes.c_int, ip, ip, ip]
    L.clover_chunk_set_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.clover_ensemble_autotune.argtypes = [vp, dp, dp]
    L.clover_ensemble_get_shape.argtypes = [vp, ctypes.c_int, ip, ip, ip]
    L.clover_ensemble_get_workgroups.argtypes = [vp, ctypes.c_int, ip]
    L.clover_ensemble_set_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.clover_hip_fetch.argtypes = [vp]
    L.clover_hip_push.argtypes = [vp]
    L.clover_hip_forget.argtypes = [vp]
    _lib = L
    return L


def check(rc: int, what: str) -> int:
    if rc < 0:
        msg = lib().clover_hip_last_error()
        raise CloverError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc
