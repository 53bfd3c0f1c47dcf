"""Host-side driver API over the device-resident hydro loop.

A :class:`Chunk` holds one problem (all tiles of one chunk) in HBM and runs
the reference's hydro loop on it through the C ABI (`clover_chunk_*`).  The
host sees only dt per step and the field_summary totals, as in the
reference's own output (timestep.f90:115-118, field_summary.f90:111-113).
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib
from .deck import Deck

FIELD_NAMES = ["density0", "density1", "energy0", "energy1", "pressure", "viscosity", "soundspeed",
               "xvel0", "xvel1", "yvel0", "yvel1", "zvel0", "zvel1", "vol_flux_x", "vol_flux_y",
               "vol_flux_z", "mass_flux_x", "mass_flux_y", "mass_flux_z", "volume", "xarea", "yarea", "zarea"]
_CLASS = {**{n: "C" for n in ("density0", "density1", "energy0", "energy1", "pressure", "viscosity",
                               "soundspeed", "volume")},
          **{n: "V" for n in ("xvel0", "xvel1", "yvel0", "yvel1", "zvel0", "zvel1")},
          "vol_flux_x": "X", "mass_flux_x": "X", "xarea": "X", "vol_flux_y": "Y", "mass_flux_y": "Y",
          "yarea": "Y", "vol_flux_z": "Z", "mass_flux_z": "Z", "zarea": "Z"}


def _extents(cls: str, nx: int, ny: int, nz: int):
    return {"C": (nx + 4, ny + 4, nz + 4), "V": (nx + 5, ny + 5, nz + 5), "X": (nx + 5, ny + 4, nz + 4),
            "Y": (nx + 4, ny + 5, nz + 4), "Z": (nx + 4, ny + 4, nz + 5)}[cls]


def to_c_deck(d: Deck) -> _lib.CloverDeck:
    cd = _lib.CloverDeck()
    for name in ("x_cells", "y_cells", "z_cells", "xmin", "ymin", "zmin", "xmax", "ymax", "zmax", "dtinit",
                 "dtmax", "dtrise", "dtmin", "dtc_safe", "dtu_safe", "dtv_safe", "dtw_safe", "dtdiv_safe",
                 "end_time", "end_step", "summary_frequency", "tiles_per_chunk", "tile_dims"):
        setattr(cd, name, getattr(d, name))
    if len(d.states) > _lib.MAX_STATES:
        raise ValueError(f"at most {_lib.MAX_STATES} states are supported")
    cd.number_of_states = len(d.states)
    for i, st in enumerate(d.states):
        for n in ("density", "energy", "xvel", "yvel", "zvel", "xmin", "xmax", "ymin", "ymax", "zmin", "zmax",
                  "radius"):
            getattr(cd, f"state_{n}")[i] = getattr(st, n)
        cd.state_geometry[i] = st.geometry
    return cd


class SmallTimestep(RuntimeError):
    """dt fell below dtmin (timestep.f90:111, 120-122); `dt` is the small
    time step that was formed (None where no single one applies)."""

    def __init__(self, msg: str, dt: Optional[float] = None):
        super().__init__(msg)
        self.dt = dt


PER_CALL = 1  # CLOVER_CHUNK_PER_CALL
FP32 = 2  # CLOVER_CHUNK_FP32
HALO_CHECK = 4  # CLOVER_CHUNK_HALO_CHECK


class Chunk:
    """One deck resident in HBM.  Construction runs the reference's start
    sequence (mesh, states, pressure, halo prime).

    per_call=False (default) runs the fused kernels; per_call=True runs one
    kernel per reference loop nest.  Both produce bit-identical fields.
    fp32=True holds the fields in FP32 (the FP32-store form, fused kernels
    only); fetch() still returns FP64 arrays.  halo_check=True is a test mode
    of the fused kernels: the halos the fused step does not refresh (and the
    reference does) get wrong values, so that reading one changes results."""

    def __init__(self, deck: Deck, device: int = 0, per_call: bool = False, fp32: bool = False,
                 halo_check: bool = False):
        self.deck = deck
        self.per_call = per_call
        self.fp32 = fp32
        self._L = _lib.lib()
        self._cdeck = to_c_deck(deck)
        h = ctypes.c_void_p()
        flags = (PER_CALL if per_call else 0) | (FP32 if fp32 else 0) | (HALO_CHECK if halo_check else 0)
        _lib.check(self._L.clover_chunk_create_ex(ctypes.byref(self._cdeck), device, flags, ctypes.byref(h)),
                   "clover_chunk_create_ex")
        self._h = h

    def close(self) -> None:
        if self._h:
            self._L.clover_chunk_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ stepping
    def step(self) -> float:
        """One hydro step (hydro.f90:48-68); returns its dt."""
        dt = ctypes.c_double()
        rc = _lib.check(self._L.clover_chunk_step(self._h, ctypes.byref(dt)), "clover_chunk_step")
        if rc == 1:
            raise SmallTimestep(f"small timestep: dt={dt.value}", dt.value)
        return dt.value

    def field_summary(self) -> Dict[str, float]:
        """field_summary.f90:27-139 totals (plus step and time)."""
        out = (ctypes.c_double * 5)()
        _lib.check(self._L.clover_chunk_field_summary(self._h, out), "clover_chunk_field_summary")
        step, time, _ = self.state()
        return dict(step=step, time=time, vol=out[0], mass=out[1], ie=out[2], ke=out[3], press=out[4])

    def visit_prepare(self) -> None:
        """The device work visit.f90:65-87 does before its dump (ideal_gas,
        pressure and velocity halos, viscosity); fetch() then returns what
        the dump writes."""
        _lib.check(self._L.clover_chunk_visit_prepare(self._h), "clover_chunk_visit_prepare")

    def run(self, max_steps: Optional[int] = None, max_rows: int = 4096) -> List[Dict[str, float]]:
        """The hydro loop to end_step/end_time; summary rows as the
        reference prints them (excluding the start row)."""
        n = ctypes.c_int()
        rows = (_lib.CloverSummary * max_rows)()
        steps = max_steps if max_steps is not None else self.deck.end_step
        rc = _lib.check(self._L.clover_chunk_run(self._h, steps, rows, max_rows, ctypes.byref(n)),
                        "clover_chunk_run")
        out = [dict(step=r.step, time=r.time, vol=r.vol, mass=r.mass, ie=r.ie, ke=r.ke, press=r.press)
               for r in rows[:n.value]]
        if rc == 1:
            raise SmallTimestep("small timestep")
        return out

    def dt_control(self):
        """(control, jdt, kdt, x_pos, y_pos) of the last step's dt, as the
        reference's step line reports them (timestep.f90:115-116)."""
        ctl, j, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        x, y = ctypes.c_double(), ctypes.c_double()
        _lib.check(self._L.clover_chunk_dt_control(self._h, ctypes.byref(ctl), ctypes.byref(j), ctypes.byref(k),
                                                   ctypes.byref(x), ctypes.byref(y)), "clover_chunk_dt_control")
        return ctl.value, j.value, k.value, x.value, y.value

    def state(self):
        step, time, dt = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        self._L.clover_chunk_state(self._h, ctypes.byref(step), ctypes.byref(time), ctypes.byref(dt))
        return step.value, time.value, dt.value

    # ------------------------------------------------------------ inspection
    @property
    def ntiles(self) -> int:
        n = ctypes.c_int()
        self._L.clover_chunk_tiles(self._h, ctypes.byref(n))
        return n.value

    def tile_info(self, t: int):
        dims = (ctypes.c_int * 3)()
        offs = (ctypes.c_int * 3)()
        nb = (ctypes.c_int * 6)()
        _lib.check(self._L.clover_chunk_tile_info(self._h, t, dims, offs, nb), "clover_chunk_tile_info")
        return tuple(dims), tuple(offs), tuple(nb)

    def fetch(self, name: str, tile: int = 0) -> np.ndarray:
        """A field of one tile as a Fortran-ordered array with the
        reference's extents (index 0 = min-2)."""
        (nx, ny, nz), _, _ = self.tile_info(tile)
        a = np.zeros(_extents(_CLASS[name], nx, ny, nz), dtype=np.float64, order="F")
        _lib.check(self._L.clover_chunk_fetch_field(self._h, tile, FIELD_NAMES.index(name),
                                                    ctypes.c_void_p(a.ctypes.data)), "clover_chunk_fetch_field")
        return a

    def device_bytes(self) -> int:
        return int(self._L.clover_chunk_device_bytes(self._h))

    def set_profiling(self, on: bool) -> None:
        self._L.clover_chunk_set_profiling(self._h, 1 if on else 0)

    def autotune(self) -> Tuple[float, float]:
        """Pick the fastest launch shape of every fused marching kernel on
        this chunk's own arrays (every field keeps its value).  Returns the
        summed kernel milliseconds of one pass before and after."""
        before, after = ctypes.c_double(), ctypes.c_double()
        _lib.check(self._L.clover_chunk_autotune(self._h, ctypes.byref(before), ctypes.byref(after)),
                   "clover_chunk_autotune")
        return before.value, after.value

    def shapes(self) -> Dict[str, Tuple[int, int, int]]:
        """kernel -> (planes per workgroup, XCD-contiguous block order,
        workgroup width)"""
        out = {}
        for k, name in enumerate(_lib.SHAPE_KERNELS):
            planes, xcd, bx = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            _lib.check(self._L.clover_chunk_get_shape(self._h, k, ctypes.byref(planes), ctypes.byref(xcd),
                                                      ctypes.byref(bx)), "clover_chunk_get_shape")
            out[name] = (planes.value, xcd.value, bx.value)
        return out

    def set_shape(self, kernel: str, planes: int, xcd: int, block_x: int = 64) -> None:
        _lib.check(self._L.clover_chunk_set_shape(self._h, _lib.SHAPE_KERNELS.index(kernel), planes, xcd, block_x),
                   "clover_chunk_set_shape")

    def profile(self) -> Dict[str, tuple]:
        ms = (ctypes.c_double * 32)()
        calls = (ctypes.c_int * 32)()
        n = self._L.clover_chunk_profile(self._h, ms, calls, 32)
        return {self._L.clover_chunk_profile_name(i).decode(): (ms[i], calls[i]) for i in range(n)}
