"""Many independent decks advanced together on one device (clover_ensemble_*).

Each deck is one tile with its own mesh, states, dt and time; every kernel
launch covers all decks, dt is a per-deck minimum formed on the device, and
the step of each advect_x parity is captured in a graph and replayed.  The
fields evolve exactly as a :class:`~cloverleaf3d_ref_amd.chunk.Chunk` running
each deck alone with the fused kernels.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import _lib
from .chunk import FIELD_NAMES, FP32, _CLASS, SmallTimestep, _extents, to_c_deck
from .deck import Deck


class Ensemble:
    def __init__(self, decks: Sequence[Deck], device: int = 0, fp32: bool = False):
        self.decks = list(decks)
        self._L = _lib.lib()
        arr = (_lib.CloverDeck * len(self.decks))(*[to_c_deck(d) for d in self.decks])
        self._arr = arr
        h = ctypes.c_void_p()
        _lib.check(self._L.clover_ensemble_create(arr, len(self.decks), device, FP32 if fp32 else 0, ctypes.byref(h)),
                   "clover_ensemble_create")
        self._h = h

    def close(self) -> None:
        if self._h:
            self._L.clover_ensemble_destroy(self._h)
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

    def run(self, steps: int, graphs: bool = True) -> None:
        """Advance every deck.  Raises SmallTimestep (after the steps ran)
        if some deck's dt has fallen below its dtmin, where the reference
        stops; state()'s small list gives the step it happened at."""
        rc = _lib.check(self._L.clover_ensemble_run(self._h, steps, 1 if graphs else 0), "clover_ensemble_run")
        if rc == 1:
            small = self.state()[3]
            raise SmallTimestep("small timestep in decks " +
                                ", ".join(f"{d} (step {s})" for d, s in enumerate(small) if s))

    def summaries(self) -> List[Dict[str, float]]:
        out = (ctypes.c_double * (5 * len(self.decks)))()
        _lib.check(self._L.clover_ensemble_summaries(self._h, out), "clover_ensemble_summaries")
        return [dict(vol=out[5 * d], mass=out[5 * d + 1], ie=out[5 * d + 2], ke=out[5 * d + 3],
                     press=out[5 * d + 4]) for d in range(len(self.decks))]

    def state(self):
        n = len(self.decks)
        step = ctypes.c_int()
        time, dt = (ctypes.c_double * n)(), (ctypes.c_double * n)()
        small = (ctypes.c_int * n)()
        _lib.check(self._L.clover_ensemble_state(self._h, ctypes.byref(step), time, dt, small), "clover_ensemble_state")
        return step.value, list(time), list(dt), list(small)

    def fetch(self, deck: int, name: str) -> np.ndarray:
        d = self.decks[deck]
        a = np.zeros(_extents(_CLASS[name], d.x_cells, d.y_cells, d.z_cells), dtype=np.float64, order="F")
        _lib.check(self._L.clover_ensemble_fetch_field(self._h, deck, FIELD_NAMES.index(name),
                                                       ctypes.c_void_p(a.ctypes.data)), "clover_ensemble_fetch_field")
        return a

    def device_bytes(self) -> int:
        return int(self._L.clover_ensemble_device_bytes(self._h))

    def autotune(self) -> Tuple[float, float]:
        """Pick the fastest launch shape of every batched marching kernel
        (every field keeps its value).  Returns the summed kernel
        milliseconds of one pass before and after."""
        before, after = ctypes.c_double(), ctypes.c_double()
        _lib.check(self._L.clover_ensemble_autotune(self._h, ctypes.byref(before), ctypes.byref(after)),
                   "clover_ensemble_autotune")
        return before.value, after.value

    def workgroups(self) -> Dict[str, int]:
        """Workgroups of each shaped batched kernel's launch: 0 for one per
        tile, else a persistent grid walking the tile list."""
        out = {}
        for k, name in enumerate(_lib.SHAPE_KERNELS):
            n = ctypes.c_int()
            _lib.check(self._L.clover_ensemble_get_workgroups(self._h, k, ctypes.byref(n)),
                       "clover_ensemble_get_workgroups")
            out[name] = n.value
        return out

    def set_shape(self, kernel: str, planes: int, xcd: int, workgroups: int = 0) -> None:
        """Launch shape of one batched kernel (planes >= 8); never changes a value."""
        _lib.check(self._L.clover_ensemble_set_shape(self._h, _lib.SHAPE_KERNELS.index(kernel), planes, xcd,
                                                     workgroups), "clover_ensemble_set_shape")

    def shapes(self) -> Dict[str, Tuple[int, int, int]]:
        out = {}
        for k, name in enumerate(_lib.SHAPE_KERNELS):
            planes, xcd, bx = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            _lib.check(self._L.clover_ensemble_get_shape(self._h, k, ctypes.byref(planes), ctypes.byref(xcd),
                                                         ctypes.byref(bx)), "clover_ensemble_get_shape")
            out[name] = (planes.value, xcd.value, bx.value)
        return out
