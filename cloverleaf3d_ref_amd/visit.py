"""VTK dumps for the device-resident driver, in visit.f90's format
(visit.f90:25-202).

Each dump first runs visit.f90's device work (ideal_gas, the pressure and
velocity halos, viscosity: Chunk.visit_prepare).  It then writes one
legacy-ASCII rectilinear-grid file per tile, clover.<tile>.<step>.vtk
(visit.f90:103-111).  A file holds:
  * the tile's vertex coordinates;
  * four cell fields: density0, energy0, pressure and viscosity;
  * three vertex fields: the level-0 velocities.
Every value is written as Fortran E12.4 on a line of its own.  Viscosity
values not above 1e-8, and velocities not above it in magnitude, are written
as 0 (visit.f90:157-196).  The 1e-8 is a default-real literal, so the
threshold is the single-precision 1e-8 widened to double.

The file names are appended to the index clover.visit (visit.f90:89-101).
Its first line is written by the first dump of a run: '!NBLOCKS ' and
parallel%task*tiles_per_chunk (visit.f90:51-58).  On one MPI rank that
product is 0, so one process writes '!NBLOCKS     0'.

The reference dumps at start (start.f90:138), every visit_frequency steps
(hydro.f90:73-75), and again after the last step (hydro.f90:88).  A last step
that is also a dump step is therefore written twice, under the same name.
clover_out.run follows that schedule.
"""
from __future__ import annotations

import os
from typing import List

import numpy as np

from .chunk import Chunk
from .clover_out import fmt_e, fmt_i
from .deck import Deck

# visit.f90:160 and :176-194 compare against the default-real literal 0.00000001
VISIT_FLOOR = float(np.float32(1.0e-8))
# parallel%task of the one process (MPI rank 0)
TASK = 0


def _f(v: int) -> float:
    """Fortran float() of a default integer: a REAL(4) value."""
    return float(np.float32(v))


def e12_4_lines(values) -> List[str]:
    """Each value as a Fortran '(e12.4)' record."""
    return [fmt_e(float(v), 12, 4) for v in values]


def vertex_coords(deck: Deck, cells: int, offset: int, axis: int) -> List[float]:
    """vertexx/y/z(min..max+1) of a tile, by the expressions of
    initialise_chunk.f90:32-44 and initialise_chunk_kernel.f90:68-70 (or the
    y and z twins): the tile's corner from its global offset, then one cell
    spacing per vertex."""
    lo = (deck.xmin, deck.ymin, deck.zmin)[axis]
    hi = (deck.xmax, deck.ymax, deck.zmax)[axis]
    n = (deck.x_cells, deck.y_cells, deck.z_cells)[axis]
    d = (hi - lo) / _f(n)
    corner = lo + d * _f(offset - 1)
    return [corner + d * _f(j) for j in range(cells + 1)]


def file_name(tile: int, step: int, task: int = 0, tiles_per_chunk: int = 1) -> str:
    """visit.f90:103-109: the tile and step numbers written '(i6)' after
    100000 is added, with the leading digit replaced by '.'."""
    chunk_name = "." + fmt_i(task * tiles_per_chunk + 100000 + tile, 6)[1:]
    step_name = "." + fmt_i(step + 100000, 6)[1:]
    return "clover" + chunk_name.strip() + step_name.strip() + ".vtk"


def vtk_text(xv: List[float], yv: List[float], zv: List[float], density0: np.ndarray, energy0: np.ndarray,
             pressure: np.ndarray, viscosity: np.ndarray, xvel0: np.ndarray, yvel0: np.ndarray,
             zvel0: np.ndarray) -> str:
    """One tile's file (visit.f90:112-198).  Cell fields have shape
    (nx, ny, nz) and vertex fields (nx+1, ny+1, nz+1), interior only; both
    are written x fastest."""
    nxc, nyc, nzc = density0.shape
    nxv, nyv, nzv = nxc + 1, nyc + 1, nzc + 1
    ncell, nvert = nxc * nyc * nzc, nxv * nyv * nzv
    out = ["# vtk DataFile Version 3.0", "vtk output", "ASCII", "DATASET RECTILINEAR_GRID",
           "DIMENSIONS" + fmt_i(nxv, 12) + fmt_i(nyv, 12) + fmt_i(nzv, 12)]
    for name, n, coords in (("X_COORDINATES ", nxv, xv), ("Y_COORDINATES ", nyv, yv), ("Z_COORDINATES ", nzv, zv)):
        out.append(name + fmt_i(n, 5) + " double")
        out += e12_4_lines(coords)
    out += ["CELL_DATA " + fmt_i(ncell, 20), "FIELD FieldData 4"]
    visc = np.where(viscosity > VISIT_FLOOR, viscosity, 0.0)
    for name, a in (("density", density0), ("energy", energy0), ("pressure", pressure), ("viscosity", visc)):
        out.append(name + " 1 " + fmt_i(ncell, 20) + " double")
        out += e12_4_lines(a.ravel(order="F"))
    out += ["POINT_DATA " + fmt_i(nvert, 20), "FIELD FieldData 3"]
    for name, a in (("x_vel", xvel0), ("y_vel", yvel0), ("z_vel", zvel0)):
        out.append(name + " 1 " + fmt_i(nvert, 20) + " double")
        out += e12_4_lines(np.where(np.abs(a) > VISIT_FLOOR, a, 0.0).ravel(order="F"))
    return "\n".join(out) + "\n"


class VisitWriter:
    """The dumps of one run into directory `path`."""

    def __init__(self, deck: Deck, path: str = "."):
        self.deck = deck
        self.path = path
        self.first_call = True

    def dump(self, chunk: Chunk) -> List[str]:
        """visit.f90:25-202 on the chunk's current state; returns the files
        written."""
        step = chunk.state()[0]
        ntiles = chunk.ntiles
        index = os.path.join(self.path, "clover.visit")
        if self.first_call:
            with open(index, "w") as f:
                f.write("!NBLOCKS " + fmt_i(TASK * self.deck.tiles_per_chunk, 5) + "\n")
            self.first_call = False
        chunk.visit_prepare()
        # tiles in the order of the reference's decomposition (clover.f90:311-346)
        names = [file_name(t + 1, step, TASK, self.deck.tiles_per_chunk) for t in range(ntiles)]
        with open(index, "a") as f:
            for n in names:
                f.write(n + "\n")
        for t, n in enumerate(names):
            (nx, ny, nz), (ox, oy, oz), _ = chunk.tile_info(t)
            cell = lambda name: chunk.fetch(name, t)[2:nx + 2, 2:ny + 2, 2:nz + 2]
            vert = lambda name: chunk.fetch(name, t)[2:nx + 3, 2:ny + 3, 2:nz + 3]
            text = vtk_text(vertex_coords(self.deck, nx, ox, 0), vertex_coords(self.deck, ny, oy, 1),
                            vertex_coords(self.deck, nz, oz, 2), cell("density0"), cell("energy0"),
                            cell("pressure"), cell("viscosity"), vert("xvel0"), vert("yvel0"), vert("zvel0"))
            with open(os.path.join(self.path, n), "w") as f:
                f.write(text)
        return names
