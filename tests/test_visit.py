"""visit.f90's VTK dump format (cloverleaf3d_ref_amd/visit.py), on the CPU.

The expected text is worked by hand from visit.f90's format statements:
'(a)' records, '(a,3i12)' DIMENSIONS, '(a,i5,a)' coordinates headers,
'(a,i20)' / '(a,i20,a)' data headers, and every value '(e12.4)' on its own
line.  The dump schedule and index file are checked with a stand-in chunk.
"""
from __future__ import annotations

import numpy as np

from cloverleaf3d_ref_amd.deck import Deck
from cloverleaf3d_ref_amd.visit import VISIT_FLOOR, VisitWriter, file_name, vertex_coords, vtk_text
from cloverleaf3d_ref_amd.clover_out import fmt_e


def test_e12_4():
    assert fmt_e(1.0, 12, 4) == "  0.1000E+01"
    assert fmt_e(0.0, 12, 4) == "  0.0000E+00"
    assert fmt_e(-2.5e-5, 12, 4) == " -0.2500E-04"
    assert fmt_e(0.2, 12, 4) == "  0.2000E+00"
    assert fmt_e(9.99996, 12, 4) == "  0.1000E+02"  # the rounding carries into the exponent
    assert fmt_e(123456.0, 12, 4) == "  0.1235E+06"
    # a three-digit exponent drops the E (Fortran Ew.d)
    assert fmt_e(1.5e150, 12, 4) == "  0.1500+151"
    assert fmt_e(-2.5e-120, 12, 4) == " -0.2500-119"


def test_file_names():
    assert file_name(1, 0) == "clover.00001.00000.vtk"
    assert file_name(1, 10) == "clover.00001.00010.vtk"
    assert file_name(8, 123, 0, 8) == "clover.00008.00123.vtk"
    # the task term: rank 1 of a run with 8 tiles per chunk numbers its tiles from 9
    assert file_name(1, 5, 1, 8) == "clover.00009.00005.vtk"


def test_floor_is_the_single_precision_literal():
    assert VISIT_FLOOR == float(np.float32(1.0e-8))
    assert VISIT_FLOOR < 1.0e-8  # so a double 1e-8 is above it and is written


def test_vertex_coords():
    d = Deck(x_cells=10, y_cells=4, z_cells=3, xmin=0.0, xmax=10.0, ymin=-2.0, ymax=2.0, zmin=0.0, zmax=0.3)
    assert vertex_coords(d, 10, 1, 0) == [float(j) for j in range(11)]
    assert vertex_coords(d, 5, 6, 0) == [5.0, 6.0, 7.0, 8.0, 9.0, 10.0]  # second of two tiles along x
    assert vertex_coords(d, 4, 1, 1) == [-2.0, -1.0, 0.0, 1.0, 2.0]
    dz = 0.3 / 3.0
    assert vertex_coords(d, 3, 1, 2) == [0.0 + dz * j for j in range(4)]


def _tiny_fields():
    """2 x 1 x 1 cells, 3 x 2 x 2 vertices."""
    rho = np.array([0.2, 1.0]).reshape(2, 1, 1, order="F")
    e = np.array([1.0, 2.5]).reshape(2, 1, 1, order="F")
    p = np.array([0.08, 1.0]).reshape(2, 1, 1, order="F")
    q = np.array([1.0e-8, 5.0e-9]).reshape(2, 1, 1, order="F")  # kept / floored
    u = np.zeros((3, 2, 2), order="F")
    u[1, 0, 0] = -2.0e-8  # kept (magnitude above the floor)
    u[2, 1, 1] = -5.0e-9  # floored
    v = np.full((3, 2, 2), 0.5, order="F")
    w = np.zeros((3, 2, 2), order="F")
    w[0, 1, 0] = 3.0
    return rho, e, p, q, u, v, w


def test_vtk_text_layout():
    rho, e, p, q, u, v, w = _tiny_fields()
    text = vtk_text([0.0, 5.0, 10.0], [0.0, 2.0], [0.0, 2.0], rho, e, p, q, u, v, w)
    lines = text.split("\n")
    assert text.endswith("\n") and lines[-1] == ""
    zero, half = "  0.0000E+00", "  0.5000E+00"
    want = ["# vtk DataFile Version 3.0", "vtk output", "ASCII", "DATASET RECTILINEAR_GRID",
            "DIMENSIONS           3           2           2",
            "X_COORDINATES     3 double", zero, "  0.5000E+01", "  0.1000E+02",
            "Y_COORDINATES     2 double", zero, "  0.2000E+01",
            "Z_COORDINATES     2 double", zero, "  0.2000E+01",
            "CELL_DATA                    2", "FIELD FieldData 4",
            "density 1                    2 double", "  0.2000E+00", "  0.1000E+01",
            "energy 1                    2 double", "  0.1000E+01", "  0.2500E+01",
            "pressure 1                    2 double", "  0.8000E-01", "  0.1000E+01",
            "viscosity 1                    2 double", "  0.1000E-07", zero,
            "POINT_DATA                   12", "FIELD FieldData 3",
            # x fastest, then y, then z
            "x_vel 1                   12 double", zero, " -0.2000E-07"] + [zero] * 10 + [
            "y_vel 1                   12 double"] + [half] * 12 + [
            "z_vel 1                   12 double", zero, zero, zero, "  0.3000E+01"] + [zero] * 8
    assert lines[:-1] == want


class _FakeChunk:
    """The parts of Chunk a dump uses: one 2x1x1 tile, a step counter, and a
    record of visit_prepare calls."""

    def __init__(self):
        self.step_no = 0
        self.prepared = 0
        rho, e, p, q, u, v, w = _tiny_fields()
        pad = lambda a: np.pad(a, 2)  # the reference's extents: index 0 is min-2
        self.f = dict(density0=pad(rho), energy0=pad(e), pressure=pad(p), viscosity=pad(q), xvel0=pad(u),
                      yvel0=pad(v), zvel0=pad(w))

    def state(self):
        return self.step_no, 0.0, 0.0

    ntiles = 1

    def tile_info(self, t):
        return (2, 1, 1), (1, 1, 1), (-1,) * 6

    def visit_prepare(self):
        self.prepared += 1

    def fetch(self, name, tile=0):
        return self.f[name]


def test_dumps_and_index(tmp_path):
    d = Deck(x_cells=2, y_cells=1, z_cells=1, xmin=0.0, xmax=10.0, ymin=0.0, ymax=2.0, zmin=0.0, zmax=2.0,
             visit_frequency=2)
    c = _FakeChunk()
    vw = VisitWriter(d, str(tmp_path))
    for step in (0, 2, 3, 3):  # start, a dump step, the last step twice (hydro.f90:73-75 then :88)
        c.step_no = step
        assert vw.dump(c) == [file_name(1, step)]
    assert c.prepared == 4
    index = (tmp_path / "clover.visit").read_text().split("\n")
    assert index == ["!NBLOCKS     0", "clover.00001.00000.vtk", "clover.00001.00002.vtk",
                     "clover.00001.00003.vtk", "clover.00001.00003.vtk", ""]
    rho, e, p, q, u, v, w = _tiny_fields()
    want = vtk_text([0.0, 5.0, 10.0], [0.0, 2.0], [0.0, 2.0], rho, e, p, q, u, v, w)
    assert (tmp_path / "clover.00001.00002.vtk").read_text() == want
