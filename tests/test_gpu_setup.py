"""Mesh and initial states painted on the device against the oracle's
initialise_chunk + generate_chunk, tile by tile, bit for bit.  Those C
restatements are pinned bit for bit against the reference's own
initialise_chunk_kernel and generate_chunk_kernel (cuboid, sphere and point
states, odd shapes, tile offsets) in test_oracle_setup_tiles.py.

A chunk runs start.f90's halo update (depth 2, every painted field) right
after painting, as the reference does, so the painted halo cells of
density, energy and the velocities are overwritten before anything reads
them; those fields are compared on the interior (cells min..max, vertices
min..max+1).  Volume and areas are not halo-updated and are compared whole."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle_lib as ol
from cloverleaf3d_ref_amd.chunk import Chunk
from cloverleaf3d_ref_amd.deck import parse_deck

pytestmark = pytest.mark.gpu

DECK = """*clover
 state 1 density=0.3 energy=0.9 {bgvel}
 state 2 density=1.7 energy=3.1 xvel=0.2 geometry=cuboid xmin=1.0 xmax=4.0 ymin=2.0 ymax=5.5 zmin=0.5 zmax=3.0
 state 3 density=0.8 energy=2.0 zvel=-0.1 geometry=sphere xmin=6.0 ymin=3.0 zmin=4.0 radius=1.6
 state 4 density=2.5 energy=4.0 yvel=0.3 geometry=point xmin=3.0 ymin=3.5 zmin=3.0
 x_cells={nx} y_cells={ny} z_cells={nz}
 xmin=0 ymin=0 zmin=0 xmax=9 ymax=7 zmax=6
 initial_timestep=0.02 end_step=5
 tiles_per_chunk {tiles}
*endclover"""

FIELDS = ["density0", "energy0", "xvel0", "yvel0", "zvel0", "volume", "xarea", "yarea", "zarea"]
HALO_UPDATED = {"density0", "energy0", "xvel0", "yvel0", "zvel0"}


def _f32(i: int) -> float:
    return float(np.float32(i))


def _oracle_tile(deck, dims, offs):
    L = ol.lib()
    nx, ny, nz = dims
    dx = (deck.xmax - deck.xmin) / _f32(deck.x_cells)
    dy = (deck.ymax - deck.ymin) / _f32(deck.y_cells)
    dz = (deck.zmax - deck.zmin) / _f32(deck.z_cells)
    x0 = deck.xmin + dx * _f32(offs[0] - 1)
    y0 = deck.ymin + dy * _f32(offs[1] - 1)
    z0 = deck.zmin + dz * _f32(offs[2] - 1)
    z = lambda *s: np.zeros(s, dtype=np.float64, order="F")  # noqa: E731
    v1 = [z(n + 5) for n in (nx, nx, ny, ny, nz, nz)]            # vertexx, vertexdx, ...
    c1 = [z(n + 4) for n in (nx, nx, ny, ny, nz, nz)]            # cellx, celldx, ...
    arr = dict(volume=z(nx + 4, ny + 4, nz + 4), xarea=z(nx + 5, ny + 4, nz + 4), yarea=z(nx + 4, ny + 5, nz + 4),
               zarea=z(nx + 4, ny + 4, nz + 5), density0=z(nx + 4, ny + 4, nz + 4),
               energy0=z(nx + 4, ny + 4, nz + 4), xvel0=z(nx + 5, ny + 5, nz + 5), yvel0=z(nx + 5, ny + 5, nz + 5),
               zvel0=z(nx + 5, ny + 5, nz + 5))
    P = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    dims6 = [1, nx, 1, ny, 1, nz]
    L.oracle_initialise_chunk_kernel(*dims6, ctypes.c_double(x0), ctypes.c_double(y0), ctypes.c_double(z0),
                                     ctypes.c_double(dx), ctypes.c_double(dy), ctypes.c_double(dz),
                                     *[P(a) for a in v1], *[P(a) for a in c1], P(arr["volume"]), P(arr["xarea"]),
                                     P(arr["yarea"]), P(arr["zarea"]))
    ns = len(deck.states)
    st = lambda name: (ctypes.c_double * ns)(*[getattr(s, name) for s in deck.states])  # noqa: E731
    geom = (ctypes.c_int * ns)(*[s.geometry for s in deck.states])
    L.oracle_generate_chunk_kernel(*dims6, P(v1[0]), P(v1[2]), P(v1[4]), P(c1[0]), P(c1[2]), P(c1[4]),
                                   P(arr["density0"]), P(arr["energy0"]), P(arr["xvel0"]), P(arr["yvel0"]),
                                   P(arr["zvel0"]), ns, st("density"), st("energy"), st("xvel"), st("yvel"),
                                   st("zvel"), st("xmin"), st("xmax"), st("ymin"), st("ymax"), st("zmin"),
                                   st("zmax"), st("radius"), geom)
    return arr


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("nx,ny,nz,tiles,bgvel", [(23, 17, 13, 1, ""), (23, 17, 13, 6, ""),
                                                  (18, 14, 12, 4, "xvel=0.05 yvel=-0.02 zvel=0.01"),
                                                  (9, 7, 6, 1, "zvel=0.3")])
def test_generated_fields_match_oracle(nx, ny, nz, tiles, bgvel, fp32):
    deck = parse_deck(DECK.format(nx=nx, ny=ny, nz=nz, tiles=tiles, bgvel=bgvel))
    print(f"[gpu] generate {nx}x{ny}x{nz} tiles={tiles} fp32={fp32}", flush=True)
    with Chunk(deck, fp32=fp32) as c:
        for t in range(c.ntiles):
            dims, offs, _ = c.tile_info(t)
            want = _oracle_tile(deck, dims, offs)
            for name in FIELDS:
                got = c.fetch(name, t)
                w = want[name].astype(np.float32).astype(np.float64) if fp32 else want[name]
                if name in HALO_UPDATED:
                    inner = tuple(slice(2, s - 2) for s in got.shape)  # min..max (cells), min..max+1 (vertices)
                    got, w = got[inner], w[inner]
                assert np.array_equal(got, w), (t, name)
