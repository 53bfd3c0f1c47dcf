"""Launch shapes and the autotuner of the fused kernels.

A shape (positions per workgroup along a kernel's march, block order)
changes only how the work is split: every shape must give the bits of the
default one, and tuning between steps must leave every field as it was."""
from __future__ import annotations

import numpy as np
import pytest

from cloverleaf3d_ref_amd import _lib
from cloverleaf3d_ref_amd.chunk import FIELD_NAMES, Chunk
from cloverleaf3d_ref_amd.deck import ensemble_decks, parse_deck
from cloverleaf3d_ref_amd.ensemble import Ensemble
from test_gpu_run import ODD_DECK

pytestmark = pytest.mark.gpu

# fields a step leaves defined (level 1 holds the previous level 0 in the
# fused path; pressure/soundspeed are compared after a summary)
STATE = ["density0", "energy0", "pressure", "viscosity", "soundspeed", "xvel0", "yvel0", "zvel0",
         "vol_flux_x", "vol_flux_y", "vol_flux_z", "mass_flux_x", "mass_flux_y", "mass_flux_z"]


def _run(deck, steps, fp32=False, shapes=None, tune_at=None):
    print(f"[gpu] shapes {deck.x_cells}x{deck.y_cells}x{deck.z_cells} tiles={deck.tiles_per_chunk} "
          f"fp32={fp32} shapes={shapes} tune_at={tune_at}", flush=True)
    with Chunk(deck, fp32=fp32) as c:
        for k, shape in (shapes or {}).items():
            c.set_shape(k, *shape)
        dts = []
        for s in range(steps):
            if s == tune_at:
                before = {(t, n): c.fetch(n, t) for t in range(c.ntiles) for n in FIELD_NAMES}
                c.autotune()
                after = {(t, n): c.fetch(n, t) for t in range(c.ntiles) for n in FIELD_NAMES}
                for key in before:
                    assert np.array_equal(before[key].view(np.uint64), after[key].view(np.uint64)), key
            dts.append(c.step())
        rows = c.field_summary()
        fields = {(t, n): c.fetch(n, t) for t in range(c.ntiles) for n in STATE}
        chosen = c.shapes()
    return np.array(dts), rows, fields, chosen


def _same(a, b):
    (d0, r0, f0, _), (d1, r1, f1, _) = a, b
    assert np.array_equal(d0, d1)
    assert r0 == r1
    for key in f0:
        inner = tuple(slice(2, s - 2) for s in f0[key].shape)
        assert np.array_equal(f0[key][inner].view(np.uint64), f1[key][inner].view(np.uint64)), key


ODD_SHAPES = {"visc_dt": (3, 1, 512), "pdv_predict": (1, 0, 128), "accel_pdv": (7, 1), "cell_x": (96, 1),
              "cell_y": (5, 1), "cell_z": (2, 0), "mom_x": (1, 1), "mom_y": (9, 0), "mom_z": (3, 1)}


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("nx,ny,nz,tiles", [(23, 17, 13, 1), (29, 21, 17, 4)])
def test_shapes_give_the_same_bits(nx, ny, nz, tiles, fp32):
    deck = parse_deck(ODD_DECK.format(nx=nx, ny=ny, nz=nz, steps=10, tiles=tiles, td="3d"))
    _same(_run(deck, 10, fp32), _run(deck, 10, fp32, shapes=ODD_SHAPES))


@pytest.mark.parametrize("tiles", [1, 4])
def test_autotune_keeps_every_field(tiles):
    deck = parse_deck(ODD_DECK.format(nx=40, ny=33, nz=70, steps=10, tiles=tiles, td="3d"))
    base = _run(deck, 10)
    tuned = _run(deck, 10, tune_at=5)
    _same(base, tuned)
    for name, (planes, xcd, bx) in tuned[3].items():
        assert planes in (8, 16, 32, 64, 128) and xcd in (0, 1) and bx in (64, 128, 256, 512), name


def test_shape_arguments_are_checked():
    deck = parse_deck(ODD_DECK.format(nx=8, ny=8, nz=8, steps=1, tiles=1, td="3d"))
    with Chunk(deck) as c:
        L = _lib.lib()
        assert L.clover_chunk_set_shape(c._h, len(_lib.SHAPE_KERNELS), 64, 0, 64) != 0
        assert L.clover_chunk_set_shape(c._h, 0, 0, 0, 64) != 0
        assert L.clover_chunk_set_shape(c._h, 0, 64, 2, 64) != 0
        assert L.clover_chunk_set_shape(c._h, 0, 64, 0, 96) != 0
        assert L.clover_chunk_set_shape(c._h, _lib.SHAPE_KERNELS.index("mom_y"), 64, 0, 256) != 0
        assert L.clover_chunk_set_shape(c._h, 0, 64, 0, 512) == 0
    with Chunk(deck, per_call=True) as c:
        with pytest.raises(_lib.CloverError):
            c.autotune()


@pytest.mark.parametrize("fp32", [False, True])
def test_ensemble_autotune_keeps_results(fp32):
    decks = ensemble_decks(12, seed=7, edges=(16, 24, 40))
    print(f"[gpu] ensemble autotune fp32={fp32}", flush=True)
    with Ensemble(decks, fp32=fp32) as e:
        e.run(8)
        base = e.summaries()
        base_state = e.state()
    with Ensemble(decks, fp32=fp32) as e:
        e.run(4)
        e.autotune()
        e.run(4)
        tuned = e.summaries()
        tuned_state = e.state()
        shapes = e.shapes()
    assert base == tuned
    assert base_state == tuned_state
    for name, (planes, xcd, bx) in shapes.items():
        assert planes in (8, 16, 32, 64, 128) and xcd in (0, 1) and bx == 64, name
