"""The ensemble driver against each deck run alone and against the oracle.

Every deck of an ensemble evolves exactly as a Chunk running it alone with
the fused kernels: after the same number of steps its summary totals, time
and dt are bit-identical, in FP64 and in the FP32-store form, whether the
steps run eagerly or as replays of the captured step graphs.  In FP64 each
deck's time and dt also equal the oracle's run of that deck bit for bit,
and its totals agree within 1e-12 (the oracle sums in another order).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import oracle_lib as ol
from cloverleaf3d_ref_amd.chunk import Chunk, SmallTimestep
from cloverleaf3d_ref_amd.deck import ensemble_decks, parse_deck
from cloverleaf3d_ref_amd.ensemble import Ensemble

pytestmark = pytest.mark.gpu

UNEVEN = """*clover
 state 1 density={rb} energy={eb}
 state 2 density={rc} energy={ec} xvel={u} geometry=cuboid xmin=1.5 xmax=4.0 ymin=2.0 ymax=6.5 zmin=0.5 zmax=3.0
 state 3 density=0.8 energy=2.0 zvel=-0.1 geometry=sphere xmin=6.0 ymin=3.0 zmin=4.0 radius=1.6
 x_cells={nx} y_cells={ny} z_cells={nz}
 xmin=0 ymin=0 zmin=0 xmax=9 ymax=7 zmax=6
 initial_timestep=0.02 timestep_rise=1.3 end_time=100 end_step=100 summary_frequency=0
*endclover"""


def _uneven():
    shapes = [(3, 3, 3), (23, 17, 13), (8, 31, 5), (70, 9, 11), (16, 16, 70), (5, 66, 4)]
    out = []
    for q, (nx, ny, nz) in enumerate(shapes):
        out.append(parse_deck(UNEVEN.format(nx=nx, ny=ny, nz=nz, rb=0.2 + 0.05 * q, eb=1.0 + 0.1 * q,
                                            rc=1.0 + 0.2 * q, ec=2.5 - 0.1 * q, u=0.05 * q)))
    return out


def _alone(deck, steps, fp32):
    with Chunk(deck, fp32=fp32) as c:
        for _ in range(steps):
            c.step()
        step, time, dt = c.state()
        s = c.field_summary()
    return s, time, dt


def _check(decks, steps, fp32, graphs, split=None):
    print(f"[gpu] ensemble of {len(decks)} decks, {steps} steps, fp32={fp32} graphs={graphs}", flush=True)
    with Ensemble(decks, fp32=fp32) as e:
        if split:
            e.run(split, graphs=graphs)
            e.run(steps - split, graphs=graphs)
        else:
            e.run(steps, graphs=graphs)
        sums = e.summaries()
        step, times, dts, small = e.state()
    assert step == steps and not any(small)
    for d, deck in enumerate(decks):
        want, time, dt = _alone(deck, steps, fp32)
        assert times[d] == time and dts[d] == dt, d
        for k in ("vol", "mass", "ie", "ke", "press"):
            assert sums[d][k] == want[k], (d, k, sums[d][k], want[k])
        if not fp32:
            rows, odts = ol.run(dataclasses.replace(deck, end_step=steps, summary_frequency=0))
            assert len(odts) == steps and dts[d] == odts[-1] and times[d] == rows[-1]["time"], d
            for k in ("vol", "mass", "ie", "ke", "press"):
                assert sums[d][k] == pytest.approx(rows[-1][k], rel=1e-12, abs=1e-300), (d, k)


@pytest.mark.parametrize("graphs", [False, True])
def test_uneven_decks_fp64(graphs):
    _check(_uneven(), 14, False, graphs, split=5)


def test_uneven_decks_fp32_graphs():
    _check(_uneven(), 14, True, True, split=9)


def test_generated_decks():
    decks = ensemble_decks(6, end_step=200, edges=(16, 24, 32))
    _check(decks, 12, False, True)


def test_graph_replays_equal_eager_fields():
    decks = _uneven()[:3]
    with Ensemble(decks) as a, Ensemble(decks) as b:
        a.run(9, graphs=True)
        b.run(9, graphs=False)
        for d in range(len(decks)):
            for name in ("density0", "energy0", "xvel0", "zvel0", "mass_flux_y"):
                assert np.array_equal(a.fetch(d, name), b.fetch(d, name)), (d, name)


def test_rejects_tiled_and_thin_decks():
    from cloverleaf3d_ref_amd._lib import CloverError
    tiled = parse_deck(UNEVEN.format(nx=8, ny=8, nz=8, rb=0.2, eb=1.0, rc=1.0, ec=2.5, u=0.0).replace(
        "*endclover", " tiles_per_chunk 2\n*endclover"))
    thin = parse_deck(UNEVEN.format(nx=8, ny=2, nz=8, rb=0.2, eb=1.0, rc=1.0, ec=2.5, u=0.0))
    for bad in (tiled, thin):
        with pytest.raises(CloverError):
            Ensemble([bad])


@pytest.mark.parametrize("wgs,xcd,fp32", [(3, 1, False), (8, 0, False), (13, 1, False), (200, 1, False),
                                          (13, 1, True)])
def test_persistent_tile_walks_keep_every_bit(wgs, xcd, fp32):
    """Short grids whose workgroups walk the tile list (fewer workgroups than
    XCDs, a stride that does not divide the tiles, one that does) give the
    fields of one workgroup per tile."""
    from cloverleaf3d_ref_amd._lib import SHAPE_KERNELS
    decks = _uneven()
    print(f"[gpu] persistent walk wgs={wgs} xcd={xcd} fp32={fp32}", flush=True)
    with Ensemble(decks, fp32=fp32) as a, Ensemble(decks, fp32=fp32) as b:
        for name in SHAPE_KERNELS:
            b.set_shape(name, 64 if name == "cell_x" else 16, xcd, wgs)
        assert set(b.workgroups().values()) == {wgs}
        a.run(9, graphs=True)
        b.run(4, graphs=False)
        b.run(5, graphs=True)
        for d in range(len(decks)):
            for name in ("density0", "energy0", "pressure", "xvel0", "yvel0", "zvel0", "viscosity"):
                assert np.array_equal(a.fetch(d, name), b.fetch(d, name)), (d, name)
        assert a.state()[2] == b.state()[2]


def test_set_shape_rejects_bad_values():
    from cloverleaf3d_ref_amd._lib import CloverError
    with Ensemble(_uneven()[:2]) as e:
        for args in (("visc_dt", 4, 0, 0), ("visc_dt", 16, 2, 0), ("mom_y", 16, 0, -1)):
            with pytest.raises(CloverError):
                e.set_shape(*args)


def test_small_timestep_deck_is_reported():
    """A deck whose dt falls below its dtmin: the reference stops there
    (timestep.f90:120-122) and Chunk.step raises; the ensemble keeps the
    batch going, records the step and raises after the run, and the other
    decks are untouched by it."""
    decks = _uneven()[:3]
    decks[1].dtmin = 1.0  # above every dt this deck can take: small at step 1
    with Chunk(decks[1]) as c:
        with pytest.raises(SmallTimestep):
            c.step()
    with Ensemble(decks) as e:
        with pytest.raises(SmallTimestep):
            e.run(3)
        step, times, dts, small = e.state()
        sums = e.summaries()
    assert step == 3 and small == [0, 1, 0]
    for d in (0, 2):
        want, time, dt = _alone(decks[d], 3, False)
        assert times[d] == time and dts[d] == dt
        for k in ("vol", "mass", "ie", "ke", "press"):
            assert sums[d][k] == want[k], (d, k)
