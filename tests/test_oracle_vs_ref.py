"""Pins the oracle (C restatement) against the reference's own kernels.

The reference's *_kernel subroutines are compiled from their own sources in
the reference tree into oracle/_ref/libclover_ref.so (oracle/Makefile).  Each
oracle kernel is called on the same seeded random inputs at small odd grid
shapes and must reproduce every output array bit for bit; the field_summary
sums (whose order depends on the OpenMP reduction) agree to 1e-13 relative.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import kernel_specs as ks
import oracle_lib as ol

pytestmark = pytest.mark.skipif(not ol.ref_available(), reason="oracle/_ref not built (needs the reference tree)")

G_SMALL = float(np.float32(1.0e-16))
G_BIG = float(np.float32(1.0e21))
CALC_DT_D = dict(g_small=G_SMALL, g_big=G_BIG, dtmin=1.0e-7, dtc_safe=0.75, dtu_safe=0.5, dtv_safe=0.5,
                 dtw_safe=0.5, dtdiv_safe=0.5)

SHAPES = [(7, 5, 6), (4, 9, 3), (12, 3, 5)]


def _fns(name):
    o = getattr(ol.lib(), "oracle_" + name)
    r = getattr(ol.ref_lib(), "ref_" + (name.lower() if name == "PdV_kernel" else name))
    o.restype = None
    r.restype = None
    return o, r


def _assert_bitwise(a: dict, b: dict, names, ctx):
    for n in names:
        x, y = a[n], b[n]
        if x.dtype == np.float64:
            same = np.array_equal(x.view(np.uint64), y.view(np.uint64))
        else:
            same = np.array_equal(x, y)
        if not same:
            diff = np.argwhere(x.view(np.uint64) != y.view(np.uint64)) if x.dtype == np.float64 else np.argwhere(x != y)
            raise AssertionError(f"{ctx}: {n} differs at {diff[:5].tolist()} (of {len(diff)})")


def _run_pair(name, dims, seed, ints=None, doubles=None, **kw):
    rng = np.random.default_rng(seed)
    vals = ks.make_args(name, dims, rng, ints or {}, doubles or {}, **kw)
    a = ks.copy_vals(vals)
    b = ks.copy_vals(vals)
    o, r = _fns(name)
    ks.call(o, name, a)
    ks.call(r, name, b)
    return a, b


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("name", ["ideal_gas_kernel", "viscosity_kernel", "revert_kernel", "reset_field_kernel"])
def test_pointwise_kernels(name, dims):
    a, b = _run_pair(name, dims, seed=hash((name, dims)) % 2**32)
    _assert_bitwise(a, b, ks.array_names(name), f"{name}{dims}")


@pytest.mark.parametrize("dims", SHAPES)
def test_calc_dt(dims):
    a, b = _run_pair("calc_dt_kernel", dims, seed=11, doubles=CALC_DT_D)
    _assert_bitwise(a, b, ks.array_names("calc_dt_kernel"), f"calc_dt{dims}")
    assert a["dt_min_val"][0] < G_BIG and a["small"][0] == 0


def test_calc_dt_small_flag():
    d = dict(CALC_DT_D, dtmin=1.0e6)
    a, b = _run_pair("calc_dt_kernel", (5, 4, 3), seed=3, doubles=d)
    _assert_bitwise(a, b, ks.array_names("calc_dt_kernel"), "calc_dt small")
    assert a["small"][0] == 1


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("predict", [0, 1])
def test_pdv(dims, predict):
    a, b = _run_pair("PdV_kernel", dims, seed=5 + predict, ints=dict(predict=predict), doubles=dict(dt=0.04))
    _assert_bitwise(a, b, ks.array_names("PdV_kernel"), f"PdV{dims} predict={predict}")


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("name", ["accelerate_kernel", "flux_calc_kernel"])
def test_dt_kernels(name, dims):
    a, b = _run_pair(name, dims, seed=17, doubles=dict(dt=0.037))
    _assert_bitwise(a, b, ks.array_names(name), f"{name}{dims}")


# (dir, sweep_number, advect_x) as advection.f90:40-149 issues them
ADVEC_CASES = [(1, 1, 1), (2, 2, 1), (3, 3, 1), (3, 1, 0), (2, 2, 0), (1, 3, 0)]


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("case", ADVEC_CASES)
def test_advec_cell(dims, case):
    d, s, ax = case
    a, b = _run_pair("advec_cell_kernel", dims, seed=23, ints=dict(dir=d, sweep_number=s, advect_x=ax))
    _assert_bitwise(a, b, ks.array_names("advec_cell_kernel"), f"advec_cell{dims}{case}")


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("case", ADVEC_CASES)
def test_advec_mom_three_calls(dims, case):
    """which_vel 1, 2, 3 in sequence sharing the work arrays, as
    advec_mom_driver.f90 does: the node fluxes and masses computed by the
    first call are reused by the next two."""
    d, s, ax = case
    name = "advec_mom_kernel"
    rng = np.random.default_rng(29)
    base = ks.make_args(name, dims, rng, dict(direction=d, sweep_number=s, advect_x=ax, which_vel=1), {})
    vels = [np.asfortranarray(ks.fill(rng, "vel", base["vel1"].shape)) for _ in range(3)]
    a = ks.copy_vals(base)
    b = ks.copy_vals(base)
    a_vels = [v.copy(order="F") for v in vels]
    b_vels = [v.copy(order="F") for v in vels]
    o, r = _fns(name)
    for w in (1, 2, 3):
        a["vel1"] = a_vels[w - 1]
        b["vel1"] = b_vels[w - 1]
        a["which_vel"] = b["which_vel"] = w
        ks.call(o, name, a)
        ks.call(r, name, b)
        _assert_bitwise(a, b, ks.array_names(name), f"advec_mom{dims}{case} which_vel={w}")


FIELD_SETS = {
    "start": [1, 3, 5, 6, 2, 4, 8, 10, 12, 9, 11, 13],
    "timestep": [5, 3, 1, 8, 10],
    "visc": [6],
    "cell": [4, 2, 14, 15, 16],
    "mom": [2, 4, 9, 11, 13, 17, 18, 19],
    "all": list(range(1, 20)),
}


@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("fset", list(FIELD_SETS))
@pytest.mark.parametrize("ext", ["all", "some"])
def test_update_halo(depth, fset, ext):
    fields = np.zeros(19, dtype=np.int32)
    for f in FIELD_SETS[fset]:
        fields[f - 1] = 1
    nb = {"chunk_neighbours": np.full(6, -1, np.int32), "tile_neighbours": np.full(6, -1, np.int32)}
    if ext == "some":
        nb["tile_neighbours"] = np.array([-1, 2, 3, -1, -1, 7], np.int32)
    a, b = _run_pair("update_halo_kernel", (7, 5, 6), seed=31 + depth, fields=fields, neighbours=nb,
                     ints=dict(depth=depth))
    _assert_bitwise(a, b, ks.array_names("update_halo_kernel"), f"halo {fset} d{depth} {ext}")


@pytest.mark.parametrize("face", list(ks.FACE_ID))
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("fset", ["all", "mom", "cell", "start"])
def test_update_tile_halo(face, depth, fset):
    fields = np.zeros(19, dtype=np.int32)
    for f in FIELD_SETS[fset]:
        fields[f - 1] = 1
    dims = (6, 5, 7)
    # neighbours share the extents across the face (clover.f90:413-426)
    ndims = {"left": (4, 5, 7), "right": (3, 5, 7), "bottom": (6, 4, 7), "top": (6, 3, 7),
             "back": (6, 5, 4), "front": (6, 5, 3)}[face]
    name = f"update_tile_halo_{face}_kernel"
    a, b = _run_pair(name, dims, seed=37, ndims=ndims, fields=fields, ints=dict(depth=depth))
    _assert_bitwise(a, b, ks.array_names(name), f"{name} d{depth} {fset}")


@pytest.mark.parametrize("dims", SHAPES)
def test_field_summary(dims):
    a, b = _run_pair("field_summary_kernel", dims, seed=41)
    for n in ("vol", "mass", "ie", "ke", "press"):
        assert a[n][0] == pytest.approx(b[n][0], rel=1e-13, abs=0), n
