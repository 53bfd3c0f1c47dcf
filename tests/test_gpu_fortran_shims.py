"""Whole runs through the Fortran side of the drop-in boundary.

oracle/_fshim/libfshim.so holds the product's Fortran shim modules
(cloverleaf3d_ref_amd/fortran/clover_hip_shims.f90: the reference's module
and subroutine names, e.g. PdV_kernel_module / PdV_kernel) behind the same
C-callable wrappers the oracle puts around the reference's own kernels.
The oracle driver runs start and the hydro loop through them in the
reference's call order, so every shim runs as the reference's drivers would
call it: PdV_kernel with .TRUE. and .FALSE., advec_cell_kernel and
advec_mom_kernel with advect_x alternating, update_halo_kernel with its
field flags and depths, the six tile-halo kernels on tiled decks, calc_dt's
and field_summary's outputs by reference, arrays passed as the base of the
caller's storage (lower bound min-2, build_field.f90:41-147).  The totals
must match the oracle's own run and the dt sequence must be bit-identical.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as ol
from cloverleaf3d_ref_amd import _lib
from cloverleaf3d_ref_amd.deck import DEFAULT_DECK, cuboid_deck, parse_deck

pytestmark = pytest.mark.gpu

FSHIM = ol.ROOT / "oracle" / "_fshim" / "libfshim.so"

DECKS = {
    "test_problem_1": DEFAULT_DECK,
    "cuboid_20": cuboid_deck(20, 16, 4),
    "cuboid_18_tiles8": cuboid_deck(18, 10, 5, tiles=8),
}


@pytest.mark.parametrize("name", list(DECKS))
def test_whole_run_through_fortran_shims(name):
    if not FSHIM.exists():
        pytest.fail(f"{FSHIM} not built (make -C oracle fshim, needs flang)")
    deck = parse_deck(DECKS[name])
    L = _lib.lib()
    _lib.check(L.clover_hip_init(0), "clover_hip_init")
    try:
        print(f"[gpu] run through the Fortran shims: {name}", flush=True)
        got, got_dt = ol.run_through(deck, FSHIM, "ref_")
    finally:
        _lib.check(L.clover_hip_finalize(), "clover_hip_finalize")
    want, want_dt = ol.run(deck)
    assert len(got_dt) == deck.end_step
    assert np.array_equal(got_dt.view(np.uint64), want_dt.view(np.uint64))
    assert len(got) == len(want)
    for g, w in zip(got, want):
        for k in ("vol", "mass", "ie", "ke", "press"):
            assert g[k] == pytest.approx(w[k], rel=1e-12, abs=1e-300), (name, g["step"], k)
    if name == "test_problem_1":
        from cloverleaf3d_ref_amd.deck import TEST_PROBLEM_1_KE
        assert abs(got[-1]["ke"] / TEST_PROBLEM_1_KE - 1) < 1e-12
