"""The batched workload's deck generator: seeded, reproducible, and drawing
from the stated distributions."""
from __future__ import annotations

import numpy as np

from cloverleaf3d_ref_amd.deck import ENSEMBLE_EDGES, SplitMix64, ensemble_decks


def test_splitmix64_known_values():
    # first outputs for seed 0 of the published splitmix64 generator
    r = SplitMix64(0)
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_generator_is_reproducible_and_in_range():
    a = ensemble_decks(64)
    b = ensemble_decks(64)
    assert [(d.x_cells, d.y_cells, d.z_cells) for d in a] == [(d.x_cells, d.y_cells, d.z_cells) for d in b]
    cells = np.array([[d.x_cells, d.y_cells, d.z_cells] for d in a])
    assert set(cells.ravel()) <= set(ENSEMBLE_EDGES)
    for d in a:
        bg, box = d.states[0], d.states[1]
        assert 0.1 <= bg.density < 0.5 and 0.5 <= bg.energy < 1.5
        assert 0.5 <= box.density < 2.0 and 1.5 <= box.energy < 5.0
        for lo, hi in ((box.xmin, box.xmax), (box.ymin, box.ymax), (box.zmin, box.zmax)):
            # read_input nudges the extents inwards by a hundredth of a cell
            assert 0.0 <= lo < hi <= 10.0 and 0.9 < hi - lo < 4.1
        assert d.tiles_per_chunk == 1 and d.end_step == 200


def test_full_ensemble_size():
    cells = np.array([[d.x_cells, d.y_cells, d.z_cells] for d in ensemble_decks(512)])
    assert len(cells) == 512
    assert abs(cells.mean() - 56.0) < 2.0  # E[edge] = 56
    assert len(set(cells.ravel())) == len(ENSEMBLE_EDGES)
