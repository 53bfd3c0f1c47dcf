"""clover.out from a device run: the default deck (test_problem 1)."""
from __future__ import annotations

import io
import re

import numpy as np
import pytest

from cloverleaf3d_ref_amd.chunk import Chunk
from cloverleaf3d_ref_amd.clover_out import run
from cloverleaf3d_ref_amd.deck import DEFAULT_DECK, parse_deck

pytestmark = pytest.mark.gpu

STEP = re.compile(r"^ Step +(\d+) time +([-0-9.]+) control +(\w+) +timestep +([-0-9.E+]+) +(\d+), +(\d+) x +"
                  r"([-0-9.E+]+) y +([-0-9.E+]+)$")


def test_default_deck_clover_out():
    deck = parse_deck(DEFAULT_DECK)
    print("[gpu] clover_out default deck", flush=True)
    buf = io.StringIO()
    steps = run(deck, buf)
    text = buf.getvalue().splitlines()
    assert steps == 75
    step_lines = [STEP.match(l) for l in text if l.startswith(" Step ")]
    assert all(step_lines) and [int(m.group(1)) for m in step_lines] == list(range(1, 76))
    with Chunk(deck) as c:
        dts = [c.step() for _ in range(75)]
    for m, dt in zip(step_lines, dts):
        assert m.group(3) == "sound"
        assert float(m.group(4)) == pytest.approx(dt, rel=6e-3)
        # calc_dt_kernel.f90:154-160 on the 10x2x2 deck: cell (1, 1), centres 0.5, 0.5
        assert (m.group(5), m.group(6), m.group(7), m.group(8)) == ("1", "1", "5.00E-01", "5.00E-01")
    times = np.cumsum([0.0] + dts[:-1])
    assert [float(m.group(2)) for m in step_lines] == pytest.approx(list(np.round(times, 7)), abs=1e-7)
    rows = [l for l in text if l.startswith(" step:")]
    assert len(rows) == 1 + 75 // deck.summary_frequency + 1  # start, every summary step, the final one
    with Chunk(deck) as c:
        c.run()
        final = c.field_summary()
    vals = [float(v) for v in rows[-1].split()[2:]]
    want = [final["vol"], final["mass"], final["mass"] / final["vol"], final["press"] / final["vol"], final["ie"],
            final["ke"], final["ie"] + final["ke"]]
    assert vals == pytest.approx(want, rel=1e-8)
    assert any("This test is considered PASSED" in l for l in text)


def test_small_timestep_clover_out(monkeypatch, tmp_path):
    """dt below dtmin at step 1: the step line is written, then
    report_error's lines (timestep.f90:112-122, report.f90:29-56), and the
    run ends there with no completion lines and a failing exit status.
    (The reference's decks cannot set dtmin, read_input.f90:63; the test
    raises it on the parsed deck.)"""
    from cloverleaf3d_ref_amd import clover_out
    from cloverleaf3d_ref_amd.chunk import SmallTimestep

    deck = parse_deck(DEFAULT_DECK)
    deck.dtmin = 1.0
    print("[gpu] clover_out small timestep", flush=True)
    buf = io.StringIO()
    with pytest.raises(SmallTimestep):
        run(deck, buf, autotune=False)
    lines = buf.getvalue().splitlines()
    steps = [l for l in lines if l.startswith(" Step ")]
    assert len(steps) == 1 and STEP.match(steps[0]).group(1) == "1"
    assert lines[lines.index(steps[0]) + 1:] == ["", " Error from timestep:", " small timestep", "",
                                                  " CLOVER is terminating.", ""]
    assert not any("Calculation complete" in l for l in lines)

    def small_deck(text):
        d = parse_deck(text)
        d.dtmin = 1.0
        return d

    monkeypatch.setattr(clover_out, "parse_deck", small_deck)
    deck_file = tmp_path / "clover.in"
    deck_file.write_text(DEFAULT_DECK)
    assert clover_out.main([str(deck_file), "-o", str(tmp_path / "clover.out"), "--no-autotune"]) == 1
