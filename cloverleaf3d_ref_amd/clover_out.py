"""clover.out for the device-resident driver: the reference's per-step and
summary lines, in its own formats, so a run here can be diffed against a
reference run line by line.

    python -m cloverleaf3d_ref_amd.clover_out clover.in [-o clover.out] [--fp32] [--per-call] [--visit-dir D]

Lines with explicit Fortran formats are reproduced character for character:
  * the step line, timestep.f90:115-116
    (' Step ', i7,' time ', f11.7,' control ',a11,' timestep  ',1pe9.2,i8,',',i8,' x ',1pe9.2,' y ',1pe9.2);
  * the summary header and row, field_summary.f90:46-47 and 111
    ('(a13,7a16)' and '(a6,i7,7e16.8)').
List-directed lines (WRITE(g_out,*): 'Time', 'Wall clock', the test-problem
verdict) have a compiler-chosen layout; they are written with the same
words and full-precision values.
"""
from __future__ import annotations

import argparse
import sys
import time as _time
from typing import List, TextIO

import numpy as np

from .chunk import Chunk, SmallTimestep
from .deck import TEST_PROBLEM_1_KE, Deck, parse_deck

_CONTROLS = {1: "sound", 2: "xvel", 3: "yvel", 4: "zvel", 5: "div"}  # calc_dt.f90
G_SMALL = float(np.float32(1.0e-16))  # data.f90:30, a default-real literal widened to double


def fmt_i(v: int, w: int) -> str:
    """Fortran Iw."""
    s = str(int(v))
    return s.rjust(w) if len(s) <= w else "*" * w


def fmt_a(s: str, w: int, length: int) -> str:
    """Fortran Aw of a CHARACTER(LEN=length) value: blank-padded to its
    length, then right-justified in w (or its first w characters)."""
    v = s.ljust(length)[:length]
    return v.rjust(w) if w >= length else v[:w]


def fmt_f(x: float, w: int, d: int) -> str:
    """Fortran Fw.d."""
    s = f"{x:.{d}f}"
    if s.startswith("-0.") and float(s) == 0.0:
        s = s[1:]
    if len(s) > w and s.startswith("0."):
        s = s[1:]
    if len(s) > w and s.startswith("-0."):
        s = "-" + s[2:]
    return s.rjust(w) if len(s) <= w else "*" * w


def fmt_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d without a scale factor: 0.d1..dd E+xx, or 0.d1..dd +xxx
    when the exponent needs three digits."""
    if x == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")  # d significant digits, correctly rounded
        digits = m.replace(".", "")
        exp = int(e) + 1
        sign = "+" if exp >= 0 else "-"
        if abs(exp) > 999:
            return "*" * w
        tail = ("E" + sign + f"{abs(exp):02d}") if abs(exp) <= 99 else (sign + f"{abs(exp):03d}")
        s = ("-" if x < 0 else "") + "0." + digits + tail
    return s.rjust(w) if len(s) <= w else "*" * w


def fmt_1pe(x: float, w: int, d: int) -> str:
    """Fortran 1PEw.d: d.dd E+xx."""
    s = f"{x:.{d}E}"
    m, e = s.split("E")
    s = m + "E" + e[0] + e[1:].zfill(2)
    return s.rjust(w) if len(s) <= w else "*" * w


def step_line(step: int, time: float, control: int, dt: float, jdt: int, kdt: int, x: float, y: float) -> str:
    """timestep.f90:115-116."""
    return (" Step " + fmt_i(step, 7) + " time " + fmt_f(time, 11, 7) + " control "
            + fmt_a(_CONTROLS.get(control, ""), 11, 8) + " timestep  " + fmt_1pe(dt, 9, 2) + fmt_i(jdt, 8)
            + "," + fmt_i(kdt, 8) + " x " + fmt_1pe(x, 9, 2) + " y " + fmt_1pe(y, 9, 2))


SUMMARY_HEADER = (" " * 13 + "".join(h.rjust(16) for h in
                  ("Volume", "Mass", "Density", "Pressure", "Internal Energy", "Kinetic Energy", "Total Energy")))


def summary_lines(step: int, time: float, s: dict) -> List[str]:
    """field_summary.f90:45-47 and 111-112."""
    vals = [s["vol"], s["mass"], s["mass"] / s["vol"], s["press"] / s["vol"], s["ie"], s["ke"], s["ie"] + s["ke"]]
    return ["", f" Time {time!r}", SUMMARY_HEADER,
            " step:" + fmt_i(step, 7) + "".join(fmt_e(v, 16, 8) for v in vals), ""]


def test_problem_lines(deck: Deck, ke: float) -> List[str]:
    """field_summary.f90:117-133."""
    if deck.test_problem != 1:
        return []
    qa_diff = abs(100.0 * (ke / TEST_PROBLEM_1_KE) - 100.0)
    verdict = "This test is considered PASSED" if qa_diff < 0.001 else "This is test is considered NOT PASSED"
    return [f" Test problem 1 is within {qa_diff!r} % of the expected solution", " " + verdict]


def error_lines(location: str, error: str) -> List[str]:
    """report.f90:29-56's lines in clover.out: list-directed records, so
    each starts with a blank, and empty WRITEs give empty lines."""
    return ["", f" Error from {location}:", f" {error}", "", " CLOVER is terminating.", ""]


def run(deck: Deck, out: TextIO, fp32: bool = False, per_call: bool = False, autotune: bool = True,
        visit_dir: str = ".") -> int:
    """The hydro loop of hydro.f90:46-95 on the device, writing clover.out's
    step and summary lines, and, when the deck sets visit_frequency, the
    VTK dumps of visit.f90 into visit_dir.  Returns the steps taken.  On a
    small time step it writes that step's line and report_error's lines, as
    timestep.f90:112-122 does, and raises SmallTimestep (the reference
    aborts there: no completion lines)."""
    from .visit import VisitWriter

    visit = VisitWriter(deck, visit_dir) if deck.visit_frequency else None
    with Chunk(deck, per_call=per_call, fp32=fp32) as c:
        if autotune and not per_call:
            c.autotune()
        w = out.write
        # start.f90:136 prints the start state's summary
        step, t, _ = c.state()
        for line in summary_lines(step, t, c.field_summary()):
            w(line + "\n")
        if visit:  # start.f90:138
            visit.dump(c)
        w(" Starting the calculation\n")
        t0 = _time.perf_counter()
        while True:
            step, t_before, _ = c.state()
            try:
                dt = c.step()
            except SmallTimestep as err:
                step, _, _ = c.state()
                control, jdt, kdt, x, y = c.dt_control()
                w(step_line(step, t_before, control, err.dt, jdt, kdt, x, y) + "\n")
                for line in error_lines("timestep", "small timestep"):
                    w(line + "\n")
                raise
            step, t, _ = c.state()
            control, jdt, kdt, x, y = c.dt_control()
            w(step_line(step, t_before, control, dt, jdt, kdt, x, y) + "\n")
            # hydro.f90:70-72, then 84-86: a last step that is also a summary
            # step is summarised twice, as the reference does
            if deck.summary_frequency and step % deck.summary_frequency == 0:
                for line in summary_lines(step, t, c.field_summary()):
                    w(line + "\n")
            if visit and step % deck.visit_frequency == 0:  # hydro.f90:73-75
                visit.dump(c)
            if t + G_SMALL > deck.end_time or step >= deck.end_step:
                s = c.field_summary()
                for line in summary_lines(step, t, s):
                    w(line + "\n")
                for line in test_problem_lines(deck, s["ke"]):
                    w(line + "\n")
                if visit:  # hydro.f90:88
                    visit.dump(c)
                break
        wall = _time.perf_counter() - t0
        w("\n Calculation complete\n Clover is finishing\n")
        w(f" Wall clock {wall!r}\n")
        return step


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m cloverleaf3d_ref_amd.clover_out")
    ap.add_argument("deck", help="an input deck in the reference's clover.in format")
    ap.add_argument("-o", "--out", default="-", help="output file (default: stdout)")
    ap.add_argument("--fp32", action="store_true", help="the FP32-store form")
    ap.add_argument("--per-call", action="store_true", help="one kernel per reference loop nest")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--visit-dir", default=".", help="where the VTK dumps go when the deck sets visit_frequency")
    a = ap.parse_args(argv)
    deck = parse_deck(open(a.deck).read())
    out = sys.stdout if a.out == "-" else open(a.out, "w")
    try:
        run(deck, out, fp32=a.fp32, per_call=a.per_call, autotune=not a.no_autotune, visit_dir=a.visit_dir)
    except SmallTimestep:
        return 1  # the reference aborts (report.f90:53, clover_abort)
    finally:
        if out is not sys.stdout:
            out.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
