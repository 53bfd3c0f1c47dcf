"""clover.out formatting: the Fortran edit descriptors of the reference's
step and summary lines (timestep.f90:115-116, field_summary.f90:46-47, 111),
checked against values worked by hand from the descriptor rules."""
from __future__ import annotations

from cloverleaf3d_ref_amd.clover_out import (SUMMARY_HEADER, fmt_1pe, fmt_a, fmt_e, fmt_f, fmt_i, step_line,
                                             summary_lines)


def test_e16_8():
    # Ew.d: optional sign, "0.", d digits, E, signed 2-digit exponent, right-justified
    assert fmt_e(1000.0, 16, 8) == "  0.10000000E+04"
    assert fmt_e(0.09897949218749998, 16, 8) == "  0.98979492E-01"
    assert fmt_e(-2.5e-7, 16, 8) == " -0.25000000E-06"
    assert fmt_e(0.0, 16, 8) == "  0.00000000E+00"
    assert fmt_e(9.999999999, 16, 8) == "  0.10000000E+02"  # rounding carries into the exponent
    assert fmt_e(247.44873046875, 16, 8) == "  0.24744873E+03"


def test_1pe9_2_f11_7_i_a():
    assert fmt_1pe(0.04, 9, 2) == " 4.00E-02"
    assert fmt_1pe(0.5, 9, 2) == " 5.00E-01"
    assert fmt_1pe(-1.234e-10, 9, 2) == "-1.23E-10"
    assert fmt_f(0.0, 11, 7) == "  0.0000000"
    assert fmt_f(12.3456789, 11, 7) == " 12.3456789"
    assert fmt_f(1234.5, 11, 7) == "*" * 11  # does not fit
    assert fmt_i(75, 7) == "     75"
    assert fmt_i(12345678, 7) == "*******"
    assert fmt_a("sound", 11, 8) == "   sound   "


def test_step_line_layout():
    line = step_line(1, 0.0, 1, 0.04, 1, 1, 0.5, 0.5)
    assert line == (" Step       1 time   0.0000000 control    sound    timestep   4.00E-02       1,       1"
                    " x  5.00E-01 y  5.00E-01")
    # ' Step ' i7 ' time ' f11.7 ' control ' a11 ' timestep  ' 1pe9.2 i8 ',' i8 ' x ' 1pe9.2 ' y ' 1pe9.2
    assert len(line) == 6 + 7 + 6 + 11 + 9 + 11 + 11 + 9 + 8 + 1 + 8 + 3 + 9 + 3 + 9


def test_summary_lines():
    assert SUMMARY_HEADER == " " * 13 + "          Volume            Mass         Density        Pressure" \
                                        " Internal Energy  Kinetic Energy    Total Energy"
    lines = summary_lines(0, 0.0, dict(vol=1000.0, mass=216.50390625, press=247.44873046875,
                                       ie=247.44873046875, ke=0.0))
    assert lines[3] == (" step:      0  0.10000000E+04  0.21650391E+03  0.21650391E+00  0.24744873E+00"
                        "  0.24744873E+03  0.00000000E+00  0.24744873E+03")
