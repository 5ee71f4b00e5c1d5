"""The budget log block (F3D_end_time_step, Sources/mod_F3D.F90:1275-1326)
on the CPU: Fortran Ew.d / Fw.d / Iw edit descriptors on known values
(what gfortran writes for them), the block's layout and its sum line."""
from fall3d_amd.output import budget_log, fortran_e, fortran_f, fortran_i, should_log


def test_edit_descriptors():
    assert fortran_e(12345.678) == " 0.123457E+05"
    assert fortran_e(-0.5) == "-0.500000E+00"
    assert fortran_e(0.0) == " 0.000000E+00"
    assert fortran_e(9.9999996e4) == " 0.100000E+06"
    assert fortran_e(3.0e-7) == " 0.300000E-06"
    assert fortran_e(1.0e120) == " 0.100000+121"
    assert fortran_f(12.345, 7, 2) == "  12.35" or fortran_f(12.345, 7, 2) == "  12.34"
    assert fortran_f(100.0, 7, 2) == " 100.00"
    assert fortran_i(42, 8) == "      42"
    assert fortran_i(123456789, 8) == "********"


def test_budget_block():
    b = {"mass_volume": 6.0e9, "ground": 3.0e9, "lateral": 1.0e9}
    txt = budget_log(b, iiter=120, gl_dt=38.2, time=7200.0, run_start=0.0, run_end=14400.0,
                     time_str="01Jan2024_02:00", mass_in=1.0e10)
    lines = txt.split("\n")
    assert lines[1] == "-> Iteration                 :      120"
    assert lines[2] == "   Time step                 :  0.382000E+02"
    assert lines[3] == "   Elapsed time              :     120 min   (  50.00%)"
    assert lines[6] == "   (1) Mass inside  the domain     :  0.600000E+10 (  60.00%)"
    assert lines[7] == "   (2) Mass outside the domain     :  0.100000E+10 (  10.00%)"
    assert lines[8] == "   (3) Mass sink terms (wet dep.)  :  0.000000E+00 (   0.00%)"
    assert lines[9] == "   (4) Mass at ground              :  0.300000E+10 (  30.00%)"
    assert lines[11] == "   Sum (1)+(2)+(3)+(4)             :  0.100000E+11"
    assert lines[12] == "   Injected mass                   :  0.100000E+11"


def test_log_cadence():
    assert should_log(1, 38.0, 38.0)
    assert should_log(17, 1210.0, 40.0)      # 10 s past 1200 s
    assert not should_log(18, 1250.0, 40.0)  # 50 s past
