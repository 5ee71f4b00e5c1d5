"""Host-side output helpers of the ADS step.

budget_log: the mass-balance block F3D_end_time_step writes to the log file
every 10 simulated minutes (Sources/mod_F3D.F90:1275-1326, formats 10 and
20), from a context's FP64 budget (fall3d_ads_get_budget). The device
computes every figure; this module only formats them as the reference's
Fortran edit descriptors do (Ew.d, Fw.d, Iw).

result_variables: the post-processed result fields of nc_IO_out_res for one
species (fall3d_ads_postprocess), keyed by the reference's variable names;
cut_variables: its concentration cuts (fall3d_ads_postprocess_cuts), named
as Sources/mod_nc_IO_names.F90:43-46 names them.
"""
from __future__ import annotations

import math

__all__ = ["fortran_e", "fortran_f", "fortran_i", "budget_log", "should_log", "result_variables", "cut_variables"]


def fortran_e(x: float, w: int = 13, d: int = 6) -> str:
    """Fortran Ew.d: 0.ddddddE+xx, right-justified in w columns (a width too
    small for the value gives w asterisks, as Fortran does)."""
    if x != x:
        s = "NaN"
    elif math.isinf(x):
        s = "Infinity" if x > 0 else "-Infinity"
    elif x == 0.0:
        s = ("-" if math.copysign(1.0, x) < 0 else "") + "0." + "0" * d + "E+00"
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")
        digits = m.replace(".", "")
        ex = int(e) + 1  # 0.d1d2... x 10^(e+1)
        mant = "0." + digits
        if abs(ex) < 100:
            es = f"E{'+' if ex >= 0 else '-'}{abs(ex):02d}"
        else:
            es = f"{'+' if ex >= 0 else '-'}{abs(ex):03d}"
        s = ("-" if x < 0 else "") + mant + es
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_f(x: float, w: int, d: int) -> str:
    s = f"{x:.{d}f}"
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_i(n: int, w: int) -> str:
    s = str(int(n))
    return s.rjust(w) if len(s) <= w else "*" * w


def should_log(iiter: int, time: float, gl_dt: float) -> bool:
    """The reference logs at the first iteration and whenever the time is
    within one step past a multiple of 600 s (mod_F3D.F90:1275)."""
    return iiter == 1 or math.fmod(time, 600.0) <= gl_dt


def budget_log(budget: dict, iiter: int, gl_dt: float, time: float, run_start: float, run_end: float,
               time_str: str, mass_in: float, sink: float = 0.0, restart_terms=(0.0, 0.0, 0.0)) -> str:
    """The log block of one logged step (formats 10 and 20).

    budget: ADS.budget() of the context (mass_volume, ground, lateral);
    mass_in: the injected mass the caller tracks (MY_TRA%gl_mass_in);
    sink: the wet-deposition sink (0: wet deposition does not run here);
    restart_terms: the (sink, ground, lateral) masses a restarted run carries
    (MY_TRA%rst_mass_sink, rst_mass_ground, rst_mass_lateral)."""
    vol = float(budget["mass_volume"])
    snk = float(sink) + restart_terms[0]
    gnd = float(budget["ground"]) + restart_terms[1]
    lat = float(budget["lateral"]) + restart_terms[2]
    tot = gnd + lat + snk + vol
    pct = lambda v: 100.0 * v / tot if tot != 0.0 else float("nan")
    lines = [
        "",
        "-> Iteration                 : " + fortran_i(iiter, 8),
        "   Time step                 : " + fortran_e(gl_dt),
        "   Elapsed time              : " + fortran_i(int((time - run_start) / 60.0), 7) + " min   (" +
        fortran_f(100.0 * ((time - run_start) / (run_end - run_start)), 7, 2) + "%)",
        "   Current time              : " + time_str,
        "   -------                     ",
        "   (1) Mass inside  the domain     : " + fortran_e(vol) + " (" + fortran_f(pct(vol), 7, 2) + "%)",
        "   (2) Mass outside the domain     : " + fortran_e(lat) + " (" + fortran_f(pct(lat), 7, 2) + "%)",
        "   (3) Mass sink terms (wet dep.)  : " + fortran_e(snk) + " (" + fortran_f(pct(snk), 7, 2) + "%)",
        "   (4) Mass at ground              : " + fortran_e(gnd) + " (" + fortran_f(pct(gnd), 7, 2) + "%)",
        "       ----------------------------  ",
        "   Sum (1)+(2)+(3)+(4)             : " + fortran_e(tot),
        "   Injected mass                   : " + fortran_e(float(mass_in)),
    ]
    return "\n".join(lines)


def result_variables(ads_ctx, species_name: str = "tephra", **kw) -> dict:
    """nc_IO_out_res's result fields of one species from the device state,
    named as in the reference's output file (<species>_con, _con_bin, _sur,
    _col_mass, _col_mass_pm, _cloud_top, _grn_load, _grn_load_bin); arrays
    at the cell corners, x fastest. Keyword arguments go to
    ADS.postprocess."""
    out = ads_ctx.postprocess(**kw)
    names = {"con": "con", "con_bins": "con_bin", "con_sur": "sur", "col_mass": "col_mass",
             "col_mass_pm": "col_mass_pm", "cloud_top": "cloud_top", "grn_load": "grn_load",
             "grn_load_bin": "grn_load_bin"}
    return {f"{species_name}_{names[k]}": v for k, v in out.items()}


def cut_variables(ads_ctx, z_c, sigma, X3max, species_name: str = "tephra", **kw) -> dict:
    """nc_IO_out_res's concentration cuts of one species from the device
    state: <species>_con_yz (x cuts), _con_xz (y cuts), _con_xy (z cuts) and
    _fl (flight-level cuts), g/m3. Keyword arguments go to
    ADS.postprocess_cuts."""
    out = ads_ctx.postprocess_cuts(z_c, sigma, X3max, **kw)
    names = {"cut_x": "con_yz", "cut_y": "con_xz", "cut_z": "con_xy", "cut_fl": "fl"}
    return {f"{species_name}_{names[k]}": v for k, v in out.items()}
