"""The resident group from Fortran (fall3d_amd/fortran/fall3d_ads_c.F90,
driven by resident_driver.F90): the ADS part of FALL3D's time loop with the
state, meteo and deposit kept on the GPU, called through ISO_C_BINDING, and
a restart snapshot restored into a second context. Compared with the FP64
oracle run through the same sequence (source, x/y sweeps in iiter-parity
order, z sweep): the FP64 state bit for bit, the FP32 state to 1e-5; the
restored run equals the unbroken one bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from fall3d_amd.problem import random_problem
from oracle.ads_oracle import Oracle
from test_fortran_dropin import _write_input

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "fall3d_amd", "lib", "fall3d_resident_driver")


def _read(path, p):
    raw = np.fromfile(path, dtype=np.float64)
    nc = (p.nx + 4) * (p.ny + 4) * (p.nz + 4) * p.nb
    na = (p.nx + 4) * (p.ny + 4) * p.nb
    shape_c, shape_a = (p.nx + 4, p.ny + 4, p.nz + 4, p.nb), (p.nx + 4, p.ny + 4, p.nb)
    o = 0
    c = raw[o:o + nc].reshape(shape_c, order="F"); o += nc
    a = raw[o:o + na].reshape(shape_a, order="F"); o += na
    flux = raw[o:o + 6]; o += 6
    budget = raw[o:o + 12]; o += 12
    c2 = raw[o:o + nc].reshape(shape_c, order="F"); o += nc
    a2 = raw[o:o + na].reshape(shape_a, order="F"); o += na
    assert o == raw.size
    return c, a, flux, budget, c2, a2


@pytest.mark.gpu
@pytest.mark.parametrize("state", [0, 2], ids=["fp32", "fp64"])
def test_fortran_resident_time_loop(tmp_path, state):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran resident driver not built (python -m fall3d_amd.build)")
    p = random_problem(14, 11, 9, 2, seed=12, umax_scale=0.4)
    nsteps, stime = 6, 0.35
    o = Oracle(p, "minmod", "rk")
    dt = o.time_step(p.u1, p.v1, p.w1, p.vs, cfl=0.6)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, p, 1, 2, nsteps, state, dt, stime, p.u1, p.v1)
    print(f"case resident state={state}: running {DRIVER}", flush=True)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    cg, ag, fg, bg, c2, a2 = _read(fout, p)
    # the driver uploads the steady winds as both snapshots, so the sweeps
    # read (1 - stime) u + stime u, which can differ from u in the last bit
    ul = np.asfortranarray((1.0 - stime) * p.u1 + stime * p.u1)
    vl = np.asfortranarray((1.0 - stime) * p.v1 + stime * p.v1)
    cr, ar = p.new_c(), p.new_acum()
    for it in range(nsteps):
        if p.s is not None:
            o.add_source(dt, cr, p.s)
        if it % 2 == 0:
            o.solve_x(dt, cr, p.k1, ul)
            o.solve_y(dt, cr, p.k2, vl)
        else:
            o.solve_y(dt, cr, p.k2, vl)
            o.solve_x(dt, cr, p.k1, ul)
        o.solve_z(dt, cr, ar, p.k3, p.w1, p.w2, stime, p.vs)
    scale = np.abs(o.flux).max() + 1e-30
    if state == 2:
        assert np.array_equal(cg, cr) and np.array_equal(ag, ar)
        assert np.all(np.abs(fg - o.flux) <= 1e-12 * scale)
        assert bg[0] == pytest.approx(o.mass_volume(cr), rel=1e-13)
    else:
        e = np.linalg.norm(cg - cr) / np.linalg.norm(cr)
        ea = np.linalg.norm(ag - ar) / max(np.linalg.norm(ar), 1e-300)
        print(f"fortran resident fp32: c relL2 {e:.3e}, acum relL2 {ea:.3e}", flush=True)
        assert e <= 1e-5 and ea <= 1e-5
        assert np.all(np.abs(fg - o.flux) <= 1e-5 * scale + 1e-5 * np.abs(o.flux))
        assert bg[0] == pytest.approx(o.mass_volume(cr), rel=1e-5)
    # the run resumed from the halfway snapshot in a second context
    assert np.array_equal(c2, cg) and np.array_equal(a2, ag)
    assert "restored at 3" in r.stdout
