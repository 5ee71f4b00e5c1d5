"""The Fortran MODULE ADS shim (fall3d_amd/fortran/mod_ADS_hip.F90) used
exactly as Sources/mod_F3D.F90:1435-1491 uses the reference module, through
ISO_C_BINDING and the C ABI, compared with the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from fall3d_amd.problem import random_problem, puff_problem
from oracle.ads_oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "fall3d_amd", "lib", "fall3d_dropin_driver")


def _write_input(path, p, limiter, tm, nsteps, state, dt, stime, u, v):
    with open(path, "wb") as f:
        np.array([p.nx, p.ny, p.nz, p.nb, limiter, tm, nsteps, state], dtype=np.int32).tofile(f)
        np.array([dt, stime], dtype=np.float64).tofile(f)
        for k in ("dX1_p", "dX1_b", "dX2_p", "dX2_b", "dX3_p", "dX3_b"):
            np.asarray(p.grid[k], dtype=np.float64).tofile(f)
        s = p.s if p.s is not None else np.zeros((p.nx, p.ny, p.nz, p.nb), order="F")
        for a in (u, v, p.w1, p.w2, p.k1, p.k2, p.k3, p.vs, s, p.c0, p.new_acum()):
            np.asarray(a, dtype=np.float64).ravel(order="F").tofile(f)


def _read_output(path, p):
    raw = np.fromfile(path, dtype=np.float64)
    nc = (p.nx + 4) * (p.ny + 4) * (p.nz + 4) * p.nb
    na = (p.nx + 4) * (p.ny + 4) * p.nb
    c = raw[:nc].reshape((p.nx + 4, p.ny + 4, p.nz + 4, p.nb), order="F")
    a = raw[nc:nc + na].reshape((p.nx + 4, p.ny + 4, p.nb), order="F")
    return c, a, raw[nc + na:nc + na + 6]


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["random", "puff"])
def test_fortran_module_ads_dropin(tmp_path, case):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran drop-in driver not built (python -m fall3d_amd.build)")
    if case == "random":
        p = random_problem(14, 11, 9, 2, seed=8, umax_scale=0.4)
        nsteps, stime = 4, 0.35
    else:
        p = puff_problem(32, 30, 16)
        nsteps, stime = 10, 0.0
    o = Oracle(p, "minmod", "rk")
    dt = o.time_step(p.u1, p.v1, p.w1, p.vs, cfl=0.6)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, p, 1, 2, nsteps, 0, dt, stime, p.u1, p.v1)
    print(f"case {case}: running {DRIVER}", flush=True)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    cg, ag, fg = _read_output(fout, p)
    # the same sequence on the oracle (steady u, v; w interpolated at stime)
    cr, ar = p.new_c(), p.new_acum()
    for it in range(nsteps):
        if p.s is not None:
            o.add_source(dt, cr, p.s)
        if it % 2 == 0:
            o.solve_x(dt, cr, p.k1, p.u1)
            o.solve_y(dt, cr, p.k2, p.v1)
        else:
            o.solve_y(dt, cr, p.k2, p.v1)
            o.solve_x(dt, cr, p.k1, p.u1)
        o.solve_z(dt, cr, ar, p.k3, p.w1, p.w2, stime, p.vs)
    e = np.linalg.norm(cg - cr) / np.linalg.norm(cr)
    ea = np.linalg.norm(ag - ar) / max(np.linalg.norm(ar), 1e-300)
    print(f"fortran drop-in {case}: c relL2 {e:.3e}, acum relL2 {ea:.3e}", flush=True)
    assert e <= 1e-5 and ea <= 1e-5
    scale = np.abs(o.flux).max() + 1e-30
    assert np.all(np.abs(fg - o.flux) <= 1e-5 * scale + 1e-5 * np.abs(o.flux))
