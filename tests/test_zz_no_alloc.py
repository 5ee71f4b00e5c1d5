"""fall3d_ads_step_noalloc through the C ABI: bitwise the results of
fall3d_ads_step, on two calls in a row where the second uses a smaller grid
(the shared static partial-sum array then still holds the first call's
larger contents), and a clear error for a step whose partial sums do not fit.
"""
import ctypes

import numpy as np
import pytest

from fall3d_amd import ads
from fall3d_amd.problem import operational_problem

pytestmark = pytest.mark.gpu


def _pair(p, state="fp32"):
    out = []
    for _ in range(2):
        g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, state)
        g.set_problem(p)
        out.append(g)
    return out


def _run_both(p, nsteps, state="fp32"):
    ref, na = _pair(p, state)
    dt = ref.time_step(0.0, 2, 0.9)
    lib = ads.load_library()
    for it in range(nsteps):
        assert lib.fall3d_ads_step(ref._h, ctypes.c_double(dt), ctypes.c_double(0.0), it) == 0
        rc = lib.fall3d_ads_step_noalloc(na._h, ctypes.c_double(dt), ctypes.c_double(0.0), it)
        assert rc == 0, lib.fall3d_ads_last_error(na._h).decode()
    c_r, a_r = ref.get_state()
    c_n, a_n = na.get_state()
    b_r, b_n = ref.budget(), na.budget()
    f_r, f_n = ref.fluxes(), na.fluxes()
    ref.release()
    na.release()
    assert np.array_equal(c_r, c_n)
    assert np.array_equal(a_r, a_n)
    assert np.array_equal(f_r, f_n)
    assert np.array_equal(b_r["ledger"], b_n["ledger"]) and b_r["injected"] == b_n["injected"]
    return float(np.abs(c_r).max())


@pytest.mark.parametrize("state", ["fp32", "fp32x2"])
def test_noalloc_step_equals_step_on_two_calls(state):
    big = operational_problem(96, 80, 30, 4, seed=2)
    small = operational_problem(40, 36, 12, 3, seed=9)
    print(f"case: noalloc {state} 96x80x30x4 then 40x36x12x3", flush=True)
    assert _run_both(big, 4, state) > 0.0
    assert _run_both(small, 3, state) > 0.0


def test_noalloc_rejects_steps_beyond_the_scratch():
    """64 bins, one bin per workgroup, 9000 rows: the z sweep alone needs
    64 * 9000 workgroups of 4 partial sums (2.3 M doubles), more than the
    2^21 doubles the entry point holds."""
    nx, ny, nz, nb = 64, 9000, 4, 64
    g = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 4, "fp32")
    for ax in (1, 2):
        g.set_tiling(ax, cells_per_thread=1, bins_per_group=1)
    lib = ads.load_library()
    p = operational_problem(nx, ny, nz, 1, seed=1)
    g.set_grid(p.grid)
    f = np.float32
    vs = np.zeros((nx, ny, nz + 3, nb), dtype=f, order="F")
    g.set_meteo(p.u1.astype(f), p.v1.astype(f), p.w1.astype(f), p.k1.astype(f), p.k2.astype(f),
                p.k3.astype(f), vs)
    rc = lib.fall3d_ads_step_noalloc(g._h, ctypes.c_double(1.0), ctypes.c_double(0.0), 0)
    msg = lib.fall3d_ads_last_error(g._h).decode()
    print(f"rc {rc}: {msg}", flush=True)
    assert rc == 1 and "allocation-free" in msg
    # the context is unharmed: the allocating step still runs
    g.step(1.0, 0)
    g.synchronize()
    g.release()
