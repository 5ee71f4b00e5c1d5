"""Restart: a snapshot of the device state and accumulators, restored into
a fresh context, resumes bitwise (the device-side counterpart of
nc_IO_out_rst / nc_IO_read_rst, Sources/mod_nc_IO.F90:4524-5178).

Run A: n1 steps, snapshot, n2 more steps. Run B: a new context with the same
grid, meteo and source, the snapshot restored, n2 steps from the stored
iteration. Concentration (ghosts included), deposit, reference fluxes and
the whole FP64 budget must be equal bit for bit.
"""
import os

import numpy as np
import pytest

from fall3d_amd import ads
from fall3d_amd.problem import operational_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("state", ["fp32", "fp32x2"])
def test_resume_is_bitwise(state, tmp_path):
    print(f"case: restart {state}", flush=True)
    p = operational_problem(64, 48, 24, 3, seed=4)
    a = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, state)
    a.set_problem(p)
    dt = a.time_step(0.0, 2, 0.9)
    n1, n2 = 7, 6
    a.run(dt, n1, 0)
    path = os.path.join(tmp_path, "ads.rst")
    a.save_restart(path, time=n1 * dt, iiter=n1)
    a.run(dt, n2, n1)
    ca, aa = a.get_state()
    fa, ba = a.fluxes(), a.budget()
    a.release()

    b = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, state)
    b.set_problem(p)  # grid, meteo and source; the state comes from the snapshot
    t, it = b.load_restart(path)
    assert t == n1 * dt and it == n1
    b.run(dt, n2, it)
    cb, ab = b.get_state()
    fb, bb = b.fluxes(), b.budget()
    b.release()
    assert np.array_equal(ca, cb)
    assert np.array_equal(aa, ab)
    assert np.array_equal(fa, fb)
    for k in ("mass_cells", "injected", "residual", "ground", "lateral"):
        assert ba[k] == bb[k], k
    assert np.array_equal(ba["ledger"], bb["ledger"])


def test_restore_rejects_another_grid():
    p = operational_problem(40, 32, 16, 2, seed=1)
    a = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
    a.set_problem(p)
    blob = a.snapshot(0.0, 0)
    a.release()
    b = ads.ADS(p.nx + 1, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
    with pytest.raises(ads.Fall3dAdsError, match="grid"):
        b.restore(blob)
    c = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32x2")
    with pytest.raises(ads.Fall3dAdsError, match="state mode"):
        c.restore(blob)
    with pytest.raises(ads.Fall3dAdsError, match="magic"):
        c.restore(b"\0" * 64 + blob[64:])
    b.release()
    c.release()


@pytest.mark.parametrize("nx,state", [(22, "fp32"), (22, "fp32x2"), (61, "fp32"), (28, "fp32")])
def test_snapshot_holds_the_padded_state_rows(nx, state):
    """The FP32 state's x rows are padded on the device to a multiple of 32
    floats (nx = 22: 26 cells with ghosts in 32 words; 61: 65 in 96; 28: 32,
    no pad). The snapshot carries those words last: after steps with sources
    every pad word is still zero, and the cells equal get_state (hi + lo
    summed in FP64, as get_state forms them) bit for bit."""
    ny, nz, nb = 9, 6, 2
    print(f"case: padded rows nx={nx} {state}", flush=True)
    p = operational_problem(nx, ny, nz, nb, seed=8)
    a = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 8, state)
    a.set_problem(p)
    dt = a.time_step(0.0, 2, 0.9)
    a.run(dt, 5, 0)
    c, _ = a.get_state()
    blob = np.frombuffer(a.snapshot(5 * dt, 5), dtype=np.uint8)
    a.release()
    pitch = (nx + 4 + 31) // 32 * 32
    ns = pitch * (ny + 4) * (nz + 4) * nb
    parts = 2 if state == "fp32x2" else 1
    words = blob[blob.size - 4 * ns * parts:].view(np.float32).reshape(parts, nb, nz + 4, ny + 4, pitch)
    assert np.count_nonzero(words[..., nx + 4:]) == 0
    got = words[0, ..., :nx + 4].astype(np.float64)
    if parts == 2:
        got = got + words[1, ..., :nx + 4].astype(np.float64)
    assert np.count_nonzero(c) > 0
    assert np.array_equal(got.reshape(-1), np.asarray(c).reshape(-1, order="F"))
