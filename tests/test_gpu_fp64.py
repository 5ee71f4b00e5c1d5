"""FP64 state (FALL3D_ADS_STATE_FP64) against the FP64 oracle.

The FP64 sweeps evaluate the reference's statements in double precision, in
the reference's order and without contraction, as the oracle does (it is
built with -ffp-contract=off): concentrations (ghost cells included) and the
ground deposit must EQUAL the oracle's (numpy equality, so +0 and -0 are the
same value), for every limiter and time marching, on uniform and stretched
metrics, through the drop-in sweeps and through resident steps with sources,
time-interpolated winds and graph replay. Boundary fluxes and ledgers are
sums over lines taken in a different order: they are compared to 1e-12 of
their scale. Mass + ledger - injected closes to FP64 rounding.
"""
import numpy as np
import pytest

from fall3d_amd import ads
from fall3d_amd.problem import operational_problem, puff_problem, random_problem
from oracle.ads_oracle import Oracle, set_num_threads
from test_gpu_parity import _flux_scale, _problem

pytestmark = pytest.mark.gpu


def _equal(gpu, ref, what):
    diff = np.count_nonzero(gpu != ref)
    if diff:
        d = np.abs(gpu - ref).max() / max(np.abs(ref).max(), 1e-300)
        print(f"{what}: {diff} values differ (max rel {d:.3e})", flush=True)
    assert diff == 0, what


@pytest.mark.parametrize("grid", ["uniform", "stretched"])
@pytest.mark.parametrize("tm", [1, 2])
@pytest.mark.parametrize("lim", [1, 2, 3])
def test_dropin_sweeps_bitwise(lim, tm, grid):
    print(f"case fp64 drop-in lim={lim} tm={tm} grid={grid}", flush=True)
    p = _problem(7 * lim + tm, grid, 13, 11, 9, 3)
    o = Oracle(p, lim, tm)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, lim, tm, 8, "fp64", grid=p.grid)
    u, v, _ = g.get_vcomps()
    u[...] = p.u1
    v[...] = p.v1
    dt = o.time_step(p.u1, p.v1, p.w1, p.vs, criterion=2, cfl=0.5)
    scale = _flux_scale(p, dt)
    cr, cg = p.new_c(), p.new_c()
    ar, ag = p.new_acum(), p.new_acum()
    o.solve_x(dt, cr, p.k1, p.u1)
    W, E = g.solve_along_x(0.0, 0.0, dt, cg, p.k1)
    _equal(cg, cr, "x sweep c")
    o.solve_y(dt, cr, p.k2, p.v1)
    S, N = g.solve_along_y(0.0, 0.0, dt, cg, p.k2)
    _equal(cg, cr, "y sweep c")
    o.solve_z(dt, cr, ar, p.k3, p.w1, p.w2, 0.3, p.vs)
    D, U = g.solve_along_z(0.0, 0.0, dt, cg, ag, p.k3, p.w1, p.w2, 0.3, p.vs)
    _equal(cg, cr, "z sweep c")
    _equal(ag, ar, "z sweep acum")
    got = np.array([W, E, S, N, D, U])
    assert np.all(np.abs(got - o.flux) <= 1e-12 * scale), (got, o.flux)
    g.release()


def test_dropin_x_sweep_more_planes_than_one_transpose_launch():
    """The FP64 x sweep runs on an (x, y)-transposed copy of the state; a
    transpose launch covers at most 65535 (z, bin) planes, so 6 z positions
    (nz + 4) x 10923 bins = 65538 planes take two launches each way."""
    print("case fp64 x sweep 6x4x2x10923", flush=True)
    p = random_problem(6, 4, 2, 10923, seed=5)
    assert (p.nz + 4) * p.nb > 65535
    o = Oracle(p, 1, 2)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, 1, 2, 8, "fp64", grid=p.grid)
    u, _, _ = g.get_vcomps()
    u[...] = p.u1
    dt = o.time_step(p.u1, p.v1, p.w1, p.vs, criterion=2, cfl=0.5)
    cr, cg = p.new_c(), p.new_c()
    o.solve_x(dt, cr, p.k1, p.u1)
    g.solve_along_x(0.0, 0.0, dt, cg, p.k1)
    _equal(cg, cr, "x sweep c, 65538 planes")
    b = g.budget()
    assert b["mass_volume"] == pytest.approx(o.mass_volume(cr), rel=1e-13)
    g.release()


@pytest.mark.parametrize("grid", ["uniform", "stretched"])
def test_time_step_bitwise(grid):
    """grid_get_time_step from the FP64 state's own fields equals the FP64
    oracle bit for bit: the rates in the reference's order (criterion 2 adds
    its six terms left to right), the max reductions are exact."""
    print(f"case fp64 time step {grid}", flush=True)
    p = _problem(5, grid, 23, 19, 11, 3)
    o = Oracle(p, 1, 2)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, 1, 2, 8, "fp64", grid=p.grid)
    g.set_problem(p)
    for stime in (0.0, 0.3):
        u = np.asfortranarray((1.0 - stime) * p.u1 + stime * p.u2)
        v = np.asfortranarray((1.0 - stime) * p.v1 + stime * p.v2)
        w = np.asfortranarray((1.0 - stime) * p.w1 + stime * p.w2)
        for crit, cfl in ((1, 0.5), (2, 0.9)):
            got = g.time_step(stime=stime, criterion=crit, cfl=cfl)
            ref = o.time_step(u, v, w, p.vs, criterion=crit, cfl=cfl)
            assert got == ref, (stime, crit, got, ref)
    g.release()


def test_dropin_real4_host_arrays():
    """rp = 4 host arrays into the FP64 state: the inputs widen exactly, so
    the sweep is the FP64 sweep of the widened field."""
    p = random_problem(10, 9, 8, 2, seed=99)
    o = Oracle(p, 1, 2)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, 1, 2, 4, "fp64", grid=p.grid)
    u, _, _ = g.get_vcomps()
    u32 = p.u1.astype(np.float32)
    u[...] = u32
    dt = 0.5 * o.time_step(p.u1, p.v1, p.w1, p.vs, cfl=0.5)
    c32 = np.asfortranarray(p.c0.astype(np.float32))
    k32 = p.k1.astype(np.float32)
    cr = np.asfortranarray(c32.astype(np.float64))
    o.solve_x(dt, cr, np.asfortranarray(k32.astype(np.float64)), np.asfortranarray(u32.astype(np.float64)))
    g.solve_along_x(0.0, 0.0, dt, c32, k32)
    _equal(c32, cr.astype(np.float32), "x sweep r4")


@pytest.mark.parametrize("grid", ["uniform", "stretched"])
@pytest.mark.parametrize("lim,tm", [(1, 2), (2, 2), (3, 1)])
def test_resident_steps_bitwise(lim, tm, grid):
    """Source injection, x/y order by parity, z with time-interpolated winds:
    7 steps (three replays of the captured step pair and one direct step)."""
    print(f"case fp64 resident lim={lim} tm={tm} grid={grid}", flush=True)
    p = _problem(31 + lim, grid, 17, 12, 10, 2, umax_scale=0.5)
    o = Oracle(p, lim, tm)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, lim, tm, 8, "fp64")
    g.set_problem(p)
    stime = 0.4
    u = np.asfortranarray((1 - stime) * p.u1 + stime * p.u2)
    v = np.asfortranarray((1 - stime) * p.v1 + stime * p.v2)
    w = np.asfortranarray((1 - stime) * p.w1 + stime * p.w2)
    dt = o.time_step(u, v, w, p.vs, cfl=0.5)
    m0 = g.budget()["mass_cells"]
    cr, ar = p.new_c(), p.new_acum()
    inj = 0.0
    for it in range(7):
        inj += o.step(dt, it, cr, ar, stime=stime)
    g.run(dt, 7, 0, stime=stime)
    cg, ag = g.get_state()
    _equal(cg, cr, "resident c")
    _equal(ag, ar, "resident acum")
    scale = _flux_scale(p, dt) * 7
    fl = g.fluxes()
    assert np.all(np.abs(fl - o.flux) <= 1e-12 * scale), (fl, o.flux)
    b = g.budget()
    assert b["mass_volume"] == pytest.approx(o.mass_volume(cr), rel=1e-13)
    if grid == "uniform":
        # the oracle returns the injected mass in the dX_p measure, the
        # library in the dX_b measure of mass_cells: equal on uniform grids
        assert b["injected"] == pytest.approx(inj, rel=1e-12)
    assert np.allclose(b["ledger"], o.ledger, rtol=1e-12, atol=1e-12 * scale)
    closure = abs(b["mass_cells"] + b["ledger"].sum() + b["residual"] - b["injected"] - m0) / m0
    print(f"fp64 closure {closure:.2e}", flush=True)
    assert b["residual"] == 0.0
    assert closure <= 1e-12
    g.release()


def test_puff_200_steps_bitwise():
    """64x64x32 Gaussian puff, minmod + RK4, CFL 0.9, 200 steps: the device
    state and deposit equal the oracle's."""
    set_num_threads(16)
    p = puff_problem()
    o = Oracle(p, "minmod", "rk")
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp64")
    g.set_problem(p)
    dt = o.time_step(p.u1, p.v1, p.w1, p.vs, cfl=0.9)
    cr, ar = p.new_c(), p.new_acum()
    for it in range(200):
        o.step(dt, it, cr, ar)
    g.run(dt, 200, 0)
    cg, ag = g.get_state()
    _equal(cg, cr, "puff c")
    _equal(ag, ar, "puff deposit")
    g.release()


def test_operational_full_size_two_steps_bitwise():
    """360x240x60, 10 bins (the operational domain), two steps."""
    set_num_threads(16)
    p = operational_problem()
    o = Oracle(p, "minmod", "rk")
    dt = o.time_step(p.u1, p.v1, p.w1, p.vs, cfl=0.9)
    cr, ar = p.new_c(), p.new_acum()
    for it in range(2):
        o.step(dt, it, cr, ar)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp64")
    g.set_problem(p)
    g.run(dt, 2, 0)
    cg, ag = g.get_state()
    _equal(cg, cr, "operational c")
    _equal(ag, ar, "operational deposit")
    g.release()


def test_fp64_restart_and_health():
    p = operational_problem(40, 32, 20, 3, seed=5)
    a = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp64")
    a.set_problem(p)
    dt = a.time_step(0.0, 2, 0.9)
    a.run(dt, 3, 0)
    blob = a.snapshot(3 * dt, 3)
    a.run(dt, 4, 3)
    ca, aa = a.get_state()
    ba = a.budget()
    b = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp64")
    b.set_problem(p)
    assert b.restore(blob) == (3 * dt, 3)
    b.run(dt, 4, 3)
    cb, ab = b.get_state()
    _equal(cb, ca, "resumed c")
    _equal(ab, aa, "resumed deposit")
    bb = b.budget()
    for k in ("mass_cells", "injected"):
        assert bb[k] == ba[k]
    # the health scan of the FP64 state counts what numpy counts on the interior
    h = b.check_state()
    inner = cb[2:-2, 2:-2, 2:-2, :]
    assert h["nonfinite"] == 0
    assert h["negative"] == int(np.count_nonzero(inner < 0.0))
    assert h["min"] == inner.min()
    a.release()
    b.release()


def test_fp64_rejects_fp32_features():
    p = random_problem(12, 10, 8, 2, seed=1)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp64")
    g.set_problem(p)
    with pytest.raises(ads.Fall3dAdsError):
        g.set_tiling(0, cells_per_thread=4)
    with pytest.raises(ads.Fall3dAdsError):
        g.step_noalloc(1.0, 0)
    with pytest.raises(ads.Fall3dAdsError):
        g.autotune(1.0)
    g.release()
