"""Parity of the gfx950 ADS kernels (through the C ABI) with the FP64 oracle.

Tolerances: the device computes reconstruction, fluxes and stage updates in
FP32 (FP64 only for the boundary/deposit accumulators, CFL and budget), so
concentrations and deposits are compared with relative L2 <= 1e-5 (the
required FP32 accuracy) and max-norm <= 1e-4 of the field's maximum; FP64-only
paths (ADS_solve_1D) with 1e-12. Mass conservation of the compensated
state (fp32x2) is checked to 1e-12 relative.
"""
import numpy as np
import pytest

from fall3d_amd import ads
from fall3d_amd.problem import operational_problem, puff_problem, random_problem
from oracle.ads_oracle import Oracle, solve_1d as ora_solve_1d, set_num_threads

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5
MAX_REL = 1e-4


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def check_field(gpu, ref, what, rel=REL_L2, mx=MAX_REL):
    e2 = rel_l2(gpu, ref)
    em = np.abs(gpu - ref).max() / max(np.abs(ref).max(), 1e-300)
    print(f"{what}: relL2={e2:.3e} max={em:.3e}", flush=True)
    assert e2 <= rel, (what, e2)
    assert em <= mx, (what, em)


def _dt(p, o, cfl=0.5):
    return o.time_step(p.u1, p.v1, p.w1, p.vs, criterion=2, cfl=cfl)


def _flux_scale(p, dt):
    vmax = max(np.abs(p.u1).max(), np.abs(p.v1).max(), np.abs(p.w1).max() + np.abs(p.vs).max())
    g = p.grid
    area = max(g["dX1_b"].max(), g["dX2_b"].max(), g["dX3_b"].max()) ** 2
    nl = max(p.nx * p.ny, p.ny * p.nz, p.nx * p.nz) * p.nb
    return vmax * np.abs(p.c0).max() * area * dt * nl


def _stretch_all_axes(p, seed):
    """Stretched metrics on every axis: a sigma grid in z (random_problem's
    stretched option, Sources/mod_Grid.F90:541-562) and independent random
    dX_p / dX_b arrays in x and y, so no two neighbouring cells share a
    spacing and dxp differs from dxb everywhere."""
    rng = np.random.default_rng(1000 + seed)
    g = p.grid
    for ax, n, d in ((1, p.nx, 1000.0), (2, p.ny, 1200.0)):
        g[f"dX{ax}_p"] = rng.uniform(0.6, 1.4, n + 4) * d
        g[f"dX{ax}_b"] = rng.uniform(0.6, 1.4, n + 3) * d
    return p


def _problem(seed, grid, *shape, **kw):
    if grid == "uniform":
        return random_problem(*shape, seed=seed, **kw)
    return _stretch_all_axes(random_problem(*shape, seed=seed, stretched=True, **kw), seed)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("grid", ["uniform", "stretched"])
@pytest.mark.parametrize("state", ["fp32", "fp32x2"])
@pytest.mark.parametrize("tm", [1, 2])
@pytest.mark.parametrize("lim", [1, 2, 3])
def test_dropin_sweeps_match_oracle(lim, tm, state, grid):
    """ADS_solve_along_x/y/z through the drop-in entry points, random problem
    (every face-velocity sign, ragged segment sizes, plateaus and zeros), on
    uniform and on stretched metrics of every axis."""
    print(f"case lim={lim} tm={tm} state={state} grid={grid}", flush=True)
    p = _problem(7 * lim + tm, grid, 13, 11, 9, 3)
    o = Oracle(p, lim, tm)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, lim, tm, 8, state, grid=p.grid)
    u, v, w = g.get_vcomps()
    u[...] = p.u1
    v[...] = p.v1
    dt = _dt(p, o)
    scale = _flux_scale(p, dt)
    cr, cg = p.new_c(), p.new_c()
    ar, ag = p.new_acum(), p.new_acum()
    o.solve_x(dt, cr, p.k1, p.u1)
    W, E = g.solve_along_x(0.0, 0.0, dt, cg, p.k1)
    check_field(cg, cr, "x sweep c")
    assert abs(W - o.flux[0]) <= 1e-5 * scale and abs(E - o.flux[1]) <= 1e-5 * scale
    o.solve_y(dt, cr, p.k2, p.v1)
    S, N = g.solve_along_y(0.0, 0.0, dt, cg, p.k2)
    check_field(cg, cr, "y sweep c")
    assert abs(S - o.flux[2]) <= 1e-5 * scale and abs(N - o.flux[3]) <= 1e-5 * scale
    o.solve_z(dt, cr, ar, p.k3, p.w1, p.w2, 0.3, p.vs)
    D, U = g.solve_along_z(0.0, 0.0, dt, cg, ag, p.k3, p.w1, p.w2, 0.3, p.vs)
    check_field(cg, cr, "z sweep c")
    check_field(ag, ar, "z sweep acum")
    assert abs(D - o.flux[4]) <= 1e-5 * scale and abs(U - o.flux[5]) <= 1e-5 * scale
    g.release()


def test_dropin_real4():
    """rp = 4 host kind (Sources/mod_KindType.F90:18-19, WITH_R4)."""
    p = random_problem(10, 9, 8, 2, seed=99)
    o = Oracle(p, 1, 2)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, 1, 2, 4, "fp32", grid=p.grid)
    u, v, w = g.get_vcomps()
    u[...] = p.u1.astype(np.float32)
    dt = _dt(p, o)
    cr = p.new_c()
    cg = np.asfortranarray(p.c0.astype(np.float32))
    o.solve_x(dt, cr, p.k1, np.asfortranarray(p.u1.astype(np.float32).astype(np.float64)))
    g.solve_along_x(0.0, 0.0, dt, cg, p.k1.astype(np.float32))
    check_field(cg.astype(np.float64), cr, "x sweep r4", rel=1e-5)


def test_update_horizontal_vcomps():
    p = random_problem(7, 6, 5, 1, seed=4)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, grid=p.grid)
    u2 = np.asfortranarray(p.u1 * 1.5 + 1.0)
    v2 = np.asfortranarray(p.v1 * 0.5 - 2.0)
    g.update_horizontal_vcomps(p.u1, u2, p.v1, v2, 0.25)
    u, v, _ = g.get_vcomps()
    np.testing.assert_allclose(u, 0.75 * p.u1 + 0.25 * u2, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(v, 0.75 * p.v1 + 0.25 * v2, rtol=1e-6, atol=1e-6)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("grid", ["uniform", "stretched"])
@pytest.mark.parametrize("state", ["fp32", "fp32x2"])
def test_resident_steps_match_oracle(state, grid):
    """fall3d_ads_step = F3D_time_step (source, x/y order by parity, z) with
    time-interpolated winds (stime != 0), several steps, on uniform and
    stretched metrics."""
    print(f"case state={state} grid={grid}", flush=True)
    p = _problem(31, grid, 17, 12, 10, 2, umax_scale=0.5)
    o = Oracle(p, 1, 2)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, 1, 2, 8, state)
    g.set_problem(p)
    stime = 0.4
    u = np.asfortranarray((1 - stime) * p.u1 + stime * p.u2)
    v = np.asfortranarray((1 - stime) * p.v1 + stime * p.v2)
    w = np.asfortranarray((1 - stime) * p.w1 + stime * p.w2)
    dt = o.time_step(u, v, w, p.vs, cfl=0.5)
    dtg = g.time_step(stime=stime, criterion=2, cfl=0.5)
    assert dtg == pytest.approx(dt, rel=1e-6)
    cr, ar = p.new_c(), p.new_acum()
    for it in range(5):
        o.step(dt, it, cr, ar, stime=stime)
        g.step(dt, it, stime=stime)
    cg, ag = g.get_state()
    check_field(cg[2:-2, 2:-2, 2:-2], cr[2:-2, 2:-2, 2:-2], "resident c")
    check_field(ag, ar, "resident acum")
    fl = g.fluxes()
    scale = _flux_scale(p, dt) * 5
    assert np.all(np.abs(fl - o.flux) <= 1e-5 * scale), (fl, o.flux)


# ---------------------------------------------------------------------------
def test_puff_200_steps_matches_fp64():
    """64x64x32 Gaussian puff, one bin (v_s = 0.5 m/s), Kh = 100, Kv = 10,
    zero-flux laterals/top, absorbing ground, CFL 0.9, minmod + RK4, 200
    steps: FP32 device vs FP64 oracle (concentration and ground deposit)."""
    set_num_threads(16)
    p = puff_problem()
    o = Oracle(p, "minmod", "rk")
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
    g.set_problem(p)
    dt = o.time_step(p.u1, p.v1, p.w1, p.vs, cfl=0.9)
    dtg = g.time_step(0.0, 2, 0.9)
    assert dtg == pytest.approx(dt, rel=1e-6)
    cr, ar = p.new_c(), p.new_acum()
    for it in range(200):
        o.step(dt, it, cr, ar)
    g.run(dt, 200, 0)
    cg, ag = g.get_state()
    check_field(cg[2:-2, 2:-2, 2:-2], cr[2:-2, 2:-2, 2:-2], "puff c")
    check_field(ag, ar, "puff deposit")
    fl = g.fluxes()
    assert fl[4] == pytest.approx(o.flux[4], rel=1e-5)
    b = g.budget()
    assert b["mass_volume"] == pytest.approx(o.mass_volume(cr), rel=1e-5)


@pytest.mark.parametrize("state,tol", [("fp32x2", 1e-12)])
def test_puff_mass_conserved(state, tol):
    """Compensated state: domain mass + mass that left (ledger) = initial mass
    to FP64 rounding, over 200 steps."""
    p = puff_problem()
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, state)
    g.set_problem(p)
    dt = g.time_step(0.0, 2, 0.9)
    m0 = g.budget()["mass_cells"]
    g.run(dt, 200, 0)
    b = g.budget()
    err = abs(b["mass_cells"] + b["ledger"].sum() - m0) / m0
    print(f"mass closure {state}: {err:.3e} (deposited {b['ledger'][4] / m0:.3f} of M0)", flush=True)
    assert err <= tol


def test_operational_small_budget_closes():
    """Fourier wind with open laterals, 4 bins, line source: M + out = M0 + injected."""
    p = operational_problem(72, 48, 30, 4, seed=3)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32x2")
    g.set_problem(p)
    dt = g.time_step(0.0, 2, 0.9)
    g.run(dt, 60, 0)
    b = g.budget()
    assert b["injected"] > 0
    err = abs(b["mass_cells"] + b["ledger"].sum() - b["injected"]) / b["injected"]
    print(f"operational closure {err:.3e}", flush=True)
    assert err <= 1e-12


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("lim", [1, 2, 3])
@pytest.mark.parametrize("bcode", [(2, 2), (1, 1), (-1, -1), (0, 2)])
def test_solve_1d_matches_oracle(lim, bcode):
    rng = np.random.default_rng(lim * 10 + bcode[0] + 5)
    n = 37
    c = rng.uniform(0.0, 1.0, n + 4)
    dxp = rng.uniform(500.0, 900.0, n + 4)
    dxb = rng.uniform(500.0, 900.0, n + 3)
    u = rng.normal(2.0, 3.0, n + 3)
    k = rng.uniform(10.0, 90.0, n + 4)
    cr = c.copy()
    fr = ora_solve_1d(cr, dxp, dxb, u, k, 20.0, (0.2, 0.7), bcode, lim)
    g = ads.ADS(4, 4, 4, 1)
    cg = c.copy()
    fg = g.solve_1d(cg, dxp, dxb, u, k, 20.0, (0.2, 0.7), bcode, 1, n, lim)
    np.testing.assert_allclose(cg, cr, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(fg, fr, rtol=1e-11, atol=1e-14)


def test_deterministic_and_tiling_invariant():
    """Two runs are bitwise equal; other tilings (cells per lane of the x
    sweep, bins per workgroup) give the same concentration bits (shared faces
    are computed by the same instruction sequence); the FP64 ledger differs
    only by the summation order of the per-block partial sums."""
    p = operational_problem(40, 36, 20, 3, seed=5)
    outs = []
    for cpt, bpg in ((0, 0), (0, 0), (3, 1), (4, 2)):
        g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32x2")
        g.set_problem(p)
        if cpt or bpg:
            for ax in range(3):
                g.set_tiling(ax, cells_per_thread=cpt if ax == 0 else 0, bins_per_group=bpg)
        dt = g.time_step(0.0, 2, 0.9)
        g.run(dt, 6, 0)
        outs.append((g.get_state()[0], g.budget()["ledger"]))
        g.release()
    assert np.array_equal(outs[1][0], outs[0][0]) and np.array_equal(outs[1][1], outs[0][1])
    for c, led in outs[2:]:
        assert np.array_equal(c, outs[0][0])
        np.testing.assert_allclose(led, outs[0][1], rtol=1e-12, atol=1e-12 * np.abs(outs[0][1]).max())


@pytest.mark.parametrize("state,tm", [("fp32", "rk"), ("fp32x2", "rk"), ("fp32", "euler")])
def test_multiwave_x_lines_bit_identical(state, tm):
    """Lines held by 2 or 3 waves (edge cells traded through LDS) give the
    same bits as lines held by one wave."""
    p = operational_problem(200, 9, 6, 2, seed=8)
    outs = []
    for cpt in (4, 2, 3):  # 51 segments: one wave; 101 and 67: two waves
        g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", tm, 8, state)
        g.set_problem(p)
        g.set_tiling(0, cells_per_thread=cpt)
        dt = g.time_step(0.0, 2, 0.9)
        g.run(dt, 5, 0)
        outs.append(g.get_state())
        g.release()
    for c, a in outs[1:]:
        assert np.array_equal(c, outs[0][0])
        assert np.array_equal(a, outs[0][1])


def test_graph_replays_and_sampled_timing_match_direct_steps():
    """fall3d_ads_run with graph replays (from an even and an odd first
    iteration) and with per-kernel events on one step in 4 gives the same
    bits as direct launches; the sampled run times exactly the steps
    iiter0, iiter0 + 4, ..."""
    p = operational_problem(40, 36, 20, 3, seed=5)
    outs = []
    for mode in ("direct", "graph", "graph_odd", "sampled"):
        g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
        g.set_problem(p)
        dt = g.time_step(0.0, 2, 0.9)
        g.set_graph(mode != "direct")
        if mode == "graph_odd":
            g.run(dt, 5, 0)
            g.run(dt, 8, 5)
        else:
            g.set_profiling(4 if mode == "sampled" else 0)
            g.run(dt, 13, 0)
        if mode == "sampled":
            for k in ("sweep_x", "sweep_y", "sweep_z", "reduce"):
                ms, n = g.kernel_time(k)
                assert n == 4 and ms > 0.0, (k, n, ms)
            g.set_profiling(0)
        outs.append((g.get_state()[0], g.budget()["ledger"]))
        g.release()
    for c, led in outs[1:]:
        assert np.array_equal(c, outs[0][0])
        assert np.array_equal(led, outs[0][1])


@pytest.mark.parametrize("dims", [(40, 36, 20, 3), (70, 20, 3, 4), (33, 9, 2, 5), (65, 4, 2, 2)])
def test_stream_bins_per_thread_invariant(dims):
    """y/z streaming sweeps with two bins per thread give the concentration
    and deposit bits of one bin per thread (RK4 and Euler), also for lines
    shorter than the stage pipeline (nz = 2, 3), odd bin counts and
    partial 64-column tiles; the RK4 result matches the FP64 oracle."""
    print(f"case dims={dims}", flush=True)
    nx, ny, nz, nb = dims
    p = operational_problem(nx, ny, nz, nb, seed=sum(dims))
    for tm in ("rk", "euler"):
        outs = []
        for bpt in (1, 2):
            g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", tm, 8, "fp32")
            g.set_problem(p)
            for ax in (1, 2):
                g.set_tiling(ax, cells_per_thread=bpt, bins_per_group=2 if nb > 1 else 1)
            assert g.tiling(1)[0] == bpt and g.tiling(2)[0] == bpt
            dt = g.time_step(0.0, 2, 0.9)
            g.run(dt, 5, 0)
            outs.append((g.get_state(), g.budget()["ledger"], dt))
            g.release()
        (c1, a1), l1, dt = outs[0]
        (c2, a2), l2, _ = outs[1]
        assert np.array_equal(c1, c2) and np.array_equal(a1, a2), tm
        np.testing.assert_allclose(l2, l1, rtol=1e-12, atol=1e-12 * max(np.abs(l1).max(), 1e-300))
        if tm == "rk":
            o = Oracle(p, "minmod", "rk")
            cr, ar = p.new_c(), p.new_acum()
            for it in range(5):
                o.step(dt, it, cr, ar)
            check_field(c2[2:-2, 2:-2, 2:-2], cr[2:-2, 2:-2, 2:-2], f"c two bins per thread {dims}")


def test_operational_full_size_two_steps():
    """360x240x60, 10 bins (the benchmark workload) for two steps vs the
    FP64 oracle, plus budget closure of the compensated state."""
    set_num_threads(16)
    p = operational_problem()
    o = Oracle(p, "minmod", "rk")
    dt = o.time_step(p.u1, p.v1, p.w1, p.vs, cfl=0.9)
    cr, ar = p.new_c(), p.new_acum()
    for it in range(2):
        o.step(dt, it, cr, ar)
    for state in ("fp32", "fp32x2"):
        g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, state)
        g.set_problem(p)
        g.run(dt, 2, 0)
        cg, ag = g.get_state()
        check_field(cg[2:-2, 2:-2, 2:-2], cr[2:-2, 2:-2, 2:-2], f"operational c {state}")
        check_field(ag, ar, f"operational deposit {state}")
        if state == "fp32x2":
            b = g.budget()
            err = abs(b["mass_cells"] + b["ledger"].sum() - b["injected"]) / b["injected"]
            assert err <= 1e-12
        g.release()


@pytest.mark.parametrize("nx,cpt", [(12, 4), (12, 3), (13, 4), (16, 16), (64, 2), (63, 2), (40, 20),
                                    (128, 2), (200, 2), (383, 3), (300, 2), (500, 2), (1024, 6)])
def test_x_segmentation_edge_cases(nx, cpt):
    """x sweep with every segmentation case: cell n at the end of a segment
    (the last segment then holds only ghosts and face n+1), in its middle,
    one segment per line, and 64-lane lines; lines over 2, 3 and 4 waves
    (nx/cpt + 1 segments > 64), including a last wave whose only segment
    starts at face n+1; against the oracle."""
    print(f"case nx={nx} cpt={cpt}", flush=True)
    p = random_problem(nx, 7, 5, 2, seed=nx + cpt)
    o = Oracle(p, 1, 2)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, 1, 2, 8, "fp32", grid=p.grid)
    g.set_tiling(0, cells_per_thread=cpt)
    u, v, w = g.get_vcomps()
    u[...] = p.u1
    v[...] = p.v1
    dt = _dt(p, o)
    cr, cg = p.new_c(), p.new_c()
    o.solve_x(dt, cr, p.k1, p.u1)
    W, E = g.solve_along_x(0.0, 0.0, dt, cg, p.k1)
    check_field(cg, cr, "x sweep c")
    scale = _flux_scale(p, dt)
    assert abs(W - o.flux[0]) <= 1e-5 * scale and abs(E - o.flux[1]) <= 1e-5 * scale
    g.release()


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("grid", ["uniform", "stretched"])
@pytest.mark.parametrize("stime", [0.0, 0.65])
def test_time_step_both_criteria(grid, stime):
    """grid_get_time_step (Sources/mod_Grid.F90:1026-1141) on the GPU with
    criterion 1 (largest per-direction rate) and criterion 2 (sum of the
    rates), for time-interpolated winds, against the FP64 oracle."""
    print(f"case grid={grid} stime={stime}", flush=True)
    p = _problem(44, grid, 23, 19, 11, 3)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
    g.set_problem(p)
    o = Oracle(p)
    u = np.asfortranarray((1 - stime) * p.u1 + stime * p.u2)
    v = np.asfortranarray((1 - stime) * p.v1 + stime * p.v2)
    w = np.asfortranarray((1 - stime) * p.w1 + stime * p.w2)
    for crit, cfl in ((1, 0.5), (2, 0.9)):
        ref = o.time_step(u, v, w, p.vs, criterion=crit, cfl=cfl)
        got = g.time_step(stime=stime, criterion=crit, cfl=cfl)
        print(f"criterion {crit}: {got!r} vs {ref!r}", flush=True)
        assert got == pytest.approx(ref, rel=1e-6)
    # the two criteria differ on this problem (the test tells them apart)
    assert o.time_step(u, v, w, p.vs, criterion=1, cfl=1.0) > 1.01 * o.time_step(u, v, w, p.vs, criterion=2, cfl=1.0)
    g.release()


def test_x_tiling_past_256_lanes_is_refused():
    """nx = 1024 at 3 or 4 cells per lane needs more than 256 lanes per x line:
    set_tiling returns an error and keeps the previous tiling."""
    g = ads.ADS(1024, 4, 4, 1)
    before = g.tiling(0)
    for cpt in (3, 4):
        with pytest.raises(ads.Fall3dAdsError, match="tiling"):
            g.set_tiling(0, cells_per_thread=cpt)
        assert g.tiling(0) == before
    g.release()


@pytest.mark.parametrize("case", ["puff", "operational", "stretched"])
def test_fp32_state_mass_closes_with_rounding_residual(case):
    """FP32 state: every final update is formed in FP64 and stored rounded to
    FP32; the mass that rounding moves is kept in the FP64 residual, so
    mass + ledger + residual - injected stays at the initial mass to FP64
    rounding (zero-flux puff, open operational domain with a source, and a
    problem stretched on every axis)."""
    if case == "puff":
        p, nsteps = puff_problem(), 200
    elif case == "operational":
        p, nsteps = operational_problem(72, 48, 30, 4, seed=3), 60
    else:
        p, nsteps = _problem(12, "stretched", 40, 33, 18, 3, umax_scale=0.5), 40
    print(f"case {case}", flush=True)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
    g.set_problem(p)
    dt = g.time_step(0.0, 2, 0.9 if case != "stretched" else 0.5)
    m0 = g.budget()["mass_cells"]
    g.run(dt, nsteps, 0)
    b = g.budget()
    g.release()
    ref = m0 + b["injected"]
    err = abs(b["mass_cells"] + b["ledger"].sum() + b["residual"] - ref) / ref
    naive = abs(b["mass_cells"] + b["ledger"].sum() - ref) / ref
    print(f"closure {err:.3e} (without the residual {naive:.3e}, residual {b['residual'] / ref:.3e})", flush=True)
    assert err <= 1e-12
    assert b["residual"] != 0.0


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dims", [(2, 2, 2, 1), (2, 5, 3, 2), (3, 2, 4, 3), (4, 3, 2, 2)])
@pytest.mark.parametrize("state", ["fp32", "fp64"])
def test_smallest_grids_match_oracle(dims, state):
    """The smallest grids the library accepts (2 cells along an axis): lines
    of 2-4 cells in every sweep, most cells next to a ghost. Resident steps
    with a source and time-interpolated winds against the oracle: the FP64
    state bit for bit, the FP32 state to the FP32 tolerances."""
    print(f"case dims={dims} state={state}", flush=True)
    p = random_problem(*dims, seed=sum(dims), umax_scale=0.5)
    o = Oracle(p, 1, 2)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, 1, 2, 8, state)
    g.set_problem(p)
    stime = 0.3
    u = np.asfortranarray((1 - stime) * p.u1 + stime * p.u2)
    v = np.asfortranarray((1 - stime) * p.v1 + stime * p.v2)
    w = np.asfortranarray((1 - stime) * p.w1 + stime * p.w2)
    dt = o.time_step(u, v, w, p.vs, cfl=0.5)
    cr, ar = p.new_c(), p.new_acum()
    for it in range(4):
        o.step(dt, it, cr, ar, stime=stime)
        g.step(dt, it, stime=stime)
    cg, ag = g.get_state()
    g.release()
    assert np.count_nonzero(cr[2:-2, 2:-2, 2:-2]) > 0
    if state == "fp64":
        assert np.array_equal(cg, cr) and np.array_equal(ag, ar)
    else:
        check_field(cg[2:-2, 2:-2, 2:-2], cr[2:-2, 2:-2, 2:-2], "smallest grid c")
        check_field(ag, ar, "smallest grid acum")
