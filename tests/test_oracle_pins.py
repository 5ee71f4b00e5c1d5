"""Pins of the FP64 CPU oracle (oracle/ads_oracle.c) before it is trusted.

The reference ships no tests, golden vectors or fixtures for this path and
its ADS sources cannot be built here (they need configure-generated code),
so every oracle function is pinned by what the mathematics fixes:

  * an independent literal restatement (oracle/ads_literal.py, the
    reference's own stage-over-all-lines control flow) on random tiny
    problems that hit every branch -> bitwise equality;
  * closed forms: linear profiles (limiter value 1 at r = 1, midpoint
    reconstruction, advective sign, division by dxb, RK4 consistency);
    the diffusion operator as a dense matrix polynomial with frozen ghosts
    (RK4 weights, frozen ghosts, face-averaged k, dxp/dxb roles); the
    discrete cosine eigenvector of the Neumann Laplacian for Euler;
  * invariants: conservation through zero-flux faces, exact ledger closure,
    mirror symmetry x -> -x, u -> -u, constant-state preservation;
  * textbook limiter values (Sweby diagram points) through one-cell probes;
  * the CFL formula on fields whose maximum is attained at a known face.
"""
import math

import numpy as np
import pytest

from fall3d_amd.problem import random_problem, puff_problem, uniform_grid, Problem, constant_diffusion
from oracle import ads_literal as L
from oracle.ads_oracle import Oracle, solve_1d as ora_solve_1d


def _line_problem(n, c_line, u_faces, k_cells, dx=1000.0, nb=1, dirn=0):
    """A problem whose sweep axis `dirn` carries the given line (2 copies of it
    on the transverse axes so the grid is non-degenerate)."""
    nx, ny, nz = [(n, 2, 2), (2, n, 2), (2, 2, n)][dirn]
    g = uniform_grid(nx, ny, nz, dx, dx, dx)
    u = np.zeros((nx + 3, ny, nz), order="F")
    v = np.zeros((ny + 3, nx, nz), order="F")
    w = np.zeros((nz + 3, nx, ny), order="F")
    k1, k2, k3 = constant_diffusion(nx, ny, nz, 0.0, 0.0)
    c = np.zeros((nx + 4, ny + 4, nz + 4, nb), order="F")
    vel = [u, v, w][dirn]
    kd = [k1, k2, k3][dirn]
    vel[...] = np.asarray(u_faces)[:, None, None]
    kd[...] = np.asarray(k_cells)[:, None, None]
    cl = np.asarray(c_line, dtype=np.float64)
    for b in range(nb):
        if dirn == 0:
            c[:, 2:4, 2:4, b] = cl[:, None, None]
        elif dirn == 1:
            c[2:4, :, 2:4, b] = cl[None, :, None]
        else:
            c[2:4, 2:4, :, b] = cl[None, None, :]
    vs = np.zeros((nx, ny, nz + 3, nb), order="F")
    return Problem(nx, ny, nz, nb, g, u, u.copy(order="F"), v, v.copy(order="F"), w, w.copy(order="F"),
                   k1, k2, k3, vs, c)


def _sweep(o, p, dirn, dt, c, acum=None, stime=0.0):
    if dirn == 0:
        o.solve_x(dt, c, p.k1, p.u1)
    elif dirn == 1:
        o.solve_y(dt, c, p.k2, p.v1)
    else:
        o.solve_z(dt, c, acum if acum is not None else p.new_acum(), p.k3, p.w1, p.w2, stime, p.vs)


def _line(c, dirn):
    if dirn == 0:
        return c[:, 2, 2, 0]
    if dirn == 1:
        return c[2, :, 2, 0]
    return c[2, 2, :, 0]


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("lim", [1, 2, 3])
@pytest.mark.parametrize("tm", [1, 2])
@pytest.mark.parametrize("stretched", [False, True])
def test_oracle_matches_literal_restatement(lim, tm, stretched):
    p = random_problem(6, 5, 4, 2, seed=100 + 10 * lim + tm, stretched=stretched)
    o = Oracle(p, limiter=lim, time_marching=tm)
    c1, c2 = p.new_c(), p.new_c()
    a1, a2 = p.new_acum(), p.new_acum()
    dt = 3.0
    o.solve_x(dt, c1, p.k1, p.u1)
    fx = [0.0, 0.0]
    L.solve_x(p, lim, tm, dt, c2, p.k1, p.u1, fx)
    assert np.array_equal(c1, c2)
    o.solve_y(dt, c1, p.k2, p.v1)
    fy = [0.0, 0.0]
    L.solve_y(p, lim, tm, dt, c2, p.k2, p.v1, fy)
    assert np.array_equal(c1, c2)
    o.solve_z(dt, c1, a1, p.k3, p.w1, p.w2, 0.37, p.vs)
    fz = [0.0, 0.0]
    L.solve_z(p, lim, tm, dt, c2, a2, p.k3, p.w1, p.w2, 0.37, fz)
    assert np.array_equal(c1, c2)
    assert np.array_equal(a1, a2)
    np.testing.assert_allclose(o.flux, np.array(fx + fy + fz), rtol=1e-13, atol=1e-300)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("lim", [1, 2, 3])
@pytest.mark.parametrize("dirn", [0, 1, 2])
@pytest.mark.parametrize("tm", [1, 2])
def test_linear_profile_is_translated(lim, dirn, tm):
    """c_i = a + b i, uniform u > 0, k = 0: every limiter gives phi(1) = 1, the
    reconstruction is the face midpoint and RHS = -u b / dx; RK4 of a
    constant RHS moves the profile by -u b dt / dx. Cells whose stencil cone
    reaches a ghost are excluded (2 per stage on each side)."""
    n, dx, uu, dt = 40, 1000.0, 7.0, 20.0
    a, b = 3.0, 0.25
    c_line = a + b * np.arange(-1, n + 3, dtype=np.float64)
    u_faces = np.full(n + 3, uu)
    p = _line_problem(n, c_line, u_faces, np.zeros(n + 4), dx=dx, dirn=dirn)
    o = Oracle(p, limiter=lim, time_marching=tm)
    c = p.new_c()
    _sweep(o, p, dirn, dt, c)
    out = _line(c, dirn)
    m = 2 if tm == 1 else 8
    idx = np.arange(1 + m, n + 1 - m)  # interior positions unaffected by the ghosts
    expect = c_line[idx + 1] - uu * b * dt / dx
    np.testing.assert_allclose(out[idx + 1], expect, rtol=0, atol=1e-13 * np.abs(expect).max())


# ---------------------------------------------------------------------------
def _diffusion_matrix(n, kcell, dxp, dxb):
    """RHS of pure diffusion with frozen ghost values g: k(y) = T y + e1*gl*a1 + en*gr*an
    built from the face coefficients D_f = (k_{f-1} + k_f)/2/dxp_{f-1}."""
    D = np.array([(kcell[f - 1 + 1] + kcell[f + 1]) * 0.5 / dxp[f - 1 + 1] for f in range(1, n + 2)])
    T = np.zeros((n, n))
    for i in range(1, n + 1):
        Dm, Dp = D[i - 1], D[i]
        T[i - 1, i - 1] = -(Dm + Dp) / dxb[i]
        if i > 1:
            T[i - 1, i - 2] = Dm / dxb[i]
        if i < n:
            T[i - 1, i] = Dp / dxb[i]
    bl = D[0] / dxb[1]
    br = D[n] / dxb[n]
    return T, bl, br


@pytest.mark.parametrize("dirn", [0, 1, 2])
def test_diffusion_rk4_matrix_polynomial_with_frozen_ghosts(dirn):
    rng = np.random.default_rng(5 + dirn)
    n, dx, dt = 17, 800.0, 900.0
    kcell = rng.uniform(50.0, 150.0, n + 4)
    c_line = rng.uniform(0.0, 1.0, n + 4)
    p = _line_problem(n, c_line, np.zeros(n + 3), kcell, dx=dx, dirn=dirn)
    o = Oracle(p, limiter=1, time_marching=2)
    c = p.new_c()
    _sweep(o, p, dirn, dt, c)
    out = _line(c, dirn)[2:n + 2]
    dxp = np.full(n + 4, dx)
    dxb = np.full(n + 3, dx)
    T, bl, br = _diffusion_matrix(n, kcell, dxp, dxb)
    c0 = c_line[2:n + 2]
    gl, gr = c0[0], c0[-1]  # zero normal velocity -> zero-gradient ghosts, frozen for the sweep

    def k(y):
        r = T @ y
        r[0] += bl * gl
        r[-1] += br * gr
        return r

    k1 = k(c0)
    k2 = k(c0 + 0.5 * dt * k1)
    k3 = k(c0 + 0.5 * dt * k2)
    k4 = k(c0 + dt * k3)
    expect = c0 + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6.0
    np.testing.assert_allclose(out, expect, rtol=1e-12, atol=1e-14)


def test_diffusion_euler_neumann_eigenvector():
    """Euler with zero-gradient ghosts: cos(m pi (i-1/2)/n) is an eigenvector of
    the discrete Neumann Laplacian, lambda = -4 K/dx^2 sin^2(m pi / 2n)."""
    n, dx, K, dt, m = 32, 500.0, 60.0, 300.0, 3
    i = np.arange(1, n + 1)
    e = np.cos(m * math.pi * (i - 0.5) / n)
    c_line = np.concatenate([[e[0], e[0]], e, [e[-1], e[-1]]])
    p = _line_problem(n, c_line, np.zeros(n + 3), np.full(n + 4, K), dx=dx, dirn=0)
    o = Oracle(p, limiter=1, time_marching=1)
    c = p.new_c()
    o.solve_x(dt, c, p.k1, p.u1)
    lam = -4.0 * K / dx ** 2 * math.sin(m * math.pi / (2 * n)) ** 2
    np.testing.assert_allclose(c[2:n + 2, 2, 2, 0], (1.0 + lam * dt) * e, rtol=0, atol=1e-13)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("lim", [1, 2, 3])
@pytest.mark.parametrize("tm", [1, 2])
def test_mirror_symmetry(lim, tm):
    """x -> -x with u -> -u maps the scheme onto itself (symmetric limiters,
    symmetric free-flow rule); data without plateaus (EPSILON branch is
    one-sided by construction)."""
    rng = np.random.default_rng(11)
    n = 23
    c_line = rng.uniform(0.1, 1.0, n + 4)
    u_f = rng.normal(0.0, 5.0, n + 3)
    kc = rng.uniform(10.0, 100.0, n + 4)
    p1 = _line_problem(n, c_line, u_f, kc)
    # mirrored: cell p -> n+1-p (array index q -> n+3-q), face f -> n+2-f
    p2 = _line_problem(n, c_line[::-1].copy(), -u_f[::-1].copy(), kc[::-1].copy())
    c1, c2 = p1.new_c(), p2.new_c()
    o1, o2 = Oracle(p1, lim, tm), Oracle(p2, lim, tm)
    o1.solve_x(10.0, c1, p1.k1, p1.u1)
    o2.solve_x(10.0, c2, p2.k1, p2.u1)
    np.testing.assert_allclose(c1[:, 2, 2, 0], c2[::-1, 2, 2, 0], rtol=1e-13, atol=1e-15)
    # boundary fluxes swap ends and change sign
    np.testing.assert_allclose(o1.flux[0], -o2.flux[1], rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("lim,r,phi", [
    (1, 0.5, 0.5), (1, 2.0, 1.0), (1, -1.0, 0.0),
    (2, 0.25, 0.5), (2, 0.75, 1.0), (2, 1.5, 1.5), (2, 3.0, 2.0), (2, -2.0, 0.0),
    (3, 1.0, 1.0), (3, 2.0, 1.5 * 6.0 / 7.0), (3, -0.5, -0.5),
])
def test_limiter_diagram_points(lim, r, phi):
    """One interior cell with d- = r, d+ = 1, u > 0 on the probed face: the
    upwind state of face i+1 is c_i + phi(r)/2 (Sweby diagram values; the
    reference's OSPRE is not clipped for r < 0)."""
    n, dx, uu, dt = 12, 1.0, 1.0, 1e-3
    c_line = np.zeros(n + 4)
    i0 = 6  # cell position
    # c_{i0-1} = 0, c_{i0} = r, c_{i0+1} = r + 1 -> d- = r, d+ = 1
    c_line[i0 + 1 - 1] = 0.0
    c_line[i0 + 1] = r
    c_line[i0 + 1 + 1] = r + 1.0
    u_f = np.zeros(n + 3)
    u_f[i0 + 1] = uu  # face between cells i0 and i0+1 only
    p = _line_problem(n, c_line, u_f, np.zeros(n + 4), dx=dx)
    o = Oracle(p, limiter=lim, time_marching=1)
    c = p.new_c()
    o.solve_x(dt, c, p.k1, p.u1)
    # Euler: cell i0 loses dt*u*cL/dx with cL = c_i0 + phi/2 * d+
    cL = (c_line[i0 + 1] - c[i0 + 1, 2, 2, 0]) * dx / (dt * uu)
    assert cL == pytest.approx(r + 0.5 * phi, rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("lim,dm,dp,phi", [
    (1, -1.0, 4e-13, 1.0),    # minmod: phi = 1 (the r-formula would give 0)
    (2, -1.0, 4e-13, 2.0),    # superbee: |d-| >= eps -> 2 (r-formula: 0)
    (2, 1e-13, 4e-13, 0.0),   # superbee: both below eps -> 0 (r-formula: 0.5)
    (3, 1e-13, 4e-13, 0.0),   # ospre: both below eps -> 0 (r-formula: 0.357)
    (3, 3e-12, -8e-13, 1.5),  # ospre: |d-| >= eps -> 1.5 (r-formula: 1.367)
])
def test_epsilon_branch(lim, dm, dp, phi):
    """|d+| < 1e-12 takes the reference's special branch
    (Sources/mod_ADS.F90:1140-1141,1171-1176,1207-1212). The probed cell holds
    c = 0 so the tiny reconstructed state is resolved to full precision."""
    n, dx, dt = 12, 1.0, 1e-3
    c_line = np.zeros(n + 4)
    i0 = 6
    c_line[i0 + 1 - 1] = -dm
    c_line[i0 + 1] = 0.0
    c_line[i0 + 1 + 1] = dp
    u_f = np.zeros(n + 3)
    u_f[i0 + 1] = 1.0
    p = _line_problem(n, c_line, u_f, np.zeros(n + 4), dx=dx)
    o = Oracle(p, limiter=lim, time_marching=1)
    c = p.new_c()
    o.solve_x(dt, c, p.k1, p.u1)
    cL = -c[i0 + 1, 2, 2, 0] * dx / dt
    assert cL == pytest.approx(0.5 * phi * dp, rel=1e-12, abs=1e-40)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("tm", [1, 2])
def test_conservation_and_ledger(tm):
    """Zero-flux laterals/top (zero normal velocity) and an absorbing ground:
    domain mass + ledger outflow is constant to FP64 rounding; with Euler the
    reference ground flux equals the ledger (single stage, zero-gradient
    outflow ghost -> no diffusive flux)."""
    p = puff_problem(16, 14, 10, dx=5000.0, dy=5000.0, dz=500.0)
    o = Oracle(p, limiter=1, time_marching=tm)
    c, acum = p.new_c(), p.new_acum()
    m0 = o.mass_cells(c)
    dt = o.time_step(p.u1, p.v1, p.w1, p.vs)
    for it in range(12):
        o.step(dt, it, c, acum)
    m = o.mass_cells(c)
    out = o.ledger.sum()
    assert abs(m + out - m0) <= 1e-13 * m0
    assert out > 0.0
    if tm == 1:
        # zero normal velocity and zero-gradient ghosts: nothing crosses the laterals/top
        np.testing.assert_allclose(o.ledger[:4], 0.0, atol=1e-20)
        assert o.ledger[5] == 0.0
    # with RK4 the ghosts stay frozen at the sweep's initial boundary values
    # (Sources/mod_ADS.F90:330), so intermediate stages see a diffusive flux
    # through the "zero-flux" faces; the ledger accounts for it exactly
    if tm == 1:
        assert o.flux[4] == pytest.approx(-o.ledger[4], rel=1e-12)
    # deposit map integrates to the ground flux
    g = p.grid
    dep = (acum[2:-2, 2:-2, 0] * g["dX1_b"][1:-2, None] * g["dX2_b"][None, 1:-2]).sum()
    if tm == 1:
        assert dep == pytest.approx(abs(o.flux[4]), rel=1e-12)


def test_constant_state_preserved_interior():
    n = 20
    c_line = np.full(n + 4, 2.5)
    u_f = np.full(n + 3, 3.0)
    u_f[1] = 0.0
    u_f[n + 1] = 0.0
    p = _line_problem(n, c_line, u_f, np.full(n + 4, 40.0))
    for lim in (1, 2, 3):
        o = Oracle(p, lim, 2)
        c = p.new_c()
        o.solve_x(5.0, c, p.k1, p.u1)
        line = c[2:n + 2, 2, 2, 0]
        np.testing.assert_allclose(line[9:-9], 2.5, rtol=0, atol=1e-14)
        # mass change of the 4 identical lines = what the ledger says left
        area = 1000.0 * 1000.0
        moved = (line.sum() - 2.5 * n) * 1000.0 * area * 4
        assert moved == pytest.approx(-o.ledger[0:2].sum(), rel=1e-5)  # sum of ~2.5 values: cancellation


# ---------------------------------------------------------------------------
def test_time_step_uses_west_faces():
    nx, ny, nz, nb = 5, 4, 3, 2
    g = uniform_grid(nx, ny, nz, 1000.0, 2000.0, 100.0)
    u = np.zeros((nx + 3, ny, nz), order="F")
    u[:, :, :] = np.arange(nx + 3, dtype=np.float64)[:, None, None]  # |u| largest on the east face nx+1
    v = np.zeros((ny + 3, nx, nz), order="F"); v[...] = 2.0
    w = np.zeros((nz + 3, nx, ny), order="F"); w[...] = 0.1
    k1, k2, k3 = constant_diffusion(nx, ny, nz, 10.0, 1.0)
    vs = np.zeros((nx, ny, nz + 3, nb), order="F"); vs[..., 0] = 0.3; vs[..., 1] = -0.2
    c = np.zeros((nx + 4, ny + 4, nz + 4, nb), order="F")
    p = Problem(nx, ny, nz, nb, g, u, u, v, v, w, w, k1, k2, k3, vs, c)
    o = Oracle(p)
    # criterion 2: west faces 1..nx -> max |u| = nx; max_b |w - vs| = 0.3
    inv = 2 * 10 / 1e6 + nx / 1000.0 + 2 * 10 / 4e6 + 2.0 / 2000.0 + 2 * 1.0 / 1e4 + 0.3 / 100.0
    assert o.time_step(u, v, w, vs, criterion=2, cfl=0.9) == pytest.approx(0.9 / inv, rel=1e-14)
    inv1 = max(2 * 10 / 1e6 + nx / 1000.0, 2 * 10 / 4e6 + 2.0 / 2000.0, 2 * 1.0 / 1e4 + 0.3 / 100.0)
    assert o.time_step(u, v, w, vs, criterion=1, cfl=0.5) == pytest.approx(0.5 / inv1, rel=1e-14)


def test_source_and_mass_volume():
    p = random_problem(4, 3, 5, 2, seed=3)
    o = Oracle(p)
    c = p.new_c()
    before = c.copy()
    inj = o.add_source(7.0, c, p.s)
    g = p.grid
    vol = g["dX1_p"][2:-2, None, None] * g["dX2_p"][None, 2:-2, None] * g["dX3_p"][None, None, 2:-2]
    expect = before[2:-2, 2:-2, 2:-2, :] + 7.0 * p.s / vol[..., None]
    np.testing.assert_allclose(c[2:-2, 2:-2, 2:-2, :], expect, rtol=1e-15)
    assert inj == pytest.approx(7.0 * p.s.sum(), rel=1e-13)
    m = o.mass_volume(c)
    assert m == pytest.approx((c[2:-2, 2:-2, 2:-2, :] * vol[..., None]).sum(), rel=1e-13)


@pytest.mark.parametrize("lim", [1, 2, 3])
def test_solve_1d_matches_sweep_on_one_line(lim):
    """ADS_solve_1D with free-flow codes integrates a line exactly like a sweep
    (same ghosts, same stages; the RK4 combination differs only by rounding)."""
    rng = np.random.default_rng(21 + lim)
    n = 19
    c_line = rng.uniform(0.0, 1.0, n + 4)
    u_f = rng.normal(0.0, 3.0, n + 3)
    kc = rng.uniform(5.0, 50.0, n + 4)
    p = _line_problem(n, c_line, u_f, kc, dx=700.0)
    o = Oracle(p, lim, 2)
    c = p.new_c()
    o.solve_x(40.0, c, p.k1, p.u1)
    cl = c_line.copy()
    f1, f2 = ora_solve_1d(cl, np.full(n + 4, 700.0), np.full(n + 3, 700.0), u_f.copy(), kc.copy(), 40.0,
                          (0.0, 0.0), (2, 2), lim)
    np.testing.assert_allclose(cl, c[:, 2, 2, 0], rtol=1e-13, atol=1e-15)
    area = 700.0 * 700.0
    assert f1 * area * 40.0 * 4 == pytest.approx(o.flux[0], rel=1e-12, abs=1e-12)
