"""Pins of the oracle's boundary conventions, its stretched-grid metric roles
and ADS_solve_1D's Dirichlet and periodic branches.

Each expected value below comes from the reference text plus mathematics,
not from the oracle's code path:

  * free-flow inflow ghost = 2*bvalue - c(first) with bvalue = 0
    (Sources/mod_ADS.F90:117,1243-1248): with constant c and K and a tiny
    inflow velocity on one boundary face, the boundary cell sees the
    diffusive flux of a zero Dirichlet face value, 2Kc/dxp, and nothing
    else moves; a zero or outflow velocity gives a zero-gradient ghost and
    no change at all (strict inequality, :1243,1256);
  * the reference accumulators add F + P of the end faces
    (Sources/mod_ADS.F90:1116-1117) times dt and the face area, from the
    LAST stage only (:394-396); with u = K/dxp the F + P and F - P
    conventions differ by a factor -3;
  * the roles of dxp(i-1) (distance of the two mass points of a face) and
    dxb(i) (size of cell i) on stretched axes: a profile linear in the
    mass-point position carries the same diffusive flux K*b through every
    face, so constant K leaves it steady and variable K gives
    b (K(i+1) - K(i-1)) / (2 dxb(i)); the divergence form conserves the
    dxb-weighted sum exactly;
  * the RK4 stage values of pure diffusion with frozen ghosts are a matrix
    polynomial (dense linear algebra, not the staged solver), which fixes
    the last-stage boundary flux and the RK-weighted ledger;
  * ADS_solve_1D (Sources/mod_ADS.F90:820-1053): Dirichlet ghosts
    2 b - c(first) and periodic wrap-around ghosts, both frozen for the
    four stages.
"""
import numpy as np
import pytest

from fall3d_amd.problem import random_problem
from oracle.ads_oracle import Oracle, solve_1d as ora_solve_1d
from test_oracle_pins import _diffusion_matrix, _line, _line_problem, _sweep


def _metrics(dirn, n, rng, stretched):
    """(dxp, dxb) of the sweep axis: index 0 of dxp is position -1, index 0
    of dxb is position 0 (the reference's halo bounds)."""
    if not stretched:
        return np.full(n + 4, 1000.0), np.full(n + 3, 1000.0)
    return rng.uniform(400.0, 1600.0, n + 4), rng.uniform(400.0, 1600.0, n + 3)


def _with_metrics(p, dirn, dxp, dxb):
    p.grid[f"dX{dirn + 1}_p"] = np.asarray(dxp, dtype=np.float64)
    p.grid[f"dX{dirn + 1}_b"] = np.asarray(dxb, dtype=np.float64)
    return p


def _rk4_frozen(T, bl, br, gl, gr, c0, dt):
    """Stage inputs y0..y3 and the RK4 result of y' = T y + e1 bl gl + en br gr
    with the ghost values gl, gr held fixed."""
    def rhs(y):
        r = T @ y
        r[0] += bl * gl
        r[-1] += br * gr
        return r

    y0 = c0
    k1 = rhs(y0)
    y1 = c0 + 0.5 * dt * k1
    k2 = rhs(y1)
    y2 = c0 + 0.5 * dt * k2
    k3 = rhs(y2)
    y3 = c0 + dt * k3
    k4 = rhs(y3)
    return (y0, y1, y2, y3), c0 + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6.0


AREA = 1000.0 * 1000.0  # transverse face area of the line problems
NLINES = 4              # the line is repeated on 2 x 2 transverse positions


# ---------------------------------------------------------------------------
# (i) free-flow inflow ghost
@pytest.mark.parametrize("dirn", [0, 1, 2])
@pytest.mark.parametrize("side", ["first", "last"])
@pytest.mark.parametrize("unorm", [1e-30, 0.0, -1e-30])
def test_free_flow_inflow_ghost_is_zero_dirichlet(dirn, side, unorm):
    n, c, K, dx, dt = 10, 2.0, 50.0, 1000.0, 100.0
    u = np.zeros(n + 3)
    # inflow: u > 0 on the first face, u < 0 on the last face
    inflow = unorm > 0.0
    if side == "first":
        u[1] = unorm
    else:
        u[n + 1] = -unorm
    p = _line_problem(n, np.full(n + 4, c), u, np.full(n + 4, K), dx=dx, dirn=dirn)
    o = Oracle(p, limiter=1, time_marching=1)
    cc, acum = p.new_c(), p.new_acum()
    _sweep(o, p, dirn, dt, cc, acum)
    out = _line(cc, dirn)[2:n + 2]
    if not inflow:
        # zero-gradient ghosts: nothing moves, no flux through the face
        assert np.array_equal(out, np.full(n, c))
        # only the advective flux u c of the outflow case (|u| = 1e-30) crosses the face
        np.testing.assert_allclose(o.flux, 0.0, atol=2.0 * abs(unorm) * c * dt * AREA * NLINES)
        return
    # ghost -c: face value 0, diffusive flux 2Kc/dxp out of the boundary cell
    drop = dt * 2.0 * K * c / (dx * dx)
    b = 0 if side == "first" else n - 1
    expect = np.full(n, c)
    expect[b] = c - drop
    np.testing.assert_allclose(out, expect, rtol=1e-14, atol=0)
    # reference accumulator of that face: F + P = 2Kc/dxp (F is ~1e-30)
    ref = 2.0 * K * c / dx
    fi = 2 * dirn + (0 if side == "first" else 1)
    sgn = 1.0 if side == "first" else -1.0  # P_p at the last face = K (gR - c)/dxp = -2Kc/dxp
    assert o.flux[fi] == pytest.approx(sgn * ref * dt * AREA * NLINES, rel=1e-14)
    # ledger: the mass that left through the face
    assert o.ledger[fi] == pytest.approx(ref * dt * AREA * NLINES, rel=1e-14)
    if dirn == 2 and side == "first":
        # ground deposit: |flux(1)| dt per unit area
        np.testing.assert_allclose(acum[2:4, 2:4, 0], ref * dt, rtol=1e-14)


# ---------------------------------------------------------------------------
# (ii) the +P sign of the reference accumulators, and F - P in the ledger
@pytest.mark.parametrize("dirn", [0, 1, 2])
def test_reference_flux_is_F_plus_P(dirn):
    n, c, K, dx, dt = 12, 2.0, 50.0, 1000.0, 100.0
    uu = K / dx  # advective and diffusive parts of the first face then differ in sign and size
    p = _line_problem(n, np.full(n + 4, c), np.full(n + 3, uu), np.full(n + 4, K), dx=dx, dirn=dirn)
    o = Oracle(p, limiter=1, time_marching=1)
    cc, acum = p.new_c(), p.new_acum()
    m0 = o.mass_cells(cc)
    _sweep(o, p, dirn, dt, cc, acum)
    out = _line(cc, dirn)[2:n + 2]
    # first face (inflow, ghost -c): upwind F = u * (-c), diffusive P = K (c - (-c))/dxp
    F1, P1 = -uu * c, 2.0 * K * c / dx
    # last face (outflow, zero gradient): F = u c, P = 0; interior faces: F = u c, P = 0
    Fn = uu * c
    expect = np.full(n, c)
    expect[0] = c + dt * (-(Fn - F1) + (0.0 - P1)) / dx
    np.testing.assert_allclose(out, expect, rtol=1e-14, atol=0)
    fb, fe = o.flux[2 * dirn], o.flux[2 * dirn + 1]
    assert fb == pytest.approx((F1 + P1) * dt * AREA * NLINES, rel=1e-14)  # +K c/dxp, not F - P = -3 K c/dxp
    assert fb > 0.0
    assert fe == pytest.approx(Fn * dt * AREA * NLINES, rel=1e-14)
    # ledger (mass out per side) = -(F - P) at the first face, F - P at the last
    assert o.ledger[2 * dirn] == pytest.approx(-(F1 - P1) * dt * AREA * NLINES, rel=1e-14)
    assert o.ledger[2 * dirn + 1] == pytest.approx(Fn * dt * AREA * NLINES, rel=1e-14)
    assert o.mass_cells(cc) == pytest.approx(m0 - o.ledger[2 * dirn:2 * dirn + 2].sum(), rel=1e-14)
    if dirn == 2:
        np.testing.assert_allclose(acum[2:4, 2:4, 0], abs(F1 + P1) * dt, rtol=1e-14)


# ---------------------------------------------------------------------------
# (ii, Q3, Q4, iii) last-stage accumulators and RK-weighted ledger under RK4,
# frozen ghosts, stretched metrics
@pytest.mark.parametrize("dirn", [0, 1, 2])
@pytest.mark.parametrize("stretched", [False, True])
def test_rk4_last_stage_flux_and_weighted_ledger(dirn, stretched):
    rng = np.random.default_rng(40 + 3 * dirn + stretched)
    n, dt = 15, 600.0
    kcell = rng.uniform(40.0, 120.0, n + 4)
    c_line = rng.uniform(0.2, 1.0, n + 4)
    u = np.zeros(n + 3)
    u[1] = 1e-30  # inflow on the first face: ghost -c(1); the last face is closed (zero gradient)
    dxp, dxb = _metrics(dirn, n, rng, stretched)
    p = _with_metrics(_line_problem(n, c_line, u, kcell, dirn=dirn), dirn, dxp, dxb)
    o = Oracle(p, limiter=1, time_marching=2)
    cc, acum = p.new_c(), p.new_acum()
    _sweep(o, p, dirn, dt, cc, acum)
    out = _line(cc, dirn)[2:n + 2]
    T, bl, br = _diffusion_matrix(n, kcell, dxp, dxb)
    c0 = c_line[2:n + 2]
    gl, gr = -c0[0], c0[-1]
    ys, expect = _rk4_frozen(T, bl, br, gl, gr, c0, dt)
    np.testing.assert_allclose(out, expect, rtol=1e-12, atol=1e-15)
    D1 = 0.5 * (kcell[1] + kcell[2]) / dxp[1]                # face 1: cells 0 and 1 over dxp(0)
    Dn = 0.5 * (kcell[n + 1] + kcell[n + 2]) / dxp[n + 1]    # face n+1: cells n, n+1 over dxp(n)
    ybar = (ys[0] + 2 * ys[1] + 2 * ys[2] + ys[3]) / 6.0
    # reference: F + P of the last stage (stage-4 input y3); F ~ 1e-30
    assert o.flux[2 * dirn] == pytest.approx(D1 * (ys[3][0] - gl) * dt * AREA * NLINES, rel=1e-12)
    assert o.flux[2 * dirn + 1] == pytest.approx(Dn * (gr - ys[3][-1]) * dt * AREA * NLINES, rel=1e-10,
                                                 abs=1e-12 * abs(o.flux[2 * dirn]))
    # ledger: RK-weighted mass that left through each face
    assert o.ledger[2 * dirn] == pytest.approx(D1 * (ybar[0] - gl) * dt * AREA * NLINES, rel=1e-12)
    assert o.ledger[2 * dirn + 1] == pytest.approx(Dn * (ybar[-1] - gr) * dt * AREA * NLINES, rel=1e-10,
                                                   abs=1e-12 * abs(o.ledger[2 * dirn]))
    # the last-stage and RK-weighted values differ: the test tells them apart
    assert abs(ys[3][0] - ybar[0]) > 1e-6 * abs(ybar[0] - gl)


# ---------------------------------------------------------------------------
# (iii) dxp / dxb roles on stretched axes
@pytest.mark.parametrize("dirn", [0, 1, 2])
@pytest.mark.parametrize("tm", [1, 2])
def test_profile_linear_in_position_is_steady(dirn, tm):
    """c = a + b x at mass points x_{i+1} - x_i = dxp(i): every face carries
    the diffusive flux K b, so constant K leaves the interior unchanged."""
    rng = np.random.default_rng(70 + dirn)
    n, K, dt = 20, 80.0, 500.0
    dxp, dxb = _metrics(dirn, n, rng, True)
    x = np.concatenate([[0.0], np.cumsum(dxp[:-1])])  # positions of -1..n+2
    c_line = 1.0 + 2e-4 * x
    p = _with_metrics(_line_problem(n, c_line, np.zeros(n + 3), np.full(n + 4, K), dirn=dirn), dirn, dxp, dxb)
    o = Oracle(p, limiter=1, time_marching=tm)
    cc = p.new_c()
    _sweep(o, p, dirn, dt, cc)
    out = _line(cc, dirn)
    m = 1 if tm == 1 else 4  # cells whose RK cone reaches a ghost
    idx = np.arange(1 + m, n + 1 - m) + 1
    np.testing.assert_allclose(out[idx], c_line[idx], rtol=1e-14, atol=0)


@pytest.mark.parametrize("dirn", [0, 1, 2])
def test_linear_profile_variable_k_closed_form(dirn):
    """Linear in position with variable K: face flux (K_{f-1} + K_f)/2 * b,
    so one Euler step adds dt b (K(i+1) - K(i-1)) / (2 dxb(i))."""
    rng = np.random.default_rng(80 + dirn)
    n, dt, b = 18, 300.0, 3e-4
    dxp, dxb = _metrics(dirn, n, rng, True)
    x = np.concatenate([[0.0], np.cumsum(dxp[:-1])])
    c_line = 0.5 + b * x
    kc = rng.uniform(20.0, 200.0, n + 4)
    p = _with_metrics(_line_problem(n, c_line, np.zeros(n + 3), kc, dirn=dirn), dirn, dxp, dxb)
    o = Oracle(p, limiter=2, time_marching=1)
    cc = p.new_c()
    _sweep(o, p, dirn, dt, cc)
    out = _line(cc, dirn)
    i = np.arange(2, n)  # interior cells (the end cells see the zero-gradient ghosts)
    expect = c_line[i + 1] + dt * b * (kc[i + 2] - kc[i]) / (2.0 * dxb[i])
    np.testing.assert_allclose(out[i + 1], expect, rtol=1e-13, atol=0)


@pytest.mark.parametrize("dirn", [0, 1, 2])
@pytest.mark.parametrize("lim", [1, 2, 3])
def test_divergence_form_conserves_dxb_measure(dirn, lim):
    """Closed end faces (zero normal velocity, zero-gradient ghosts), random
    interior winds, K and stretched metrics: one Euler step keeps
    sum dxb(i) c(i) to rounding (the RHS of cell i is divided by its own
    size dxb(i)); with RK4 the frozen ghosts let some mass through the ends
    and sum dxb c * area + ledger is what stays constant."""
    rng = np.random.default_rng(90 + 3 * dirn + lim)
    n = 25
    dxp, dxb = _metrics(dirn, n, rng, True)
    u = rng.normal(0.0, 4.0, n + 3)
    u[1] = 0.0
    u[n + 1] = 0.0
    kc = rng.uniform(10.0, 100.0, n + 4)
    c_line = rng.uniform(0.0, 1.0, n + 4)
    for tm in (1, 2):
        p = _with_metrics(_line_problem(n, c_line, u, kc, dirn=dirn), dirn, dxp, dxb)
        o = Oracle(p, limiter=lim, time_marching=tm)
        cc = p.new_c()
        m0 = float(np.dot(dxb[1:n + 1], c_line[2:n + 2]))
        _sweep(o, p, dirn, 20.0, cc)
        m1 = float(np.dot(dxb[1:n + 1], _line(cc, dirn)[2:n + 2]))
        if tm == 1:
            assert m1 == pytest.approx(m0, rel=1e-14)
        else:
            out = o.ledger[2 * dirn:2 * dirn + 2].sum() / (AREA * NLINES)
            assert m1 + out == pytest.approx(m0, rel=1e-14)


@pytest.mark.parametrize("stretched", [False, True])
def test_source_and_mass_volume_use_mass_point_spacings(stretched):
    """Q13 and the source add (Sources/mod_F3D.F90:1456-1473,
    mod_Grid.F90:1364-1375) divide and weigh by dX_p; on a stretched axis
    dX3_p and dX3_b differ, so the measure is told apart."""
    p = random_problem(4, 3, 6, 2, seed=8, stretched=stretched)
    g = p.grid
    if stretched:
        assert not np.allclose(g["dX3_p"][2:-2], g["dX3_b"][1:-2])
    o = Oracle(p)
    c = p.new_c()
    before = c.copy()
    o.add_source(5.0, c, p.s)
    vol = g["dX1_p"][2:-2, None, None] * g["dX2_p"][None, 2:-2, None] * g["dX3_p"][None, None, 2:-2]
    np.testing.assert_allclose(c[2:-2, 2:-2, 2:-2, :], before[2:-2, 2:-2, 2:-2, :] + 5.0 * p.s / vol[..., None],
                               rtol=1e-15)
    assert o.mass_volume(c) == pytest.approx(float((c[2:-2, 2:-2, 2:-2, :] * vol[..., None]).sum()), rel=1e-13)


def test_source_with_map_factors():
    """The source divides by dX1_p dX2_p dX3_p Hm1_p(j) Hm2_p(j) Hm3_p(i,j)
    (Sources/mod_F3D.F90:1458-1470)."""
    p = random_problem(5, 4, 3, 2, seed=12)
    rng = np.random.default_rng(3)
    hm1 = rng.uniform(0.3, 1.0, p.ny)
    hm2 = np.ones(p.ny + 4)
    hm3 = rng.uniform(0.6, 1.0, (p.nx, p.ny))
    o = Oracle(p)
    c = p.new_c()
    before = c.copy()
    o.add_source(3.0, c, p.s, hm1=hm1, hm2=hm2, hm3=hm3)
    g = p.grid
    vol = (g["dX1_p"][2:-2, None, None] * g["dX2_p"][None, 2:-2, None] * g["dX3_p"][None, None, 2:-2]
           * (hm1[None, :] * hm2[2:-2][None, :] * hm3)[:, :, None])
    np.testing.assert_allclose(c[2:-2, 2:-2, 2:-2, :], before[2:-2, 2:-2, 2:-2, :] + 3.0 * p.s / vol[..., None],
                               rtol=1e-15)


# ---------------------------------------------------------------------------
# (iv) ADS_solve_1D: Dirichlet and periodic branches
def _line1d(n, rng, stretched=True):
    dxp, dxb = _metrics(0, n, rng, stretched)
    return rng.uniform(0.1, 1.0, n + 4), dxp, dxb


def test_solve_1d_dirichlet_constant_boundary_value_is_steady():
    rng = np.random.default_rng(1)
    n, bval = 14, 0.7
    _, dxp, dxb = _line1d(n, rng)
    c = np.full(n + 4, bval)
    k = rng.uniform(10.0, 80.0, n + 4)
    for uu in (2.5, -1.7, 0.0):  # uniform (divergence-free) wind: c = b is a steady state
        for lim in (1, 2, 3):
            cl = c.copy()
            ora_solve_1d(cl, dxp, dxb, np.full(n + 3, uu), k, 50.0, (bval, bval), (1, 1), lim)
            np.testing.assert_allclose(cl, bval, rtol=1e-15, atol=0)


@pytest.mark.parametrize("bvals", [(0.3, -0.2), (0.0, 0.0), (1.5, 0.9)])
def test_solve_1d_dirichlet_diffusion_closed_form(bvals):
    """Ghost 2 b - c(first), frozen for the four stages: the dense-matrix RK4
    polynomial with those ghost values; flux = F + P of the last stage."""
    rng = np.random.default_rng(2)
    n, dt = 16, 400.0
    c, dxp, dxb = _line1d(n, rng)
    k = rng.uniform(20.0, 90.0, n + 4)
    cl = c.copy()
    f1, f2 = ora_solve_1d(cl, dxp, dxb, np.zeros(n + 3), k, dt, bvals, (1, 1), 1)
    T, bl, br = _diffusion_matrix(n, k, dxp, dxb)
    gl, gr = 2 * bvals[0] - c[2], 2 * bvals[1] - c[n + 1]
    ys, expect = _rk4_frozen(T, bl, br, gl, gr, c[2:n + 2], dt)
    np.testing.assert_allclose(cl[2:n + 2], expect, rtol=1e-12, atol=1e-15)
    # ghosts as set by the rule (second ghost copies the first)
    np.testing.assert_array_equal(cl[:2], [gl, gl])
    np.testing.assert_array_equal(cl[n + 2:], [gr, gr])
    D1 = 0.5 * (k[1] + k[2]) / dxp[1]
    Dn = 0.5 * (k[n + 1] + k[n + 2]) / dxp[n + 1]
    assert f1 == pytest.approx(D1 * (ys[3][0] - gl), rel=1e-12)
    assert f2 == pytest.approx(Dn * (gr - ys[3][-1]), rel=1e-12)


def test_solve_1d_periodic_diffusion_closed_form():
    """bcode -1 with inflow on the first face (u > 0): ghosts wrap around,
    c(0) = c(n) and c(n+1) = c(1), frozen for the four stages."""
    rng = np.random.default_rng(4)
    n, dt = 13, 350.0
    c, dxp, dxb = _line1d(n, rng)
    k = rng.uniform(20.0, 90.0, n + 4)
    u = np.zeros(n + 3)
    u[1] = 1e-30
    cl = c.copy()
    ora_solve_1d(cl, dxp, dxb, u, k, dt, (0.0, 0.0), (-1, -1), 1)
    T, bl, br = _diffusion_matrix(n, k, dxp, dxb)
    _, expect = _rk4_frozen(T, bl, br, c[n + 1], c[2], c[2:n + 2], dt)
    np.testing.assert_allclose(cl[2:n + 2], expect, rtol=1e-12, atol=1e-15)
    # all four ghosts wrapped: c(-1) = c(n-1), c(0) = c(n), c(n+1) = c(1), c(n+2) = c(2)
    np.testing.assert_array_equal(cl[[0, 1, n + 2, n + 3]], c[[n, n + 1, 2, 3]])


@pytest.mark.parametrize("ufirst,ulast,wrap", [(2.0, 1.0, True), (-1.0, -2.0, True), (-1.0, 2.0, False),
                                               (0.0, 0.0, False)])
def test_solve_1d_periodic_ghost_rule(ufirst, ulast, wrap):
    """Periodic ghosts are set only when the first face flows in (u > 0) or
    the last face flows in (u < 0); otherwise the ghosts are left as they
    are. Either way the line then integrates exactly like bcode 0 (ghosts
    untouched) on the same ghost values, with advection through every face."""
    rng = np.random.default_rng(6)
    n, dt = 17, 30.0
    c, dxp, dxb = _line1d(n, rng)
    u = rng.normal(0.0, 3.0, n + 3)
    u[1], u[n + 1] = ufirst, ulast
    k = rng.uniform(5.0, 40.0, n + 4)
    for lim in (1, 2, 3):
        a = c.copy()
        fa = ora_solve_1d(a, dxp, dxb, u, k, dt, (0.0, 0.0), (-1, -1), lim)
        b = c.copy()
        if wrap:
            b[[0, 1, n + 2, n + 3]] = c[[n, n + 1, 2, 3]]
        fb = ora_solve_1d(b, dxp, dxb, u, k, dt, (0.0, 0.0), (0, 0), lim)
        np.testing.assert_array_equal(a, b)
        assert fa == fb


def test_solve_1d_dirichlet_second_ghost_and_advection():
    """Dirichlet ghosts (2 b - c(first), second ghost a copy) with advection
    in both directions integrate exactly like bcode 0 on those ghosts."""
    rng = np.random.default_rng(9)
    n, dt = 15, 30.0
    c, dxp, dxb = _line1d(n, rng)
    u = rng.normal(0.0, 3.0, n + 3)
    k = rng.uniform(5.0, 40.0, n + 4)
    bv = (0.25, -0.4)
    for lim in (1, 2, 3):
        a = c.copy()
        fa = ora_solve_1d(a, dxp, dxb, u, k, dt, bv, (1, 1), lim)
        b = c.copy()
        b[0] = b[1] = 2 * bv[0] - c[2]
        b[n + 2] = b[n + 3] = 2 * bv[1] - c[n + 1]
        fb = ora_solve_1d(b, dxp, dxb, u, k, dt, (0.0, 0.0), (0, 0), lim)
        np.testing.assert_array_equal(a, b)
        assert fa == fb
