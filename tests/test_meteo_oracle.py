"""Pins of the meteo-update restatement (oracle/meteo_oracle.py) on the CPU:
each expected value comes from a closed form or an independent computation,
not from the restatement's own statements.

* staggering and corner->mass averages of fields linear in the indices are
  the exact face/centre values; the ghost faces copy the first/last face;
* CMAQ and RAMS horizontal diffusion on uniform-shear winds equal their
  closed forms in the shear rates;
* the similarity profile inside the boundary layer equals
  kappa z u* (1 - z/h)^2 / phi_h with the Businger-Dyer phi_h, and above it
  the mixing-length form with the Richardson stability function, for
  columns built to give known z, h, L, shear and Ri;
* the drag laws: tiny spheres settle at Stokes' velocity g d^2 rho_p /
  (18 mu) (models whose sphere limit is Stokes), large ones at the Newton
  plateau where the law has one (Arastopour cd = 0.44, Pfeiffer cd = 1);
  every iterative law's result is the root of its own fixed-point equation
  found by an independent bracketing solver; Sutherland's viscosity equals
  visa0 at Ta0;
* metric scaling: Cartesian maps leave the fields unscaled; on sigma maps
  the terrain correction (w - hx u - hy v) / Hm3 on uniform winds.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

from fall3d_amd.problem import meteo_corner_problem
from oracle import meteo_oracle as M


def _lin(nx, ny, nz, a, b, c, d):
    i = np.arange(1, nx + 2)[:, None, None]
    j = np.arange(1, ny + 2)[None, :, None]
    k = np.arange(1, nz + 2)[None, None, :]
    return np.asfortranarray(a + b * i + c * j + d * k + 0.0 * (i + j + k))


def test_stagger_and_c2p_of_linear_fields_are_exact():
    nx, ny, nz = 6, 5, 4
    a, b, c, d = 1.5, 0.25, -0.5, 2.0
    f = _lin(nx, ny, nz, a, b, c, d)
    u = M.stagger_u(f)
    for ii in range(nx + 3):
        i = min(max(ii, 1), nx + 1)
        for j in range(1, ny + 1):
            for k in range(1, nz + 1):
                assert u[ii, j - 1, k - 1] == a + b * i + c * (j + 0.5) + d * (k + 0.5)
    v = M.stagger_v(f)
    for i in range(1, nx + 1):
        for jj in range(ny + 3):
            j = min(max(jj, 1), ny + 1)
            assert v[i - 1, jj, 0] == a + b * (i + 0.5) + c * j + d * 1.5
    w = M.stagger_w(f)
    assert w[2, 3, 1] == a + b * 3.5 + c * 4.5 + d * 2
    cp = M.c2p(f)
    assert cp[1, 2, 3] == a + b * 2.5 + c * 3.5 + d * 4.5
    c2 = M.c2p_2d(f[:, :, 0])
    assert c2[4, 1] == a + b * 5.5 + c * 2.5 + d


@pytest.mark.parametrize("modkh", [2, 4])
def test_cmaq_and_rams_on_uniform_shear(modkh):
    """u = A x + B y, v = C x + D y on a uniform grid: the corner-difference
    derivatives are the shear rates."""
    nx, ny, nz = 5, 4, 3
    dx, dy = 2000.0, 1500.0
    A, B, C, D = 2e-4, -3e-4, 1e-4, 5e-4
    x = (np.arange(nx + 1) * dx)[:, None, None] + np.zeros((1, ny + 1, nz + 1))
    y = (np.arange(ny + 1) * dy)[None, :, None] + np.zeros((nx + 1, 1, nz + 1))
    uc, vc = A * x + B * y, C * x + D * y
    opts = {"modkh": modkh, "map_h": 0, "dlon": 3000.0, "dlat": 2500.0, "rams_cs": 0.3}
    kh = M.kh_points(opts, uc, vc, np.full(nx + 4, dx), np.full(ny + 4, dy), np.ones(ny), np.ones(ny + 4))
    if modkh == 2:
        kht = 0.28 ** 2 * dx * dy * np.hypot(A - D, C + B)
        khn = 8000.0 * (4000.0 / 3000.0) * (4000.0 / 2500.0)
        want = max(kht * khn / (kht + khn), 100.0)
    else:
        want = max(0.3 ** 2 * dx * dy * np.sqrt((B + C) ** 2 + 2 * (A * A + D * D)), 0.075 * (dx * dy) ** (2 / 3), 100.0)
    np.testing.assert_allclose(kh, want, rtol=1e-12)


def _similarity_column(z_mass, pblh, L, ust, shear, dtv):
    """One 1x1 column (2x2 corners) with flat terrain: corner heights such
    that the mass-point heights are z_mass, winds u = shear * z (v = 0),
    virtual temperature tv = 300 + dtv * z."""
    nz = len(z_mass)
    zc = np.empty(nz + 1)
    zc[0] = 0.0
    for k in range(nz):
        zc[k + 1] = 2 * z_mass[k] - zc[k]
    Z = np.broadcast_to(zc, (2, 2, nz + 1)).copy()
    uc = shear * Z
    vc = np.zeros_like(Z)
    tvc = 300.0 + dtv * Z
    flat = np.zeros((2, 2))
    opts = {"modkv": 3}
    k = M.kv_points(opts, uc, vc, tvc, np.full((2, 2), pblh), np.full((2, 2), ust), np.full((2, 2), L), Z, flat)
    return k[0, 0, :], zc


@pytest.mark.parametrize("L", [150.0, -80.0, 3.0])
def test_similarity_inside_boundary_layer(L):
    z = np.array([50.0, 150.0, 400.0])
    h, ust = 1000.0, 0.4
    k, _ = _similarity_column(z, h, L, ust, 0.01, 0.0)
    Le = L if abs(L) >= 10 else np.copysign(10.0, L)   # |L| < 10 is raised to 10
    zl = z / Le
    phi = np.where(zl < 0, np.abs(1 - 15 * zl) ** -0.5, 1 + 5 * zl)
    np.testing.assert_allclose(k, 0.4 * z * ust * (1 - z / h) ** 2 / phi, rtol=1e-13)


def test_similarity_above_boundary_layer_neutral_and_stable():
    z = np.array([600.0, 900.0, 1300.0])
    shear = 0.005
    k, zc = _similarity_column(z, 200.0, 100.0, 0.3, shear, 0.0)  # uniform tv: Ri = 0, f_c = 1
    lc = 1.0 / (1.0 / (0.4 * z) + 1.0 / 30.0)
    np.testing.assert_allclose(k, lc * lc * shear, rtol=1e-12)
    dtv = 0.01                                              # stable: Ri = g dz dtv / (umod^2 tv_mean)
    k2, _ = _similarity_column(z, 200.0, 100.0, 0.3, shear, dtv)
    dz = zc[1:] - zc[:-1]
    umod = 0.5 * shear * (zc[1:] + zc[:-1])
    tvm = 0.5 * ((300 + dtv * zc[1:]) + (300 + dtv * zc[:-1]))
    ri = 9.81 * dz * (dtv * dz) / (umod ** 2 * tvm)
    np.testing.assert_allclose(k2, lc * lc * shear / (1 + 10 * ri * (1 + 8 * ri)), rtol=1e-12)


def _stokes(d, rhop, mu):
    return 9.81 * d * d * rhop / (18.0 * mu)


@pytest.mark.parametrize("model", [1, 2, 3, 5, 7])
def test_tiny_spheres_settle_at_stokes_velocity(model):
    d, rhop, rhoa, mu = 2e-6, 2000.0, 1.2, 1.8e-5
    v, ok = M.vset_point(d, rhop, 1.0, np.array([rhoa]), np.array([mu]), model)
    assert ok.all()
    np.testing.assert_allclose(v, _stokes(d, rhop, mu), rtol=2e-3)


@pytest.mark.parametrize("model,cd", [(1, 0.44), (5, 1.0)])
def test_newton_plateau(model, cd):
    d, rhop, rhoa, mu = 8e-3, 2500.0, 1.2, 1.8e-5
    v, ok = M.vset_point(d, rhop, 0.8, np.array([rhoa]), np.array([mu]), model)
    assert ok.all()
    np.testing.assert_allclose(v, np.sqrt(4 * 9.81 * d * rhop / (3 * cd * rhoa)), rtol=1e-6)


@pytest.mark.parametrize("model", [1, 2, 3, 5, 6, 7])
@pytest.mark.parametrize("d", [5e-6, 1e-4, 2e-3])
def test_iteration_reaches_the_root_of_its_equation(model, d):
    """v = sqrt(4 g d rho_p / (3 cd(Re(v)) rho_a)) solved by bracketing: the
    iterate agrees within the iteration's own tolerance."""
    rhop, rhoa, mu, psi = 1800.0, 1.1, 1.75e-5, 0.85
    v, ok = M.vset_point(d, rhop, psi, np.array([rhoa]), np.array([mu]), model)
    assert ok.all()

    def g(x):
        # the fixed-point residual at velocity x with the published drag law
        rey = rhoa * x * d / mu
        return x - np.sqrt(4 * 9.81 * d * rhop / (3 * _cd(model, rey, psi) * rhoa))

    root = brentq(g, 1e-9, 200.0, xtol=1e-14, rtol=1e-14)
    np.testing.assert_allclose(v[0], root, rtol=5e-6)


def _cd(model, rey, psi):
    """Drag coefficients as published (Schiller-Naumann; Ganser 1993;
    Wilson & Huang 1979; Pfeiffer et al. 2005; Dioguardi et al. 2017, 2018),
    written independently of the restatement for the root check."""
    if model == 1:
        return 24 / rey * (1 + 0.15 * rey ** 0.687) if rey <= 988.947 else 0.44
    if model == 2:
        k1 = 3 / (1 + 2 * psi ** -0.5)
        k2 = 10 ** (1.8148 * (-np.log10(psi)) ** 0.5743)
        return 24 / (rey * k1) * (1 + 0.1118 * (rey * k1 * k2) ** 0.6567) + 0.4305 * k2 / (1 + 3305 / (rey * k1 * k2))
    if model == 3:
        return 24 / rey * psi ** -0.828 + 2 * np.sqrt(1.07 - psi)
    if model == 5:
        c100 = 0.24 * psi ** -0.828 + 2 * np.sqrt(1.07 - psi)
        if rey <= 100:
            return 24 / rey * psi ** -0.828 + 2 * np.sqrt(1.07 - psi)
        if rey < 1000:
            return c100 + (1.0 - c100) * (rey - 100) / 900
        return 1.0
    if model == 6:
        sph = 24 / rey * (1 + 0.15 * rey ** 0.687) + 0.42 / (1 + 42500 / rey ** 1.16)
        return 0.74533 * rey ** 0.146012 * sph / psi ** (0.5134 / max(rey, 0.01) ** 0.2)
    return 24 / rey * (((1 - psi) / rey + 1) ** 0.25 + 0.1806 * rey ** 0.6459 * psi ** -(rey ** 0.08)) + \
        0.4251 / (1 + 6880.95 / rey * psi ** 5.03)


def test_dellino_closed_form():
    d, rhop, rhoa, mu, psi = 3e-4, 2400.0, 1.0, 1.8e-5, 0.7
    v, _ = M.vset_point(d, rhop, psi, np.array([rhoa]), np.array([mu]), 4)
    want = 1.2065 * mu * (d ** 3 * 9.81 * rhop * rhoa * psi ** 1.6 / mu ** 2) ** 0.5206 / (d * rhoa)
    np.testing.assert_allclose(v, want, rtol=1e-13)


def test_sutherland_viscosity_at_reference_temperature():
    rhoc = np.full((2, 2, 2), 1.2)
    tc = np.full((2, 2, 2), M.TA0)
    opts = {"cat": [1], "aggregate": [0], "diam": [1e-6], "rho": [1000.0], "psi": [1.0], "modv": 1,
            "vset_fac": 1.0}
    vs, ok = M.vs_field(opts, rhoc, tc, 1)
    np.testing.assert_allclose(vs, _stokes(1e-6, 1000.0, M.VISA0), rtol=2e-3)


def test_cartesian_scaling_is_identity_and_sigma_correction():
    nx, ny, nz, nb = 6, 5, 4, 3
    g, corner, mg, opts = meteo_corner_problem(nx, ny, nz, nb, seed=3, map_v=0, modkh=1, modkv=1)
    out = M.update_meteo(nx, ny, nz, nb, g["dX1_p"], g["dX2_p"], corner, mg, opts)
    uc = corner["uc"][0].copy()
    uc[:, :, 0] = 0.0
    np.testing.assert_array_equal(out["u1"], M.stagger_u(uc))
    assert np.all(out["k1"] == 250.0) and np.all(out["k3"] == 15.0)
    # sigma without decay, uniform winds U, V, W at every corner above ground
    g, corner, mg, opts = meteo_corner_problem(nx, ny, nz, nb, seed=3, map_v=1, modkh=1, modkv=1)
    U, V, W = 7.0, -3.0, 0.25
    for t in (0, 1):
        corner["uc"][t][...] = U
        corner["vc"][t][...] = V
        corner["wc"][t][...] = W
    out = M.update_meteo(nx, ny, nz, nb, g["dX1_p"], g["dX2_p"], corner, mg, opts)
    hx, hy, h3 = mg["dhdx_p"], mg["dhdy_p"], mg["Hm3_p"]
    w = out["w1"]                                   # (nz+3, nx, ny); faces 2..nz+1 see no ground corner
    for f in range(2, nz + 2):
        np.testing.assert_allclose(w[f], (W - hx * U - hy * V) / h3, rtol=1e-13)
    # k3 on the sigma map: (kv + (hx^2 + hy^2) kh) / Hm3^2 with constant kh, kv
    np.testing.assert_allclose(out["k3"][2], (15.0 + hx * hx * 250.0 + hy * hy * 250.0) / (h3 * h3), rtol=1e-13)


# ---------------------------------------------------------------------------
# dry deposition (Feng 2008), oracle/meteo_oracle.py:drydep_velocity
def test_drydep_modes_and_land_classes():
    assert [M.drydep_mode(d) for d in (2e-4, 100.0001e-6, 100e-6, 10.0001e-6, 10e-6, 2.5e-6, 1e-7, 5e-8)] == \
        [0, 0, 1, 1, 2, 3, 4, 4]  # strict ">" at every threshold
    want = {1: 1, 10: 3, 19: 3, 16: 5, 20: 2, 21: 2, 22: 2, 23: 2, 24: 4}
    for c in range(0, 26):
        assert M.land_class(c) == want.get(c, 6), c  # everything else (and unknown codes) rural
    assert M.land_class(16.9) == 5  # INT() truncates


def _uniform_column(nx, ny, ust, T, rho, dz, land):
    f2 = lambda v: np.full((nx + 1, ny + 1), float(v))
    tc = np.full((nx + 1, ny + 1, 2), float(T))
    rhoc = np.full((nx + 1, ny + 1, 2), float(rho))
    zc = np.stack([f2(0.0), f2(dz)], axis=2)
    return f2(ust), tc, rhoc, zc, f2(land)


def test_drydep_velocity_structure():
    # T = Ta0 makes the viscosity visa0 exactly; dz puts Re* at the Gaussian's centre 40300
    ust, rho = 0.3, 1.1
    dz = 2.0 * 40300.0 * M.VISA0 / (ust * rho)
    diam = [5e-5, 5e-6, 1e-6, 5e-8, 2e-4]  # modes 1, 2, 3, 4 and none
    for land_code, land in ((10, 3), (16, 5), (1, 1)):
        u, tc, rhoc, zc, lu = _uniform_column(3, 2, ust, M.TA0, rho, dz, land_code)
        vd = M.drydep_velocity(diam, u, tc, rhoc, zc, lu)
        ra = 1.0 / (0.4 * ust)
        for b, mode in enumerate((1, 2, 3, 4)):
            # surface term from the table row of the land class and column of the mode
            vs = 0.0226 * ust + M.DRYDEP_A[land - 1, mode - 1] * ust ** M.DRYDEP_B[land - 1, mode - 1]
            v = vd[:, :, b]
            assert np.all(v < 1.0 / ra) and np.all(v < vs)               # resistances in series
            np.testing.assert_allclose(1.0 / v, ra + 1.0 / vs, rtol=1e-12)
        assert np.all(vd[:, :, 4] == 0.0)                                # > 100 um: none
    # table orientation: desert / coarse is the steepest entry of its row
    assert (M.DRYDEP_A[2, 1], M.DRYDEP_B[2, 1]) == (0.2928, 3.8581)
    # the Re* term is a Gaussian centred at 40300: equal at 40300 -+ delta, largest at the centre
    vals = []
    for f in (0.7, 1.0, 1.3):
        u, tc, rhoc, zc, lu = _uniform_column(2, 2, ust, M.TA0, rho, dz * f, 1)
        vals.append(M.drydep_velocity([5e-5], u, tc, rhoc, zc, lu)[0, 0, 0])
    assert vals[1] > vals[0] and vals[1] > vals[2] and abs(vals[0] - vals[2]) <= 1e-12 * vals[1]
    # no friction velocity, no deposition velocity
    u, tc, rhoc, zc, lu = _uniform_column(2, 2, 0.0, 280.0, rho, 50.0, 7)
    assert np.all(M.drydep_velocity([1e-6], u, tc, rhoc, zc, lu) == 0.0)


def test_drydep_enters_the_two_lowest_faces_only():
    nx, ny, nz = 6, 5, 4
    diam = np.array([5e-8, 1e-6, 5e-6, 5e-5, 2e-4])
    g, corner, mg, opts = meteo_corner_problem(nx, ny, nz, 5, seed=4, map_v=1, modkv=1, modv=2)
    opts = dict(opts, diam=diam)
    corner = dict(corner, lusec=np.asfortranarray(np.random.default_rng(1).integers(0, 26, (nx + 1, ny + 1)) * 1.0))
    base = M.update_meteo(nx, ny, nz, 5, g["dX1_p"], g["dX2_p"], corner, mg, opts)["vs"]
    dd = M.update_meteo(nx, ny, nz, 5, g["dX1_p"], g["dX2_p"], corner, mg, dict(opts, dry_deposition=1))["vs"]
    assert np.array_equal(dd[:, :, 2:, :], base[:, :, 2:, :])   # faces above the ground untouched
    assert np.array_equal(dd[..., 4], base[..., 4])              # the 200 um bin has no deposition velocity
    add = dd[:, :, :2, :4] - base[:, :, :2, :4]
    assert np.all(add > 0.0) and np.allclose(add[:, :, 0], add[:, :, 1], rtol=1e-12)
    assert np.all(dd[:, :, :2, 0] > 0.0) and np.all(base[:, :, :, 0] == 0.0)  # the aerosol bin moves too
