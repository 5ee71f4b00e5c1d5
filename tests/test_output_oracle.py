"""Pins of the output post-processing restatement (oracle/output_oracle.py)
against closed forms: corner averages of constant and linear fields, the
clip at 0, division by the map factors, the trapezoid column load of a
profile linear in height (exact), the cloud top of a profile crossing the
50 ug/m3 threshold at a known level, PM selection by diameter, the Dobson
conversion, and the ground-load corner average with clamped edges."""
import numpy as np

from oracle import output_oracle as O


def _c(nx, ny, nz, nb, f):
    i = np.arange(-1, nx + 3)[:, None, None, None]
    j = np.arange(-1, ny + 3)[None, :, None, None]
    k = np.arange(-1, nz + 3)[None, None, :, None]
    b = np.arange(nb)[None, None, None, :]
    return np.asfortranarray(f(i, j, k, b) + 0.0 * (i + j + k + b))


def test_linear_field_averages_to_corner_values_and_clips():
    nx, ny, nz, nb = 5, 4, 3, 2
    c = _c(nx, ny, nz, nb, lambda i, j, k, b: 1e-3 * (2.0 + 0.5 * i + 0.25 * j + 1.0 * k + 3.0 * b))
    out = O.postprocess(c, np.zeros((nx + 4, ny + 4, nb)))
    cc = out["con_bins"]
    for i in range(1, nx + 2):
        for j in range(1, ny + 2):
            for k in range(1, nz + 2):
                for b in range(nb):  # corner (i, j, k) sits at mass point (i - 1/2, j - 1/2, k - 1/2)
                    want = 2.0 + 0.5 * (i - 0.5) + 0.25 * (j - 0.5) + (k - 0.5) + 3.0 * b
                    assert abs(cc[i - 1, j - 1, k - 1, b] - want) <= 1e-13 * want
    neg = _c(nx, ny, nz, 1, lambda i, j, k, b: -1.0 + 0 * i)
    assert np.all(O.postprocess(neg, np.zeros((nx + 4, ny + 4, 1)))["con"] == 0.0)


def test_map_factors_divide():
    nx, ny, nz = 4, 3, 2
    c = _c(nx, ny, nz, 1, lambda i, j, k, b: 2e-3 + 0 * i)
    hm1, hm2, hm3 = np.full(ny, 2.0), np.full(ny + 4, 0.5), np.full((nx, ny), 4.0)
    out = O.postprocess(c, np.zeros((nx + 4, ny + 4, 1)), hm1=hm1, hm2=hm2, hm3=hm3)
    np.testing.assert_allclose(out["con"], 1e3 * 2e-3 / 4.0, rtol=1e-15)


def test_column_load_trapezoid_and_cloud_top():
    """w(z) = a + s z on irregular corner heights: the trapezoid rule is exact,
    so the load is a (ztop - z0) + s (ztop^2 - z0^2) / 2; the cloud top is the
    highest corner height with w > 50 ug/m3."""
    nx, ny, nz = 2, 2, 6
    zlev = np.array([0.0, 150.0, 400.0, 800.0, 1500.0, 2600.0, 4000.0])  # corners k = 1..nz+1
    # mass points k = 0..nz+1 placed so that corner k is the midpoint of mass
    # points k-1 and k: the corner average of a profile linear in height is
    # then its value at the corner height
    zmass = np.empty(nz + 2)
    zmass[0] = -75.0
    for k in range(1, nz + 2):
        zmass[k] = 2.0 * zlev[k - 1] - zmass[k - 1]
    a, s = 2.4e-4, -5e-8  # positive at every mass point (no clipping); crosses 50 ug/m3 below 4000 m
    c = np.zeros((nx + 4, ny + 4, nz + 4, 1))
    for k in range(nz + 2):
        c[:, :, k + 1, 0] = 1e-3 * (a + s * zmass[k])
    z_c = np.broadcast_to(zlev, (nx + 1, ny + 1, nz + 1)).copy()
    out = O.postprocess(c, np.zeros((nx + 4, ny + 4, 1)), z_c=z_c)
    want = a * (zlev[-1] - zlev[0]) + s * (zlev[-1] ** 2 - zlev[0] ** 2) / 2
    np.testing.assert_allclose(out["col_mass"], want, rtol=1e-12)
    w = a + s * zlev
    top = zlev[np.nonzero(w > 50e-6)[0].max()]
    assert np.all(out["cloud_top"] == top)
    du = O.postprocess(c, np.zeros((nx + 4, ny + 4, 1)), z_c=z_c, so2_du=True)["col_mass"]
    np.testing.assert_allclose(du, want / 64.0 * 2238.0, rtol=1e-12)


def test_pm_loads_select_bins_by_diameter():
    nx, ny, nz, nb = 2, 2, 2, 3
    c = _c(nx, ny, nz, nb, lambda i, j, k, b: 1e-3 * (b + 1.0))
    z_c = np.broadcast_to(np.array([0.0, 100.0, 300.0]), (nx + 1, ny + 1, nz + 1)).copy()
    diam = np.array([2e-6, 8e-6, 40e-6])
    out = O.postprocess(c, np.zeros((nx + 4, ny + 4, nb)), z_c=z_c, pm_um=(2.5, 10.0), bin_diam=diam)
    np.testing.assert_allclose(out["col_mass_pm"][..., 0], 1.0 * 300.0, rtol=1e-14)        # bin 0
    np.testing.assert_allclose(out["col_mass_pm"][..., 1], (1.0 + 2.0) * 300.0, rtol=1e-14)  # bins 0, 1
    aer = O.postprocess(c, np.zeros((nx + 4, ny + 4, nb)), z_c=z_c, pm_um=(2.5,), bin_diam=diam, species_category=2)
    assert np.all(aer["col_mass_pm"] == 0.0)


def test_ground_load_corner_average_with_clamped_edges():
    nx, ny, nb = 3, 2, 2
    acum = np.full((nx + 4, ny + 4, nb), np.nan)  # ghosts are never read
    for i in range(1, nx + 1):
        for j in range(1, ny + 1):
            acum[i + 1, j + 1, :] = [10.0 * i + j, 1.0]
    out = O.postprocess(np.zeros((nx + 4, ny + 4, 5, nb)), acum)
    g = out["grn_load_bin"][..., 0]
    val = lambda i, j: 10.0 * min(max(i, 1), nx) + min(max(j, 1), ny)
    for i in range(1, nx + 2):
        for j in range(1, ny + 2):
            want = (val(i, j) + val(i - 1, j) + val(i, j - 1) + val(i - 1, j - 1)) / 4.0
            assert g[i - 1, j - 1] == want
    np.testing.assert_array_equal(out["grn_load"], g + 1.0)


def test_shapez_intervals_weights_and_ends():
    v = [0.0, 10.0, 30.0, 60.0]
    assert O.shapez(v, 5.0) == (1, 0.5)
    assert O.shapez(v, 10.0) == (1, 1.0)  # a node value: the first interval holding it, searched upward
    assert O.shapez(v, 20.0) == (2, 0.5)
    assert O.shapez(v, 45.0) == (3, 0.5)
    assert O.shapez(v, 0.0) == (1, 0.0)
    assert O.shapez(v, -1.0) == (1, 0.0)  # below the mesh: the bottom value
    assert O.shapez(v, 61.0) == (3, 1.0)  # above: the top value


def _terrain(nx, ny, nz, X3max=6000.0, seed=2):
    rng = np.random.default_rng(seed)
    h = rng.uniform(0.0, 900.0, (nx + 1, ny + 1))
    sigma = np.concatenate([[0.0], np.cumsum(rng.uniform(1.0, 2.0, nz))])
    sigma /= sigma[-1]
    zc = h[:, :, None] + sigma[None, None, :] * (X3max - h[:, :, None])  # terrain-following corner heights
    return sigma, zc, h


def test_z_levels_are_exact_for_a_field_linear_in_height():
    nx, ny, nz, X3max = 5, 4, 7, 6000.0
    sigma, zc, h = _terrain(nx, ny, nz, X3max)
    con = 3.0 + 2e-3 * zc  # linear in height a.s.l.
    out = O.cuts(con, zc, sigma, X3max, z_cut=sigma * X3max)
    cz = out["cut_z"]  # cut k at level k is that level of the z-level field
    for i in range(nx + 1):
        for j in range(ny + 1):
            for k in range(nz + 1):
                z = sigma[k] * X3max
                want = 3.0 + 2e-3 * (z if z >= h[i, j] else h[i, j])  # under the terrain: the lowest corner
                assert abs(cz[i, j, k] - want) <= 1e-13 * want, (i, j, k)


def test_x_y_z_and_flight_level_cuts_of_linear_fields():
    nx, ny, nz, X3max = 6, 5, 4, 8000.0
    sigma = np.linspace(0.0, 1.0, nz + 1)
    zc = np.broadcast_to(sigma * X3max, (nx + 1, ny + 1, nz + 1)).copy()  # flat terrain: z levels = corners
    lonmin, dlon, latmin, dlat = -10.0, 0.5, 40.0, 0.25
    i = np.arange(nx + 1)[:, None, None]
    j = np.arange(ny + 1)[None, :, None]
    x, y = lonmin + i * dlon, latmin + j * dlat
    con = 1.0 + 0.3 * x + 0.7 * y + 1e-4 * zc
    xs, ys, zs, fls = [-9.3, -7.25], [40.1, 41.0], [1234.0, 7999.0], [50.0, 200.0]
    out = O.cuts(con, zc, sigma, X3max, lonmin, dlon, latmin, dlat, xs, ys, zs, fls)
    xx, yy, zz = x[:, 0, 0], y[0, :, 0], sigma * X3max
    for q, xv in enumerate(xs):
        want = 1.0 + 0.3 * xv + 0.7 * yy[:, None] + 1e-4 * zz[None, :]
        np.testing.assert_allclose(out["cut_x"][:, :, q], want, rtol=1e-13)
    for q, yv in enumerate(ys):
        want = 1.0 + 0.3 * xx[:, None] + 0.7 * yv + 1e-4 * zz[None, :]
        np.testing.assert_allclose(out["cut_y"][:, :, q], want, rtol=1e-13)
    for key, heights in (("cut_z", zs), ("cut_fl", [f * 30.48 for f in fls])):  # FL 50 = 1524 m, FL 200 = 6096 m
        for q, zv in enumerate(heights):
            want = 1.0 + 0.3 * xx[:, None] + 0.7 * yy[None, :] + 1e-4 * zv
            np.testing.assert_allclose(out[key][:, :, q], want, rtol=1e-13)
    # a cut outside the mesh takes the edge value
    edge = O.cuts(con, zc, sigma, X3max, lonmin, dlon, latmin, dlat, x_cut=[lonmin - 5.0])["cut_x"][:, :, 0]
    np.testing.assert_array_equal(edge, con[0, :, :])
