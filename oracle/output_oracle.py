"""CPU restatement (numpy, FP64) of the output post-processing of
nc_IO_out_res (Sources/mod_nc_IO.F90:3206-4335). TEST INFRASTRUCTURE ONLY:
tests and the benchmark's CPU baseline may call it; the product path never
does.

c(-1:nx+2,-1:ny+2,-1:nz+2,nb) and acum(-1:nx+2,-1:ny+2,nb) in the reference
layouts; outputs at the cell corners (nx+1, ny+1[, nz+1]).

Pins (tests/test_output_oracle.py): constant and linear fields average to
their exact corner values, negative values clip to 0, map factors divide,
the trapezoid column load of a profile linear in height is exact, the cloud
top is the highest level over the threshold, PM selection by diameter, the
Dobson conversion factor, and the ground-load corner averages of a field
with clamped edges; for the cuts, grid_get_shapez's intervals, weights and
out-of-mesh ends, the exact z-level values of a field linear in height over
terrain-following columns, and x, y, z and flight-level cuts of fields
linear in x, y and z.
"""
from __future__ import annotations

import numpy as np


def unscale_clip(c, hm1=None, hm2=None, hm3=None):
    """max(0, 1e3 c / (Hm1 Hm2 Hm3)) on mass points -1..n+2 with the map
    factors at the clamped interior index (mod_nc_IO.F90:3397-3416)."""
    nx, ny = c.shape[0] - 4, c.shape[1] - 4
    jj = np.clip(np.arange(-1, ny + 3), 1, ny)
    ii = np.clip(np.arange(-1, nx + 3), 1, nx)
    h1 = np.ones(ny) if hm1 is None else np.asarray(hm1, dtype=np.float64)
    h2 = np.ones(ny + 4) if hm2 is None else np.asarray(hm2, dtype=np.float64)
    h3 = np.ones((nx, ny)) if hm3 is None else np.asarray(hm3, dtype=np.float64).reshape(nx, ny, order="F")
    H1 = h1[jj - 1][None, :]
    H2 = h2[jj + 1][None, :]
    H3 = h3[np.ix_(ii - 1, jj - 1)]
    den = H1 * H2 * H3
    v = 1e3 * c / den[:, :, None, None]
    return np.maximum(0.0, v)


def p2c(cs):
    """8-point average of mass points (i-1..i, j-1..j, k-1..k) at corner
    (i, j, k), i in 1..nx+1 (mod_Grid.F90:1150-1181); cs with 2 ghosts."""
    a = cs[1:-1, 1:-1, 1:-1]  # mass points 0..n+1
    return (a[1:, 1:, 1:] + a[:-1, 1:, 1:] + a[1:, :-1, 1:] + a[:-1, :-1, 1:] +
            a[1:, 1:, :-1] + a[:-1, 1:, :-1] + a[1:, :-1, :-1] + a[:-1, :-1, :-1]) / 8.0


def p2c_2d(acum):
    """4-point corner average after copying the first/last interior rows
    into the ghost rows (mod_Grid.F90:1190-1220); acum (nx+4, ny+4, ...)."""
    nx, ny = acum.shape[0] - 4, acum.shape[1] - 4
    a = acum[1:-1, 1:-1].copy()  # mass points 0..n+1
    a[0] = a[1]
    a[-1] = a[-2]
    a[:, 0] = a[:, 1]
    a[:, -1] = a[:, -2]
    return (a[1:, 1:] + a[:-1, 1:] + a[1:, :-1] + a[:-1, :-1]) / 4.0


def postprocess(c, acum, z_c=None, hm1=None, hm2=None, hm3=None, bin_species=None, species=0, species_category=1,
                so2_du=False, pm_um=(), bin_diam=None):
    nb = c.shape[3]
    spe = np.zeros(nb, dtype=int) if bin_species is None else np.asarray(bin_species)
    species = 0 if bin_species is None else species
    cs = unscale_clip(np.asarray(c, dtype=np.float64), hm1, hm2, hm3)
    cc = np.stack([p2c(cs[..., b]) for b in range(nb)], axis=-1)
    ac = np.stack([p2c_2d(np.asarray(acum, dtype=np.float64)[..., b]) for b in range(nb)], axis=-1)
    sel = [b for b in range(nb) if spe[b] == species]
    w = np.zeros(cc.shape[:3])
    for b in sel:
        w = w + cc[..., b]
    out = {"con_bins": cc, "con": w, "con_sur": w[:, :, 0].copy(), "grn_load_bin": ac}
    g = np.zeros(ac.shape[:2])
    for b in sel:
        g = g + ac[..., b]
    out["grn_load"] = g
    if z_c is not None:
        z = np.asarray(z_c, dtype=np.float64)
        nz1 = w.shape[2]
        load = np.zeros(w.shape[:2])
        for k in range(nz1 - 1):
            load = load + 0.5 * (w[:, :, k] + w[:, :, k + 1]) * (z[:, :, k + 1] - z[:, :, k])
        if so2_du:
            load = load / 64.0 * 2.238e3
        out["col_mass"] = load
        top = np.zeros(w.shape[:2])
        found = np.zeros(w.shape[:2], dtype=bool)
        for k in range(nz1 - 1, -1, -1):
            hit = (~found) & (w[:, :, k] > 50e-6)
            top = np.where(hit, z[:, :, k], top)
            found |= hit
        out["cloud_top"] = top
        pm = np.zeros(w.shape[:2] + (max(len(pm_um), 1),))
        if len(pm_um) and species_category != 2:
            for p, th in enumerate(pm_um):
                wp = np.zeros(w.shape)
                for b in sel:
                    if bin_diam[b] <= th * 1e-6:
                        wp = wp + cc[..., b]
                acc = np.zeros(w.shape[:2])
                for k in range(nz1 - 1):
                    acc = acc + 0.5 * (wp[:, :, k] + wp[:, :, k + 1]) * (z[:, :, k + 1] - z[:, :, k])
                pm[..., p] = acc
        out["col_mass_pm"] = pm
    return out


def shapez(v, z):
    """grid_get_shapez (mod_Grid.F90:1287-1329) over v(1..n): (ik, shape)
    of the first interval [v(k), v(k+1)] holding z, searched upward; below
    v(1) it is (1, 0), above v(n) (n-1, 1)."""
    n = len(v)
    if z < v[0]:
        return 1, 0.0
    if z > v[n - 1]:
        return n - 1, 1.0
    for k in range(1, n):
        if v[k - 1] <= z <= v[k]:
            return k, (z - v[k - 1]) / (v[k] - v[k - 1])
    raise ValueError("levels not ordered upward")


def cuts(con, z_c, sigma, X3max, lonmin=0.0, dlon=1.0, latmin=0.0, dlat=1.0, x_cut=(), y_cut=(), z_cut=(),
         fl_cut=()):
    """Concentration cuts of nc_IO_out_res (mod_nc_IO.F90:3292-3393 and
    4084-4332) from the species' corner concentration con (nx+1, ny+1, nz+1):
    con on the z levels sigma(k) X3max per corner column (interpolated over
    the column's heights z_c), then interpolated to the cut positions."""
    nbx, nby, nbz = con.shape
    zl = [float(sigma[k]) * X3max for k in range(nbz)]
    cz = np.zeros_like(con)
    for j in range(nby):
        for i in range(nbx):
            col = [float(z) for z in z_c[i, j, :]]
            for k in range(nbz):
                kk, s = shapez(col, zl[k])
                cz[i, j, k] = (1.0 - s) * con[i, j, kk - 1] + s * con[i, j, kk] if kk > 0 else 0.0
    xc = [lonmin + float(i) * dlon for i in range(nbx)]
    yc = [latmin + float(j) * dlat for j in range(nby)]
    out = {"cut_x": np.zeros((nby, nbz, len(x_cut))), "cut_y": np.zeros((nbx, nbz, len(y_cut))),
           "cut_z": np.zeros((nbx, nby, len(z_cut))), "cut_fl": np.zeros((nbx, nby, len(fl_cut)))}
    for q, x in enumerate(x_cut):
        ii, s = shapez(xc, x)
        out["cut_x"][:, :, q] = (1.0 - s) * cz[ii - 1, :, :] + s * cz[ii, :, :]
    for q, y in enumerate(y_cut):
        jj, s = shapez(yc, y)
        out["cut_y"][:, :, q] = (1.0 - s) * cz[:, jj - 1, :] + s * cz[:, jj, :]
    for key, zs in (("cut_z", list(z_cut)), ("cut_fl", [f * 30.48 for f in fl_cut])):
        for q, z in enumerate(zs):
            kk, s = shapez(zl, z)
            out[key][:, :, q] = (1.0 - s) * cz[:, :, kk - 1] + s * cz[:, :, kk]
    return out
