"""CPU restatement (numpy, FP64) of the per-meteo-interval preprocessing of
F3D_update_meteo_term (Sources/mod_F3D.F90:564-969). TEST INFRASTRUCTURE
ONLY: tests and the benchmark's CPU baseline may call it; the product path
never does.

Inputs follow the reference: corner fields (nx+1, ny+1, nz+1) and 2-D corner
fields (nx+1, ny+1) at the two meteo times bounding the interval, the map
factors Hm1_p(1:ny), Hm2_p(-1:ny+2), Hm3_p(1:nx,1:ny), slopes, heights and
the physics options. Outputs are the fields in the layouts the ADS step
reads (those of fall3d_ads_set_meteo): u (nx+3,ny,nz), v (ny+3,nx,nz),
w (nz+3,nx,ny) for both times, k1 (nx+4,ny,nz), k2 (ny+4,nx,nz),
k3 (nz+4,nx,ny), vs (nx,ny,nz+3,nb).

Each statement keeps the reference's operand order (numpy evaluates left to
right, one rounding per operation, no fused multiply-add).

Pins (tests/test_meteo_oracle.py): linear corner fields stagger and average
to their exact face/centre values; CMAQ and RAMS on uniform-shear winds
against their closed forms; the similarity branches at chosen heights
against hand-evaluated profiles; the drag laws against published sphere
limits and Stokes' law, the iteration's fixed point against an independent
root finder, Dellino's closed form; the metric scalings on Cartesian maps
reduce to the identity and on sigma maps follow the hand-evaluated terrain
correction; the dry deposition velocity: mode thresholds, land classes and
table orientation, resistances in series, the Re* Gaussian centred at 40300,
zero u*, and only the two lowest faces of the bins up to 100 um changed.
"""
from __future__ import annotations

import numpy as np

GI = 9.81            # Sources/mod_KindType.F90:45
KARMAN = 0.4         # :46
VISA0 = 1.827e-5     # Sources/mod_Phys.F90:67
TA0 = 291.15         # :68
CAT_AEROSOL = 2      # Sources/mod_KindType.F90:55


def _F(a):
    return np.asfortranarray(a)


# ---------------------------------------------------------------------------
# corner -> face winds, Sources/mod_Grid.F90:685-752
def stagger_u(uc):
    nx1, ny1, nz1 = uc.shape
    f = 0.25 * (uc[:, :-1, :-1] + uc[:, 1:, :-1] + uc[:, :-1, 1:] + uc[:, 1:, 1:])  # (nx+1, ny, nz)
    out = np.empty((nx1 + 2, ny1 - 1, nz1 - 1))
    out[1:-1] = f
    out[0] = f[0]
    out[-1] = f[-1]
    return out                                   # faces 0..nx+2


def stagger_v(vc):
    nx1, ny1, nz1 = vc.shape
    f = 0.25 * (vc[:-1, :, :-1] + vc[1:, :, :-1] + vc[:-1, :, 1:] + vc[1:, :, 1:])  # (nx, ny+1, nz)
    out = np.empty((nx1 - 1, ny1 + 2, nz1 - 1))
    out[:, 1:-1] = f
    out[:, 0] = f[:, 0]
    out[:, -1] = f[:, -1]
    return out                                   # (nx, faces 0..ny+2, nz)


def stagger_w(wc):
    f = 0.25 * (wc[:-1, :-1, :] + wc[1:, :-1, :] + wc[:-1, 1:, :] + wc[1:, 1:, :])  # (nx, ny, nz+1)
    return f


def _ghost_faces_z(f):
    out = np.empty(f.shape[:2] + (f.shape[2] + 2,))
    out[:, :, 1:-1] = f
    out[:, :, 0] = f[:, :, 0]
    out[:, :, -1] = f[:, :, -1]
    return out


# corner -> mass points, Sources/mod_Grid.F90:1229-1277
def c2p(cc):
    return (cc[:-1, :-1, :-1] + cc[1:, :-1, :-1] + cc[:-1, 1:, :-1] + cc[1:, 1:, :-1] +
            cc[:-1, :-1, 1:] + cc[1:, :-1, 1:] + cc[:-1, 1:, 1:] + cc[1:, 1:, 1:]) / 8.0


def c2p_2d(cc):
    return (cc[:-1, :-1] + cc[1:, :-1] + cc[:-1, 1:] + cc[1:, 1:]) / 4.0


# ---------------------------------------------------------------------------
# horizontal diffusion, Sources/mod_Phys.F90:445-575 (mass points, unscaled)
def kh_points(opts, uc, vc, dX1_p, dX2_p, hm1, hm2):
    """uc, vc: corner winds interpolated in time; dX1_p(-1:nx+2) at index
    p+1; hm2 = Hm2_p(-1:ny+2) at index p+1. Returns (nx, ny, nz)."""
    nx, ny, nz = uc.shape[0] - 1, uc.shape[1] - 1, uc.shape[2] - 1
    if opts["modkh"] == 1:
        return np.full((nx, ny, nz), float(opts["kh0"]))
    i = np.arange(1, nx + 1)[:, None, None]
    j = np.arange(1, ny + 1)[None, :, None]
    H1 = hm1[j - 1]
    H2 = hm2[j + 1]
    dx = dX1_p[i + 1] * H1
    dy = dX2_p[j + 1] * H2
    U = uc[:, :, :-1]
    V = vc[:, :, :-1]
    u00, u10, u01, u11 = U[:-1, :-1], U[1:, :-1], U[:-1, 1:], U[1:, 1:]
    v00, v10, v01, v11 = V[:-1, :-1], V[1:, :-1], V[:-1, 1:], V[1:, 1:]
    dudx = 0.5 * (u10 - u00 + u11 - u01) / dx
    dvdx = 0.5 * (v10 - v00 + v11 - v01) / dx
    dudy = 0.5 * (u01 - u00 + u11 - u10) / dy
    dvdy = 0.5 * (v01 - v00 + v11 - v10) / dy
    alfa, khf, rkhmin = 0.28, 8000.0, 100.0
    if opts["modkh"] == 2:
        khn0 = 4000.0 if opts["map_h"] == 0 else 0.036
        d1 = dudx - dvdy
        d2 = dvdx + dudy
        kht = alfa * alfa * dx * dy * np.sqrt(d1 * d1 + d2 * d2)
        khn = khf * (khn0 / (H1 * opts["dlon"])) * (khn0 / (H2 * opts["dlat"]))
        khn = np.broadcast_to(khn, kht.shape)
        with np.errstate(divide="ignore", invalid="ignore"):
            k = np.maximum(kht * khn / (kht + khn), rkhmin)
        return np.where((kht < 1.0) & (khn < 1.0), rkhmin, k)
    if opts["modkh"] == 4:
        kmh = 0.075 * np.power(dx * dy, 2.0 / 3.0)
        d1 = dudy + dvdx
        cs = opts["rams_cs"]
        k = cs * cs * dx * dy * np.sqrt(d1 * d1 + 2.0 * (dudx * dudx + dvdy * dvdy))
        return np.maximum(np.maximum(k, kmh), rkhmin)
    raise ValueError("horizontal diffusion model")


# similarity helpers, Sources/mod_Phys.F90:1684-1741
def fih(z, L):
    zl = z / L
    return np.where(zl < 0.0, 1.0 / np.sqrt(np.maximum(1.0 - 15.0 * zl, 0.0)), 1.0 + 5.0 * zl)


def rib(tv1, tv0, umod, z1, z0):
    um2 = np.maximum(umod * umod, 1e-18)
    return GI * (z1 - z0) * (tv1 - tv0) / (um2 * 0.5 * (tv1 + tv0))


def fc(ri):
    with np.errstate(invalid="ignore"):
        st = 1.0 / (1.0 + 10.0 * ri * (1.0 + 8.0 * ri))
        un = np.sqrt(1.0 - 18.0 * ri)
    return np.where(ri > 0.0, st, un)


def kv_points(opts, uc, vc, tvc, pblhc, ustc, monc, zc, hc):
    """Vertical diffusion at mass points (unscaled), Sources/mod_Phys.F90:577-653.
    Inputs are already interpolated in time (corner fields)."""
    nx, ny, nz = uc.shape[0] - 1, uc.shape[1] - 1, uc.shape[2] - 1
    if opts["modkv"] == 1:
        return np.full((nx, ny, nz), float(opts["kv0"]))
    if opts["modkv"] != 3:
        raise ValueError("vertical diffusion model")
    pblh = c2p_2d(pblhc)[:, :, None]
    ust = c2p_2d(ustc)[:, :, None]
    mon = c2p_2d(monc)
    mon = np.where(np.abs(mon) < 10, np.copysign(10.0, mon), mon)[:, :, None]
    zp = c2p(zc)
    hp = c2p_2d(hc)[:, :, None]
    z = np.maximum(0.0, zp - hp)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = fih(z, mon)
        k_in = KARMAN * z * ust * (1.0 - z / pblh) * (1.0 - z / pblh) / f
        landac = 30.0
        lc = 1.0 / (KARMAN * z) + 1.0 / landac
        lc = 1.0 / lc

        def face(x, k0):  # 4-corner average at corner level k0 (0-based slices)
            return 0.25 * (x[:-1, :-1, k0] + x[1:, :-1, k0] + x[:-1, 1:, k0] + x[1:, 1:, k0])

        z1 = face(zc, slice(0, nz)) - hp
        z2 = face(zc, slice(1, nz + 1)) - hp
        u1, v1 = face(uc, slice(0, nz)), face(vc, slice(0, nz))
        u2, v2 = face(uc, slice(1, nz + 1)), face(vc, slice(1, nz + 1))
        umod1 = np.sqrt(u1 * u1 + v1 * v1)
        umod2 = np.sqrt(u2 * u2 + v2 * v2)
        umod = 0.5 * (umod1 + umod2)
        tv1, tv2 = face(tvc, slice(0, nz)), face(tvc, slice(1, nz + 1))
        ri = rib(tv2, tv1, umod, z2, z1)
        k_out = lc * lc * np.abs((umod2 - umod1) / (z2 - z1)) * fc(ri)
    k = np.where(z < pblh, k_in, k_out)
    return np.maximum(0.0, k)


# ---------------------------------------------------------------------------
# terminal velocity, Sources/mod_Phys.F90:752-960 (vectorized over points;
# each point iterates until its own convergence, as the scalar routine)
def vset_point(diam, rhop, psi, rhoa, visc, model):
    rhoa = np.asarray(rhoa, dtype=np.float64)
    visc = np.asarray(visc, dtype=np.float64)
    shape = np.broadcast(rhoa, visc).shape
    rhoa = np.broadcast_to(rhoa, shape)
    visc = np.broadcast_to(visc, shape)
    atol, rtol, maxit = 1e-12, 1e-6, 1000
    if model == 4:  # Dellino
        v = np.power((diam ** 3 * GI * rhop * rhoa * (psi ** 1.6)) / (visc ** 2), 0.5206)
        v = (1.2065 * visc * v) / (diam * rhoa)
        return v, np.ones(shape, dtype=bool)
    cd = 1.0
    v = np.sqrt(4.0 * GI * diam * rhop / (3.0 * cd * rhoa))
    vold = v.copy()
    done = np.zeros(shape, dtype=bool)
    out = v.copy()
    cd100 = 0.24 * psi ** (-0.828) + 2.0 * np.sqrt(1.07 - psi) if psi <= 1.07 else np.nan
    cd1000 = 1.0
    pa = (cd1000 - cd100) / 900.0
    pb = cd1000 - 1000.0 * pa
    with np.errstate(all="ignore"):
        for _ in range(maxit):
            rey = rhoa * v * diam / visc
            if model == 1:
                cd = np.where(rey <= 988.947, 24.0 / rey * (1.0 + 0.15 * rey ** 0.687), 0.44)
            elif model == 2:
                rk1 = 3.0 / (1.0 + 2.0 / np.sqrt(psi))
                rk2 = 10.0 ** (1.8148 * (-np.log10(psi)) ** 0.5743)
                cd = 24.0 / (rey * rk1) * (1.0 + 0.1118 * (rey * rk1 * rk2) ** 0.6567) + \
                    0.4305 * rk2 / (1.0 + 3305.0 / (rey * rk1 * rk2))
            elif model == 3:
                cd = 24.0 / rey * psi ** (-0.828) + 2.0 * np.sqrt(1.07 - psi)
            elif model == 5:
                cd = np.where(rey <= 100.0, 24.0 / rey * psi ** (-0.828) + 2.0 * np.sqrt(1.07 - psi),
                              np.where((rey > 100.0) & (rey < 1000.0), pa * rey + pb, cd1000))
            elif model == 6:
                cd = 24.0 * (1.0 + 0.15 * rey ** 0.687) / rey + 0.42 / (1.0 + 42500.0 / rey ** 1.16)
                cd = 0.74533 * rey ** 0.146012 * cd / psi ** (0.5134 / np.maximum(rey, 0.01) ** 0.2)
            elif model == 7:
                cd = 24.0 / rey * (((1.0 - psi) / rey + 1.0) ** 0.25 +
                                   (0.1806 * rey ** 0.6459) * psi ** (-(rey ** 0.08))) + \
                    0.4251 / (1.0 + 6880.95 / rey * psi ** 5.03)
            else:
                raise ValueError("settling velocity model")
            v = np.sqrt(4.0 * GI * diam * rhop / (3.0 * cd * rhoa))
            conv = (np.abs(v - vold) <= atol + rtol * np.abs(v)) & ~done
            out = np.where(conv, v, out)
            done |= conv
            if done.all():
                break
            vold = np.where(done, vold, v)
            v = np.where(done, out, v)
    return out, done


def vs_field(opts, rhoc, tc, nb):
    """(nx, ny, nz+1, nb) settling velocities at the w-boundaries (unscaled, no
    ghosts) from time-interpolated corner density and temperature."""
    rhoa = 0.25 * (rhoc[:-1, :-1, :] + rhoc[1:, :-1, :] + rhoc[1:, 1:, :] + rhoc[:-1, 1:, :])
    ta = 0.25 * (tc[:-1, :-1, :] + tc[1:, :-1, :] + tc[1:, 1:, :] + tc[:-1, 1:, :])
    visa = VISA0 * ((TA0 + 120.0) / (ta + 120.0)) * np.power(ta / TA0, 1.5)
    out = np.zeros(rhoa.shape + (nb,))
    ok = True
    for b in range(nb):
        if int(opts["cat"][b]) != CAT_AEROSOL:
            v, done = vset_point(float(opts["diam"][b]), float(opts["rho"][b]), float(opts["psi"][b]), rhoa, visa,
                                 int(opts["modv"]))
            out[..., b] = v
            ok = ok and bool(done.all())
        if int(opts["aggregate"][b]):
            out[..., b] = opts["vset_fac"] * out[..., b]
    return out, ok


# ---------------------------------------------------------------------------
# Dry deposition velocity (Feng 2008), phys_dry_deposition, mod_Phys.F90:1479-1673
# coefficients a, b of each land class (rows 1..6: urban, remote continental,
# desert, polar, marine, rural) and aerosol mode (columns 1..4: giant, coarse,
# accumulation, nuclei), mod_Phys.F90:1521-1571
DRYDEP_A = np.array([[0.0048, 0.0315, 1.2891, 1.0338], [0.0037, 0.0120, 1.3977, 1.0707],
                     [0.0042, 0.2928, 1.3970, 0.9155], [0.0032, 0.1201, 1.1838, 1.0096],
                     [0.0043, 0.1337, 1.2834, 1.1595], [0.0045, 0.0925, 1.2654, 1.0891]])
DRYDEP_B = np.array([[1.0, 2.7925, 2.6878, 1.2644], [1.0, 2.2413, 2.5838, 1.3247],
                     [1.0, 3.8581, 2.5580, 1.0364], [1.0, 3.4407, 2.8033, 1.2069],
                     [1.0, 3.5456, 2.7157, 1.4863], [1.0, 3.2920, 2.7227, 1.4240]])


def drydep_mode(diam):
    """Aerosol mode of a bin by diameter: 0 above 100 um (no deposition
    velocity), 1 giant (> 10 um), 2 coarse (> 2.5 um), 3 accumulation
    (> 0.1 um), 4 nuclei (mod_Phys.F90:1575-1587)."""
    if diam > 100e-6:
        return 0
    if diam > 10e-6:
        return 1
    if diam > 2.5e-6:
        return 2
    if diam > 0.1e-6:
        return 3
    return 4


def land_class(category):
    """Feng land class of a USGS 24-category land use (mod_Phys.F90:1592-1608)."""
    c = int(category)
    if c == 1:
        return 1
    if c in (10, 19):
        return 3
    if c == 16:
        return 5
    if c in (20, 21, 22, 23):
        return 2
    if c == 24:
        return 4
    return 6


def drydep_velocity(diam, ustc, tc, rhoc, zc, lusec):
    """(nx, ny, nb) deposition velocity of each mass-point column and bin
    (0 for bins above 100 um) from the time-interpolated corner u*, the
    lowest corner temperature and density, the two lowest corner heights
    and the land use at corner (i, j) (mod_Phys.F90:1613-1670)."""
    avg = lambda f: 0.25 * (f[:-1, :-1] + f[1:, :-1] + f[1:, 1:] + f[:-1, 1:])
    ust = avg(ustc)
    with np.errstate(divide="ignore"):
        ra = np.where(ust > 0.0, 1.0 / (KARMAN * np.where(ust > 0.0, ust, 1.0)), 1e20)
    T = avg(tc[:, :, 0])
    visc = VISA0 * ((TA0 + 120.0) / (T + 120.0)) * np.power(T / TA0, 1.5)
    rho = avg(rhoc[:, :, 0])
    z1, z2 = avg(zc[:, :, 0]), avg(zc[:, :, 1])
    Re = ust * (z2 - z1) * 0.5 * rho / visc
    vd1 = -0.5 * ((Re - 40300.0) / 15330.0) * ((Re - 40300.0) / 15330.0)
    vd1 = 0.0226 * ust * np.exp(vd1)
    land = np.vectorize(land_class)(np.asarray(lusec)[:-1, :-1]) - 1
    out = np.zeros(ust.shape + (len(diam),))
    for b, d in enumerate(diam):
        m = drydep_mode(float(d))
        if m == 0:
            continue
        a, bb = DRYDEP_A[land, m - 1], DRYDEP_B[land, m - 1]
        vd = vd1 + a * np.power(ust, bb)
        with np.errstate(divide="ignore"):
            out[:, :, b] = 1.0 / (ra + 1.0 / vd)
    return out


# ---------------------------------------------------------------------------
def update_meteo(nx, ny, nz, nb, dX1_p, dX2_p, corner, grid, opts):
    """The whole update; returns a dict of the ADS input fields (see the
    module docstring) and "converged"."""
    s = float(opts["s_time"])
    uc = [np.array(corner["uc"][t], dtype=np.float64, order="F") for t in (0, 1)]
    vc = [np.array(corner["vc"][t], dtype=np.float64, order="F") for t in (0, 1)]
    wc = [np.array(corner["wc"][t], dtype=np.float64, order="F") for t in (0, 1)]
    for t in (0, 1):  # ground corners, Sources/mod_F3D.F90:662-669
        uc[t][:, :, 0] = 0.0
        vc[t][:, :, 0] = 0.0
        wc[t][:, :, 0] = 0.0
    hm1 = np.asarray(grid["Hm1_p"], dtype=np.float64)
    hm2 = np.asarray(grid["Hm2_p"], dtype=np.float64)
    hm3 = np.asarray(grid["Hm3_p"], dtype=np.float64).reshape(nx, ny, order="F")
    map_v = int(opts["map_v"])
    if map_v:
        dhdx = np.asarray(grid["dhdx_p"], dtype=np.float64).reshape(nx, ny, order="F")
        dhdy = np.asarray(grid["dhdy_p"], dtype=np.float64).reshape(nx, ny, order="F")
    out = {}
    for t in (0, 1):
        u = stagger_u(uc[t]) / hm1[None, :, None]
        v = stagger_v(vc[t])
        hm2v = 0.5 * (hm2[1:] + hm2[:-1])             # Hm2 at faces 0..ny+2 (Hm2_p(p) at index p+1)
        v = v / hm2v[None, :, None]
        w = stagger_w(wc[t])                            # (nx, ny, nz+1)
        if map_v:
            if map_v == 2:
                hz = (1.0 - np.asarray(grid["X3_c"], dtype=np.float64) / opts["x3max"])[None, None, :]
                hx = dhdx[:, :, None] * hz
                hy = dhdy[:, :, None] * hz
            else:
                hx = dhdx[:, :, None]
                hy = dhdy[:, :, None]
            uw = 0.25 * (uc[t][:-1, :-1, :] + uc[t][1:, :-1, :] + uc[t][:-1, 1:, :] + uc[t][1:, 1:, :])
            vw = 0.25 * (vc[t][:-1, :-1, :] + vc[t][1:, :-1, :] + vc[t][:-1, 1:, :] + vc[t][1:, 1:, :])
            w = (w - hx * uw - hy * vw) / hm3[:, :, None]
        w = _ghost_faces_z(w)
        out[f"u{t + 1}"] = _F(u)
        out[f"v{t + 1}"] = _F(np.transpose(v, (1, 0, 2)))
        out[f"w{t + 1}"] = _F(np.transpose(w, (2, 0, 1)))
    ucl = (1.0 - s) * uc[0] + s * uc[1]
    vcl = (1.0 - s) * vc[0] + s * vc[1]
    kh = kh_points(opts, ucl, vcl, np.asarray(dX1_p, dtype=np.float64), np.asarray(dX2_p, dtype=np.float64), hm1, hm2)
    if int(opts["modkv"]) == 3:
        lerp = lambda key: (1.0 - s) * np.asarray(corner[key][0], dtype=np.float64) + \
            s * np.asarray(corner[key][1], dtype=np.float64)
        kv = kv_points(opts, ucl, vcl, lerp("tvc"), lerp("pblhc"), lerp("ustc"), lerp("monc"),
                       np.asarray(grid["z_c"], dtype=np.float64), np.asarray(grid["h_c"], dtype=np.float64))
    else:
        kv = kv_points(opts, ucl, vcl, None, None, None, None, None, None)
    # k3 terrain correction (uses the unscaled k1 = k2), Sources/mod_Grid.F90:920-960
    if map_v:
        if map_v == 2:
            hz = (1.0 - np.asarray(grid["X3_p"], dtype=np.float64) / opts["x3max"])[None, None, :]
            hx = dhdx[:, :, None] * hz
            hy = dhdy[:, :, None] * hz
        else:
            hx = dhdx[:, :, None]
            hy = dhdy[:, :, None]
        kv = (kv + hx * hx * kh + hy * hy * kh) / (hm3[:, :, None] * hm3[:, :, None])
    k1 = np.empty((nx + 4, ny, nz))
    k1[2:-2] = kh
    k1[0] = k1[1] = kh[0]
    k1[-1] = k1[-2] = kh[-1]
    k1 = k1 / (hm1[None, :, None] * hm1[None, :, None])
    k2 = np.empty((nx, ny + 4, nz))
    k2[:, 2:-2] = kh
    k2[:, 0] = k2[:, 1] = kh[:, 0]
    k2[:, -1] = k2[:, -2] = kh[:, -1]
    k2 = k2 / (hm2[None, :, None] * hm2[None, :, None])
    k3 = np.empty((nx, ny, nz + 4))
    k3[:, :, 2:-2] = kv
    k3[:, :, 0] = k3[:, :, 1] = kv[:, :, 0]
    k3[:, :, -1] = k3[:, :, -2] = kv[:, :, -1]
    out["k1"] = _F(k1)
    out["k2"] = _F(np.transpose(k2, (1, 0, 2)))
    out["k3"] = _F(np.transpose(k3, (2, 0, 1)))
    rhocl = (1.0 - s) * np.asarray(corner["rhoc"][0], dtype=np.float64) + s * np.asarray(corner["rhoc"][1], dtype=np.float64)
    tcl = (1.0 - s) * np.asarray(corner["tc"][0], dtype=np.float64) + s * np.asarray(corner["tc"][1], dtype=np.float64)
    vs, ok = vs_field(opts, rhocl, tcl, nb)
    vsg = np.empty((nx, ny, nz + 3, nb))
    vsg[:, :, 1:-1] = vs
    vsg[:, :, 0] = vs[:, :, 0]
    vsg[:, :, -1] = vs[:, :, -1]
    if int(opts.get("dry_deposition", 0)):  # added at the ground face and the one below it, before scaling
        ustl = (1.0 - s) * np.asarray(corner["ustc"][0], dtype=np.float64) + \
            s * np.asarray(corner["ustc"][1], dtype=np.float64)
        vd = drydep_velocity(opts["diam"], ustl, tcl, rhocl, np.asarray(grid["z_c"], dtype=np.float64),
                             corner["lusec"])
        for b in range(nb):
            if drydep_mode(float(opts["diam"][b])) > 0:
                vsg[:, :, 0, b] = vsg[:, :, 0, b] + vd[:, :, b]
                vsg[:, :, 1, b] = vsg[:, :, 1, b] + vd[:, :, b]
    out["vs"] = _F(vsg / hm3[:, :, None, None])
    out["converged"] = ok
    return out
