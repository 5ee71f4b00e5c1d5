"""Synthetic ADS problems (grid, winds, diffusion, settling, sources).

Every array is float64 in the layout the reference passes to its ADS
routines (Fortran column-major, explicit halo bounds):

  c    (-1:nx+2, -1:ny+2, -1:nz+2, nb)  Sources/mod_KindType.F90:389
  s    ( 1:nx,    1:ny,    1:nz,   nb)  Sources/mod_KindType.F90:390
  vs   ( 1:nx,    1:ny,    0:nz+2, nb)  Sources/mod_KindType.F90:391
  acum (-1:nx+2, -1:ny+2, nb)           Sources/mod_KindType.F90:392
  u    ( 0:nx+2,  1:ny,    1:nz)        Sources/mod_ADS.F90:110 (x fastest)
  v    ( 0:ny+2,  1:nx,    1:nz)        Sources/mod_ADS.F90:111 (y fastest)
  w    ( 0:nz+2,  1:nx,    1:ny)        Sources/mod_ADS.F90:112 (z fastest)
  k1   (-1:nx+2,  1:ny,    1:nz)        Sources/mod_ADS.F90:249
  k2   (-1:ny+2,  1:nx,    1:nz)        Sources/mod_ADS.F90:438
  k3   (-1:nz+2,  1:nx,    1:ny)        Sources/mod_ADS.F90:634
  dX1_p(-1:nx+2) dX1_b(0:nx+2) ...      Sources/mod_KindType.F90:486-491

Face index f of a sweep axis is the face on the low side of cell f, so the
boundary faces of an n-cell axis are f = 1 and f = n + 1; f = 0 and n + 2
are halo faces that the solver never reads.

The generators build the workloads the benchmark and the tests use:
  puff_problem       -- 64x64x32, one bin, Gaussian puff, analytic shear wind
                        with zero-flux lateral/top faces and an absorbing ground
  operational_problem -- seeded Fourier-mode wind + shear, log-spaced settling
                        velocities, vertical line source(s)
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np


@dataclasses.dataclass
class Problem:
    nx: int
    ny: int
    nz: int
    nb: int
    grid: dict
    u1: np.ndarray
    u2: np.ndarray
    v1: np.ndarray
    v2: np.ndarray
    w1: np.ndarray
    w2: np.ndarray
    k1: np.ndarray
    k2: np.ndarray
    k3: np.ndarray
    vs: np.ndarray
    c0: np.ndarray
    s: Optional[np.ndarray] = None
    name: str = ""
    meta: dict = dataclasses.field(default_factory=dict)

    def new_c(self) -> np.ndarray:
        return np.array(self.c0, order="F", copy=True)

    def new_acum(self) -> np.ndarray:
        return np.zeros((self.nx + 4, self.ny + 4, self.nb), order="F")

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def source_cells(self):
        """Nonzero source entries as (i, j, k, b, rate) with 1-based i, j, k and 0-based b."""
        if self.s is None:
            return []
        idx = np.argwhere(self.s != 0.0)
        return [(int(i) + 1, int(j) + 1, int(k) + 1, int(b), float(self.s[i, j, k, b])) for i, j, k, b in idx]


def _F(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


def uniform_grid(nx, ny, nz, dx, dy, dz):
    """Cartesian metrics of Sources/mod_Grid.F90:485-562 for uniform spacing."""
    return {
        "dX1_p": np.full(nx + 4, float(dx)), "dX1_b": np.full(nx + 3, float(dx)),
        "dX2_p": np.full(ny + 4, float(dy)), "dX2_b": np.full(ny + 3, float(dy)),
        "dX3_p": np.full(nz + 4, float(dz)), "dX3_b": np.full(nz + 3, float(dz)),
    }


def sigma_grid(nx, ny, nz, dx, dy, sigma, x3max):
    """Vertical metrics from sigma levels gl_sigma(1:nz+1), Sources/mod_Grid.F90:541-562."""
    sigma = np.asarray(sigma, dtype=np.float64)
    assert sigma.shape == (nz + 1,)
    g = uniform_grid(nx, ny, nz, dx, dy, 1.0)
    db = np.zeros(nz + 3)  # index 0..nz+2
    for k in range(1, nz + 1):
        db[k] = (sigma[k] - sigma[k - 1]) * x3max
    db[0] = db[1]
    db[nz + 1] = db[nz]
    db[nz + 2] = db[nz]
    dp = np.zeros(nz + 4)  # index -1..nz+2 -> dp[k+1]
    for k in range(1, nz):
        zp = 0.5 * (sigma[k] + sigma[k - 1]) * x3max
        zp1 = 0.5 * (sigma[k + 1] + sigma[k]) * x3max
        dp[k + 1] = zp1 - zp
    dp[0] = dp[2]
    dp[1] = dp[2]
    dp[nz + 1] = dp[nz]
    dp[nz + 2] = dp[nz]
    dp[nz + 3] = dp[nz]
    g["dX3_b"] = db
    g["dX3_p"] = dp
    return g


def constant_diffusion(nx, ny, nz, kh, kv, dtype=np.float64):
    """Constant K fills every element, halos included (Sources/mod_Phys.F90:479-480,590)."""
    k1 = _F((nx + 4, ny, nz), dtype); k1[...] = kh
    k2 = _F((ny + 4, nx, nz), dtype); k2[...] = kh
    k3 = _F((nz + 4, nx, ny), dtype); k3[...] = kv
    return k1, k2, k3


def puff_problem(nx=64, ny=64, nz=32, dx=5000.0, dy=5000.0, dz=500.0, nb=1, vset=0.5,
                 kh=100.0, kv=10.0, sigma_cells=4.0, peak=1.0, zc=8000.0) -> Problem:
    """Gaussian puff in a sheared wind with zero-flux laterals/top and an absorbing ground.

    u = 10 + 5 sin(2 pi z_k / H) (H = domain top), v = 3, w = 0; boundary-face
    normal velocities are 0 (zero flux, Sources/mod_ADS.F90:1243-1321); v_s on
    every w-face except the top face (zero-flux top), so the ground face has
    w_eff = -v_s (deposition).
    """
    H = nz * dz
    g = uniform_grid(nx, ny, nz, dx, dy, dz)
    zk = (np.arange(1, nz + 1) - 0.5) * dz
    u = _F((nx + 3, ny, nz))
    u[...] = (10.0 + 5.0 * np.sin(2.0 * math.pi * zk / H))[None, None, :]
    u[0, :, :] = 0.0; u[1, :, :] = 0.0; u[nx + 1, :, :] = 0.0; u[nx + 2, :, :] = 0.0
    v = _F((ny + 3, nx, nz)); v[...] = 3.0
    v[0, :, :] = 0.0; v[1, :, :] = 0.0; v[ny + 1, :, :] = 0.0; v[ny + 2, :, :] = 0.0
    w = _F((nz + 3, nx, ny))
    k1, k2, k3 = constant_diffusion(nx, ny, nz, kh, kv)
    vs = _F((nx, ny, nz + 3, nb))
    vset = np.broadcast_to(np.asarray(vset, dtype=np.float64), (nb,))
    for b in range(nb):
        vs[:, :, 0:nz + 1, b] = vset[b]
    c = _F((nx + 4, ny + 4, nz + 4, nb))
    xc, yc = 0.5 * nx * dx, 0.5 * ny * dy
    xi = (np.arange(1, nx + 1) - 0.5) * dx
    yj = (np.arange(1, ny + 1) - 0.5) * dy
    sx, sy, sz = sigma_cells * dx, sigma_cells * dy, sigma_cells * dz
    gx = np.exp(-0.5 * ((xi - xc) / sx) ** 2)
    gy = np.exp(-0.5 * ((yj - yc) / sy) ** 2)
    gz = np.exp(-0.5 * ((zk - zc) / sz) ** 2)
    for b in range(nb):
        c[2:nx + 2, 2:ny + 2, 2:nz + 2, b] = peak * gx[:, None, None] * gy[None, :, None] * gz[None, None, :]
    return Problem(nx, ny, nz, nb, g, u, u.copy(order="F"), v, v.copy(order="F"), w, w.copy(order="F"),
                   k1, k2, k3, vs, c, None, name=f"puff{nx}x{ny}x{nz}x{nb}",
                   meta={"dx": dx, "dy": dy, "dz": dz, "kh": kh, "kv": kv})


class FourierWind:
    """Divergence-free horizontal wind from a streamfunction of random modes.

    psi(x, y) = sum_m a_m sin(kx_m x + ky_m y + phi_m); u = dpsi/dy, v = -dpsi/dx,
    wavelengths log-uniform in [lmin, lmax], random directions and phases
    (numpy PCG64 seeded). The field is scaled so that its maximum horizontal
    speed over the domain equals `amplitude`; a linear shear `shear` (m/s per
    km of height) is added to u.
    """

    def __init__(self, seed, nmodes=16, amplitude=15.0, lmin=50e3, lmax=1500e3, shear=2.0,
                 amp_factors=None):
        rng = np.random.default_rng(seed)
        lam = np.exp(rng.uniform(math.log(lmin), math.log(lmax), nmodes))
        theta = rng.uniform(0.0, 2.0 * math.pi, nmodes)
        self.phi = rng.uniform(0.0, 2.0 * math.pi, nmodes)
        kmag = 2.0 * math.pi / lam
        self.kx = kmag * np.cos(theta)
        self.ky = kmag * np.sin(theta)
        # equal velocity amplitude per mode: a_m * |k_m| = 1
        self.a = 1.0 / kmag
        # per-mode perturbation factors (ensembles); applied after the
        # amplitude calibration, which uses the unperturbed modes
        self.factors = None if amp_factors is None else np.asarray(amp_factors, dtype=np.float64)
        self.amplitude = amplitude
        self.shear = shear
        self.scale = 1.0

    def _raw(self, x, y, perturbed=True):
        a = self.a * self.factors if (perturbed and self.factors is not None) else self.a
        ph = np.multiply.outer(x, self.kx) + np.multiply.outer(y, self.ky) + self.phi
        cs = np.cos(ph)
        u = (cs * (a * self.ky)).sum(-1)
        v = -(cs * (a * self.kx)).sum(-1)
        return u, v

    def calibrate(self, Lx, Ly, n=96):
        xs = np.linspace(0.0, Lx, n)
        ys = np.linspace(0.0, Ly, n)
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        u, v = self._raw(X, Y, perturbed=False)
        sp = np.sqrt(u * u + v * v).max()
        self.scale = self.amplitude / sp if sp > 0 else 1.0

    def uv(self, x, y):
        u, v = self._raw(x, y)
        return self.scale * u, self.scale * v


def operational_problem(nx=360, ny=240, nz=60, nb=10, dx=5000.0, dy=5000.0, dz=300.0, seed=1,
                        vmin=0.01, vmax=3.0, kh=100.0, kv=10.0, column_top=15000.0, mass_rate=1.0e6,
                        sources=None, nsources=1, amp_factors=None, dtype_c=np.float64) -> Problem:
    """Operational-size single domain (Fourier wind + shear, line source, log-spaced v_s).

    Wind: FourierWind(seed) evaluated at each face centre; w = 0. Settling
    velocities v_b = vmin*(vmax/vmin)^((b-1)/(nb-1)) on every w-face. Line
    source(s): TOP-HAT-like uniform split of `mass_rate` over bins and over
    the column cells whose centre lies below `column_top`
    (Sources/mod_Src.F90:1406-1423 places a top-hat column; here each source
    column sits in one cell column). With nsources > 1 the source columns
    are drawn from the seeded generator. `sources` may override the list of
    (i, j, column_top, mass_rate) tuples.
    """
    g = uniform_grid(nx, ny, nz, dx, dy, dz)
    Lx, Ly = nx * dx, ny * dy
    fw = FourierWind(seed, amp_factors=amp_factors)
    fw.calibrate(Lx, Ly)
    zk = (np.arange(1, nz + 1) - 0.5) * dz
    # u at x-faces: x = (f-1) dx, y = (j-1/2) dy
    xf = (np.arange(0, nx + 3) - 1.0) * dx
    yc = (np.arange(1, ny + 1) - 0.5) * dy
    X, Y = np.meshgrid(xf, yc, indexing="ij")
    u2d, _ = fw.uv(X, Y)
    u = _F((nx + 3, ny, nz))
    u[...] = u2d[:, :, None] + (fw.shear * 1e-3) * zk[None, None, :]
    # v at y-faces (v layout is y-fastest: v(f, i, k))
    yf = (np.arange(0, ny + 3) - 1.0) * dy
    xc = (np.arange(1, nx + 1) - 0.5) * dx
    Yf, Xc = np.meshgrid(yf, xc, indexing="ij")
    _, v2d = fw.uv(Xc, Yf)
    v = _F((ny + 3, nx, nz))
    v[...] = v2d[:, :, None]
    w = _F((nz + 3, nx, ny))
    k1, k2, k3 = constant_diffusion(nx, ny, nz, kh, kv)
    if nb > 1:
        vset = vmin * (vmax / vmin) ** (np.arange(nb) / (nb - 1))
    else:
        vset = np.array([vmin])
    vs = _F((nx, ny, nz + 3, nb))
    vs[...] = vset[None, None, None, :]
    c = _F((nx + 4, ny + 4, nz + 4, nb))
    s = _F((nx, ny, nz, nb))
    if sources is None:
        if nsources == 1:
            sources = [(nx // 2, ny // 2, column_top, mass_rate)]
        else:
            rng = np.random.default_rng(seed + 7919)
            sources = []
            for _ in range(nsources):
                i = int(rng.integers(nx // 5, 4 * nx // 5)) + 1
                j = int(rng.integers(ny // 5, 4 * ny // 5)) + 1
                sources.append((i, j, column_top, mass_rate))
    for (i, j, top, rate) in sources:
        ks = [k for k in range(1, nz + 1) if zk[k - 1] <= top]
        per = rate / (nb * len(ks))
        for k in ks:
            s[i - 1, j - 1, k - 1, :] += per
    meta = {"dx": dx, "dy": dy, "dz": dz, "kh": kh, "kv": kv, "seed": seed, "vset": vset.tolist(),
            "sources": sources}
    return Problem(nx, ny, nz, nb, g, u, u.copy(order="F"), v, v.copy(order="F"), w, w.copy(order="F"),
                   k1, k2, k3, vs, c, s, name=f"oper{nx}x{ny}x{nz}x{nb}s{seed}", meta=meta)


def config1_wind_problem(nx, ny, nz, nb, dx=5000.0, dz=500.0, seed=0, coarsen=1, vmin=0.05, vmax=2.0,
                         kh=100.0, kv=10.0, nsources=1, column_top=15000.0, mass_rate=1.0e6, puff=True,
                         sigma_cells=4.0, peak=1.0, zc=8000.0, dtype=np.float64) -> Problem:
    """The puff problem's wind generator over open boundaries, with seeded
    direction and sources; fine/coarse pairs for the learned correction.

    Wind: the speed profile u = 10 + 5 sin(2 pi z/H), v = 3 (the puff
    problem's), rotated by a seeded angle; w = 0; boundary faces keep their
    wind (free-flow boundaries). v_s log-spaced vmin..vmax over the bins.
    Sources: nsources line sources (0..column_top, mass_rate split evenly over
    bins and column cells) at seeded cells of the middle half; puff=True adds
    the Gaussian puff (sigma_cells fine cells) at a seeded centre.
    coarsen=2 builds the same physical problem on the grid coarsened 2x
    horizontally (nx, ny are the fine sizes): winds at the coarse faces, each
    line source in the coarse cell that holds its fine cell, the initial field
    as 2x2 means of the fine one.
    """
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.0, 2.0 * math.pi)
    fx = [(int(rng.integers(nx // 4, 3 * nx // 4)) + 1, int(rng.integers(ny // 4, 3 * ny // 4)) + 1)
          for _ in range(nsources)]
    pc = (rng.uniform(0.3, 0.7) * nx * dx, rng.uniform(0.3, 0.7) * ny * dx)
    assert nx % coarsen == 0 and ny % coarsen == 0
    cx, cy, d = nx // coarsen, ny // coarsen, dx * coarsen
    H = nz * dz
    g = uniform_grid(cx, cy, nz, d, d, dz)
    zk = (np.arange(1, nz + 1) - 0.5) * dz
    sp = 10.0 + 5.0 * np.sin(2.0 * math.pi * zk / H)
    ct, st = math.cos(theta), math.sin(theta)
    u = _F((cx + 3, cy, nz), dtype)
    u[...] = (sp * ct - 3.0 * st)[None, None, :]
    v = _F((cy + 3, cx, nz), dtype)
    v[...] = (sp * st + 3.0 * ct)[None, None, :]
    w = _F((nz + 3, cx, cy), dtype)
    k1, k2, k3 = constant_diffusion(cx, cy, nz, kh, kv, dtype)
    vset = vmin * (vmax / vmin) ** (np.arange(nb) / max(nb - 1, 1)) if nb > 1 else np.array([0.5])
    vs = _F((cx, cy, nz + 3, nb), dtype)
    vs[...] = vset[None, None, None, :]
    s = _F((cx, cy, nz, nb), dtype)
    for (i, j) in fx:
        ks = [k for k in range(1, nz + 1) if zk[k - 1] <= column_top]
        ic, jc = (i - 1) // coarsen + 1, (j - 1) // coarsen + 1
        s[ic - 1, jc - 1, [k - 1 for k in ks], :] += mass_rate / (nb * len(ks))
    c = _F((cx + 4, cy + 4, nz + 4, nb), dtype)
    if puff:
        xf = (np.arange(1, nx + 1) - 0.5) * dx
        sx = sigma_cells * dx
        gx = np.exp(-0.5 * ((xf - pc[0]) / sx) ** 2)
        gy = np.exp(-0.5 * (((np.arange(1, ny + 1) - 0.5) * dx - pc[1]) / sx) ** 2)
        gz = np.exp(-0.5 * ((zk - zc) / (sigma_cells * dz)) ** 2)
        if coarsen > 1:
            gx = gx.reshape(cx, coarsen).mean(1)
            gy = gy.reshape(cy, coarsen).mean(1)
        for b in range(nb):
            c[2:cx + 2, 2:cy + 2, 2:nz + 2, b] = peak * gx[:, None, None] * gy[None, :, None] * gz[None, None, :]
    return Problem(cx, cy, nz, nb, g, u, u.copy(order="F"), v, v.copy(order="F"), w, w.copy(order="F"),
                   k1, k2, k3, vs, c, s, name=f"c1wind{cx}x{cy}x{nz}x{nb}s{seed}",
                   meta={"dx": d, "dz": dz, "seed": seed, "theta": theta, "sources": fx, "vset": vset.tolist()})


def latin_hypercube(n, dims, rng):
    """n samples in [-1, 1]^dims with exactly one sample per 1/n stratum along
    every dimension (the reference draws its ensemble perturbations by Latin
    hypercube sampling, Sources/mod_Ensemble.F90:763-766)."""
    strata = np.stack([rng.permutation(n) for _ in range(dims)], axis=1)
    u = (strata + rng.random((n, dims))) / n
    return 2.0 * u - 1.0


@dataclasses.dataclass
class EnsembleMember:
    """One ensemble member: grid size and the perturbed operational-problem
    parameters; problem() builds it (member 0 is the unperturbed run)."""
    nx: int
    ny: int
    nz: int
    nb: int
    column_top: float
    mass_rate: float
    amp_factors: np.ndarray
    wind_seed: int

    def problem(self, dtype_c=np.float64) -> "Problem":
        return operational_problem(self.nx, self.ny, self.nz, self.nb, seed=self.wind_seed,
                                   column_top=self.column_top, mass_rate=self.mass_rate,
                                   amp_factors=self.amp_factors, dtype_c=dtype_c)


def ensemble_members(nmembers=64, nb=6, seed=3, nxy=(96, 256), nzr=(40, 60), column_top=15000.0,
                     mass_rate=1.0e6, height_range=0.30, wind_range=0.20, rate_sigma=0.5, nmodes=16):
    """Ensemble of operational problems of uneven sizes.

    Sizes: nx, ny uniform in nxy, nz uniform in nzr (inclusive). Perturbations
    follow Sources/mod_Ensemble.F90:888-935 (relative: var (1 + r range),
    clipped at 0 for the column height): r in [-1, 1] from a Latin hypercube,
    column height +-height_range, each wind Fourier-mode amplitude
    +-wind_range; the mass rate is multiplied by exp(rate_sigma z), z standard
    normal (a log-normal factor). Member 0 is unperturbed, as the reference's
    first member (Sources/mod_Ensemble.F90:774-781). All members share the
    base wind (seed) and the settling velocities of operational_problem.
    """
    rng = np.random.default_rng(seed)
    nx = rng.integers(nxy[0], nxy[1] + 1, nmembers)
    ny = rng.integers(nxy[0], nxy[1] + 1, nmembers)
    nz = rng.integers(nzr[0], nzr[1] + 1, nmembers)
    lhs = latin_hypercube(max(nmembers - 1, 1), 1 + nmodes, rng)
    z = rng.standard_normal(nmembers)
    out = []
    for m in range(nmembers):
        r = np.zeros(1 + nmodes) if m == 0 else lhs[m - 1]
        top = max(0.0, column_top * (1.0 + height_range * r[0]))
        rate = mass_rate * (1.0 if m == 0 else math.exp(rate_sigma * z[m]))
        out.append(EnsembleMember(int(nx[m]), int(ny[m]), int(nz[m]), nb, top, rate,
                                  1.0 + wind_range * r[1:], seed))
    return out


def random_problem(nx, ny, nz, nb, seed=0, dx=1000.0, dy=1200.0, dz=150.0, wscale=5.0, stretched=False,
                   umax_scale=1.0) -> Problem:
    """Small random problem exercising every branch (signs of every face velocity,
    non-uniform diffusion, rough concentrations with exact zeros and plateaus
    that hit the limiter's EPSILON branch). Used by the parity tests."""
    rng = np.random.default_rng(seed)
    if stretched:
        s = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, nz))])
        s = s / s[-1]
        g = sigma_grid(nx, ny, nz, dx, dy, s, nz * dz)
    else:
        g = uniform_grid(nx, ny, nz, dx, dy, dz)
    u = np.asfortranarray(rng.normal(0.0, wscale * umax_scale, (nx + 3, ny, nz)))
    v = np.asfortranarray(rng.normal(0.0, wscale * umax_scale, (ny + 3, nx, nz)))
    w = np.asfortranarray(rng.normal(0.0, 0.3 * wscale * umax_scale, (nz + 3, nx, ny)))
    w2 = np.asfortranarray(w + rng.normal(0.0, 0.1 * wscale, w.shape))
    k1 = np.asfortranarray(rng.uniform(10.0, 200.0, (nx + 4, ny, nz)))
    k2 = np.asfortranarray(rng.uniform(10.0, 200.0, (ny + 4, nx, nz)))
    k3 = np.asfortranarray(rng.uniform(1.0, 20.0, (nz + 4, nx, ny)))
    vs = np.asfortranarray(rng.uniform(0.0, 1.0, (nx, ny, nz + 3, nb)))
    c = np.asfortranarray(rng.uniform(0.0, 1.0, (nx + 4, ny + 4, nz + 4, nb)))
    # plateaus (EPSILON branch) and exact zeros
    mask = rng.uniform(size=c.shape) < 0.15
    c[mask] = 0.0
    c[2:5, 2:5, :, :] = 0.25
    s = np.asfortranarray(np.where(rng.uniform(size=(nx, ny, nz, nb)) < 0.05,
                                   rng.uniform(0.0, 1e3, (nx, ny, nz, nb)), 0.0))
    return Problem(nx, ny, nz, nb, g, u, u.copy(order="F"), v, v.copy(order="F"), w, w2, k1, k2, k3, vs, c, s,
                   name=f"rand{nx}x{ny}x{nz}x{nb}s{seed}")


def meteo_corner_problem(nx, ny, nz, nb, seed=0, map_v=0, map_h=0, modkh=2, modkv=3, modv=2, dx=5000.0, dy=5000.0,
                         ztop=12000.0, s_time=0.5, dry_deposition=False, diam=None):
    """Synthetic inputs of the meteo update (F3D_update_meteo_term): corner
    fields at two meteo times (sheared winds with small-scale noise, a
    stratified temperature and density, boundary-layer height, u* and a
    Monin-Obukhov length of both signs, |L| < 10 included), terrain with
    slopes for the sigma maps, map factors away from 1, and nb particle
    bins (1 um - 2 mm, densities 800-2600 kg/m3, shape factors 0.5-1, one
    aerosol bin, one aggregate bin; diam overrides the diameters), and USGS
    land-use codes 0-25 at the corners (unknown codes included) for dry
    deposition. Returns (grid, corner, mgrid, opts): grid the ADS spacing
    arrays, the rest the update's arguments."""
    rng = np.random.default_rng(seed)
    x3c = np.linspace(0.0, ztop, nz + 1)                       # sigma levels (corner heights above terrain)
    x3p = 0.5 * (x3c[1:] + x3c[:-1])
    X = (np.arange(nx + 1) * dx)[:, None]
    Y = (np.arange(ny + 1) * dy)[None, :]
    h_c = 400.0 + 300.0 * np.sin(X / (7.0 * dx)) * np.cos(Y / (5.0 * dy)) if map_v else np.zeros((nx + 1, ny + 1))
    hm3_c = 1.0 - h_c / ztop
    z_c = h_c[:, :, None] + x3c[None, None, :] * hm3_c[:, :, None]
    corner = {k: [None, None] for k in ("uc", "vc", "wc", "tc", "tvc", "rhoc", "pblhc", "ustc", "monc")}
    for t in (0, 1):
        sh = 1.0 + 0.2 * t
        zz = z_c / 1000.0
        corner["uc"][t] = np.asfortranarray(sh * (5.0 + 1.5 * zz) + 0.002 * X[:, :, None] / dx + rng.normal(0, 0.5, z_c.shape))
        corner["vc"][t] = np.asfortranarray(-2.0 + 0.8 * zz + 0.003 * Y[:, :, None] / dy + rng.normal(0, 0.5, z_c.shape))
        corner["wc"][t] = np.asfortranarray(rng.normal(0.0, 0.05, z_c.shape))
        tc = 288.0 + 2.0 * t - 6.5e-3 * z_c + rng.normal(0, 0.3, z_c.shape)
        corner["tc"][t] = np.asfortranarray(tc)
        corner["tvc"][t] = np.asfortranarray(tc * (1.0 + 0.61 * 0.005) + rng.normal(0, 0.05, z_c.shape))
        corner["rhoc"][t] = np.asfortranarray(1.225 * np.exp(-z_c / 8500.0) * (1.0 + rng.normal(0, 0.01, z_c.shape)))
        corner["pblhc"][t] = np.asfortranarray(rng.uniform(300.0, 1800.0, (nx + 1, ny + 1)))
        corner["ustc"][t] = np.asfortranarray(rng.uniform(0.05, 0.6, (nx + 1, ny + 1)))
        mon = rng.uniform(-400.0, 400.0, (nx + 1, ny + 1))
        mon[::3, ::4] = rng.uniform(-12.0, 12.0, mon[::3, ::4].shape)  # |L| < 10 clamps
        corner["monc"][t] = np.asfortranarray(mon)
    corner["lusec"] = np.asfortranarray(rng.integers(0, 26, (nx + 1, ny + 1)).astype(np.float64))
    hm3 = 1.0 - c2p_2d_np(h_c) / ztop
    mgrid = {
        "Hm1_p": rng.uniform(0.8, 1.0, ny) if map_h else np.ones(ny),
        "Hm2_p": rng.uniform(0.95, 1.05, ny + 4) if map_h else np.ones(ny + 4),
        "Hm3_p": np.asfortranarray(hm3) if map_v else np.ones((nx, ny), order="F"),
        "dhdx_p": np.asfortranarray(rng.normal(0.0, 0.02, (nx, ny))),
        "dhdy_p": np.asfortranarray(rng.normal(0.0, 0.02, (nx, ny))),
        "z_c": np.asfortranarray(z_c), "h_c": np.asfortranarray(h_c), "X3_c": x3c, "X3_p": x3p,
    }
    g = uniform_grid(nx, ny, nz, dx, dy, ztop / nz)
    diam = np.geomspace(1e-6, 2e-3, nb) if diam is None else np.asarray(diam, dtype=np.float64)
    cat = np.ones(nb, dtype=np.int32)
    agg = np.zeros(nb, dtype=np.int32)
    if nb > 2:
        cat[0] = 2          # an aerosol bin
        agg[nb // 2] = 1    # an aggregate bin
    opts = {"map_h": map_h, "map_v": map_v, "modkh": modkh, "modkv": modkv, "modv": modv,
            "kh0": 250.0, "kv0": 15.0, "rams_cs": 0.2275, "dlon": dx, "dlat": dy, "x3max": ztop,
            "s_time": s_time, "vset_fac": 0.85, "diam": diam, "rho": rng.uniform(800.0, 2600.0, nb),
            "psi": rng.uniform(0.5, 1.0, nb), "cat": cat, "aggregate": agg, "dry_deposition": int(dry_deposition)}
    return g, corner, mgrid, opts


def c2p_2d_np(cc):
    """4-corner average of a 2-D corner field (mass points)."""
    return (cc[:-1, :-1] + cc[1:, :-1] + cc[:-1, 1:] + cc[1:, 1:]) / 4.0
