"""Training and evaluation of the learned flux correction (host orchestration).

The solver generates its own training data: a fine run and a coarse run on
the grid coarsened 2x horizontally (fall3d_amd.problem.config1_wind_problem
builds both from one seed); fall3d_ads_learned_collect turns one coarse step
of the coarse-grained fine state into (face features -> flux correction)
rows on the device, and the network trains there (fall3d_mlp_train). Plain
KT on the coarse grid is the fallback and the baseline the correction is
measured against; the reference solution is the fine run averaged onto the
coarse grid.
"""
from __future__ import annotations

import numpy as np

from . import ads
from .problem import config1_wind_problem


def make_pair(nx, ny, nz, nb, seed, dx=5000.0, dz=500.0, nsources=1, puff=True, state="fp32", real_kind=8,
              **kw):
    """Fine and coarse contexts of one seeded problem (fine nx x ny, coarse
    half). real_kind 4 builds the host arrays in float32 (large grids)."""
    dtype = np.float32 if real_kind == 4 else np.float64
    pf = config1_wind_problem(nx, ny, nz, nb, dx, dz, seed, nsources=nsources, puff=puff, dtype=dtype, **kw)
    fine = ads.ADS(pf.nx, pf.ny, pf.nz, pf.nb, "minmod", "rk", real_kind, state)
    fine.set_problem(pf)
    del pf
    pc = config1_wind_problem(nx, ny, nz, nb, dx, dz, seed, coarsen=2, nsources=nsources, puff=puff, dtype=dtype,
                              **kw)
    coarse = ads.ADS(pc.nx, pc.ny, pc.nz, pc.nb, "minmod", "rk", real_kind, "fp32")
    coarse.set_problem(pc)
    return fine, coarse


def coarse_dt(fine, coarse, cfl=0.9):
    """Coarse step: the CFL step of both grids (the fine one takes dt/2)."""
    return min(coarse.time_step(0.0, 2, cfl), 2.0 * fine.time_step(0.0, 2, cfl))


def collect(net, seeds, size, nwarm=20, nsamples=8, gap=2, smin=1e-12, cfl=0.9, **kw):
    """Append training rows from the seeded problems: after nwarm fine
    steps, nsamples coarse steps, gap fine steps apart. Returns rows added."""
    nx, ny, nz, nb = size
    added = 0
    for seed in seeds:
        fine, coarse = make_pair(nx, ny, nz, nb, seed, **kw)
        dt = coarse_dt(fine, coarse, cfl)
        fine.run(0.5 * dt, nwarm, 0)
        it = nwarm // 2
        for _ in range(nsamples):
            added += ads.learned_collect(fine, coarse, net, dt, 0.0, it, smin)
            it += 1
            if gap > 0:
                fine.run(0.5 * dt, 2 * gap, 2 * it)
                it += gap
        fine.release()
        coarse.release()
    return added


def _coarse_mean(a, nx, ny):
    """2x2 horizontal means of the interior of a (nx+4, ny+4, ...) field."""
    return _pairs_mean(a[2:nx + 2, 2:ny + 2])


def _pairs_mean(a):
    """Means over 2x2 blocks of the first two axes (fine i = 2 I + d)."""
    nx, ny = a.shape[0], a.shape[1]
    sh = (2, nx // 2, 2, ny // 2) + a.shape[2:]
    return np.asarray(a).reshape(sh, order="F").mean(axis=(0, 2))


def column_load(c, nx, ny, nz, dz):
    """sum_k c dz over the interior, summed over bins (kg/m^2)."""
    return (c[2:nx + 2, 2:ny + 2, 2:nz + 2, :].sum(axis=(2, 3))) * dz


def rel_l1(a, ref):
    return float(np.abs(a - ref).sum() / max(np.abs(ref).sum(), 1e-300))


def _run_reporting(g, dt, nsteps, progress, what, chunk=200):
    """g.run(dt, nsteps, 0) in chunks of an even number of steps (the x/y
    order follows the iteration count), calling progress(message) after each
    chunk when given."""
    if progress is None:
        g.run(dt, nsteps, 0)
        return
    done = 0
    while done < nsteps:
        n = min(chunk, nsteps - done)
        g.run(dt, n, done)
        g.synchronize()
        done += n
        progress(f"{what}: {done}/{nsteps} steps")


def evaluate(net, seed, size, nsteps, smin=1e-12, cfl=0.9, truth_state="fp32", gain=1.0, fine_fp32=False,
             gains=None, variants=None, real_kind=8, dtype=None, progress=None, **kw):
    """Run fine (2 nsteps steps), coarse plain and coarse learned (nsteps steps
    each) from the same initial state; compare the coarse column loads and
    ground deposits with the fine ones averaged onto the coarse grid.
    fine_fp32=True also runs the fine grid with the FP32 state and compares it
    the same way (when the truth uses the compensated state). gains: several
    correction gains at once (keys "learned_<gain>"); variants: name ->
    keyword arguments of ADS.set_learned (smin, gain, flux_limit).
    real_kind 4 (dtype float32) keeps the host arrays of large grids in
    float32. progress: optional callable given a line of text every few
    hundred steps of each run. Returns a dict of relative L1 errors."""
    nx, ny, nz, nb = size
    dtype = dtype or (np.float32 if real_kind == 4 else np.float64)
    fine, coarse = make_pair(nx, ny, nz, nb, seed, state=truth_state, real_kind=real_kind, **kw)
    dt = coarse_dt(fine, coarse, cfl)
    coarse.release()
    dz = kw.get("dz", 500.0)
    _run_reporting(fine, 0.5 * dt, 2 * nsteps, progress, f"fine {truth_state}")
    cf, af = fine.get_state()
    fine.release()
    lf = _pairs_mean(column_load(cf, nx, ny, nz, dz))
    dep_f = _coarse_mean(af, nx, ny).sum(axis=-1)
    out = {"dt": dt}
    if fine_fp32:
        g = make_pair(nx, ny, nz, nb, seed, state="fp32", real_kind=real_kind, **kw)
        g[1].release()
        g = g[0]
        _run_reporting(g, 0.5 * dt, 2 * nsteps, progress, "fine fp32")
        c2, a2 = g.get_state()
        g.release()
        out["fine_fp32"] = {"load_l1": rel_l1(_pairs_mean(column_load(c2, nx, ny, nz, dz)), lf),
                            "deposit_l1": rel_l1(_coarse_mean(a2, nx, ny).sum(axis=-1), dep_f)}
        del c2, a2
    runs = [("plain", None, {})]
    if variants is not None:
        runs += [(k, net, v) for k, v in variants.items()]
    elif gains is None:
        runs.append(("learned", net, {"smin": smin, "gain": gain}))
    else:
        runs += [(f"learned_{g}", net, {"smin": smin, "gain": g}) for g in gains]
    for name, learned, opts in runs:
        pc = config1_wind_problem(nx, ny, nz, nb, kw.get("dx", 5000.0), dz, seed, coarsen=2, dtype=dtype,
                                  **{k: v for k, v in kw.items() if k not in ("dx", "dz")})
        g = ads.ADS(pc.nx, pc.ny, pc.nz, pc.nb, "minmod", "rk", real_kind, "fp32")
        g.set_problem(pc)
        del pc
        g.set_learned(learned, **opts) if learned is not None else g.set_learned(None)
        _run_reporting(g, dt, nsteps, progress, f"coarse {name}")
        cc, ac = g.get_state()
        g.release()
        lc = column_load(cc, nx // 2, ny // 2, nz, dz)
        dep_c = ac[2:nx // 2 + 2, 2:ny // 2 + 2, :].sum(axis=-1)
        out[name] = {"load_l1": rel_l1(lc, lf), "deposit_l1": rel_l1(dep_c, dep_f)}
    return out
