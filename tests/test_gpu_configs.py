"""The benchmark configurations at their own sizes on the GPU, through the C ABI.

* Operational ensemble: 64 members, nx, ny uniform in 96..256, nz in 40..60,
  6 bins, perturbed column height / mass rate / wind modes. Every member is
  bit-identical to its own single-context run with the same tile shapes, and
  three members match the FP64 oracle.
* Learned-correction training data at 512 x 512 x 40 x 4 (coarse grid
  256 x 256): two seeds of solver-generated rows; a random sample of rows is
  bit-identical to the CPU restatement of the features and equal to the
  restated targets; the network's loss falls when it trains on them.
* The large single domain, 1024 x 768 x 80 x 20: FP32 against the
  compensated state (rel L2 <= 1e-5 of concentration and deposit), mass of
  the compensated state closed to 1e-12, concentration bits independent of
  the tile shapes; single sweeps of full-length x, y and z lines of that
  configuration against the FP64 oracle.
"""
import numpy as np
import pytest

from fall3d_amd import ads, learned
from fall3d_amd.problem import config1_wind_problem, ensemble_members
from oracle.ads_oracle import Oracle, set_num_threads

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5


def _rel_l2_chunked(a, b):
    """Relative L2 of two 4-D arrays, accumulated in float64 bin by bin."""
    num = den = 0.0
    for k in range(a.shape[-1]):
        x = np.asarray(a[..., k], dtype=np.float64)
        y = np.asarray(b[..., k], dtype=np.float64)
        num += float(np.sum((x - y) ** 2))
        den += float(np.sum(y * y))
    return (num / max(den, 1e-300)) ** 0.5


# ---------------------------------------------------------------------------
def test_ensemble_64_members_full_size():
    set_num_threads(16)
    specs = ensemble_members(64, nb=6, seed=3)
    print(f"case: ensemble of {len(specs)} members, {sum(s.nx * s.ny * s.nz * s.nb for s in specs)} cell-bins",
          flush=True)
    ens = ads.Ensemble([(s.nx, s.ny, s.nz) for s in specs], 6, "minmod", "rk", 8, "fp32")
    dts = []
    for g, sp in zip(ens.members, specs):
        p = sp.problem()
        g.set_problem(p)
        dts.append(g.time_step(0.0, 2, 0.9))
    dts = np.array(dts)
    nsteps = 6
    ens.run(dts, nsteps, 0)
    ens.synchronize()
    tiling = ens.tiling()
    # three members against the oracle: the smallest, the largest, member 0
    sizes = [s.nx * s.ny * s.nz for s in specs]
    oracle_members = {0, int(np.argmin(sizes)), int(np.argmax(sizes))}
    for m, sp in enumerate(specs):
        c_e, a_e = ens.members[m].get_state()
        p = sp.problem()
        g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
        g.set_problem(p)
        g.set_tiling(0, cells_per_thread=tiling[0], bins_per_group=p.nb)
        g.set_tiling(1, bins_per_group=tiling[1])
        g.set_tiling(2, bins_per_group=tiling[2])
        g.run(dts[m], nsteps, 0)
        c_s, a_s = g.get_state()
        g.release()
        assert np.array_equal(c_e, c_s), m
        assert np.array_equal(a_e, a_s), m
        if m in oracle_members:
            o = Oracle(p, "minmod", "rk")
            cr, ar = p.new_c(), p.new_acum()
            for it in range(nsteps):
                o.step(dts[m], it, cr, ar)
            e_c = np.linalg.norm(c_e[2:-2, 2:-2, 2:-2] - cr[2:-2, 2:-2, 2:-2]) / np.linalg.norm(cr[2:-2, 2:-2, 2:-2])
            e_a = np.linalg.norm(a_e - ar) / np.linalg.norm(ar)
            print(f"member {m} {sp.nx}x{sp.ny}x{sp.nz}: rel L2 c {e_c:.2e}, deposit {e_a:.2e}", flush=True)
            assert e_c <= REL_L2 and e_a <= REL_L2
    ens.release()


# ---------------------------------------------------------------------------
def test_learned_training_rows_at_512x512x40x4():
    from oracle.learned_oracle import FaceFeatures, targets

    size = (512, 512, 40, 4)
    kw = {"dx": 5000.0, "dz": 500.0, "nsources": 2, "puff": True, "vmin": 0.005, "vmax": 5.0}
    print(f"case: learned rows {size}", flush=True)
    net = ads.Network(capacity=40_000_000)
    # seed 1: rows checked against the CPU restatement
    fine, coarse = learned.make_pair(*size, 1, **kw)
    dt = learned.coarse_dt(fine, coarse)
    fine.run(0.5 * dt, 40, 0)
    ads.coarse_grain(fine, coarse)
    c0, _ = coarse.get_state()
    n = ads.learned_collect(fine, coarse, net, dt, 0.0, 20, 1e-12)
    print(f"seed 1: {n} rows", flush=True)
    assert n > 100_000
    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(n, 3000, replace=False))
    x, t = net.rows()
    ref, den = net.refs()
    pc = config1_wind_problem(*size, seed=1, coarsen=2, **kw)
    ff = FaceFeatures(pc, c0, dt)
    for r in sample:
        q = ref[r]
        bits, S, d = ff.face(q[0] & 1, q[0] >> 1, q[1], q[2], q[3])
        assert np.array_equal(bits, x[r]), (r, q)
        assert d == den[r], (r, q)
    tr = targets(ref[sample], den[sample], (fine.face_flux(0), fine.face_flux(1)),
                 (coarse.face_flux(0), coarse.face_flux(1)))
    assert np.allclose(tr, t[sample], rtol=1e-6, atol=1e-6 * np.abs(tr).max())
    fine.release()
    coarse.release()
    # seed 2 through the collection loop, then training
    n2 = learned.collect(net, [2], size, nwarm=40, nsamples=2, gap=20, **kw)
    print(f"seed 2: {n2} rows", flush=True)
    assert n2 > 100_000 and net.count() == n + n2
    loss = net.train(300, batch=65536, lr=1e-3, seed=5)
    print(f"loss {loss[:10].mean():.4e} -> {loss[-20:].mean():.4e}", flush=True)
    assert np.isfinite(loss).all()
    assert loss[-20:].mean() < 0.8 * loss[:10].mean()
    net.release()


# ---------------------------------------------------------------------------
C5 = (1024, 768, 80, 20)
C5_KW = {"dx": 5000.0, "dz": 250.0, "nsources": 3, "puff": True, "sigma_cells": 60.0, "zc": 6000.0,
         "vmin": 0.005, "vmax": 5.0}


def test_config5_fp32_against_compensated_and_tile_shapes():
    print(f"case: config 5 {C5}", flush=True)
    p = config1_wind_problem(*C5, seed=1234, dtype=np.float32, **C5_KW)
    nsteps = 8
    res = {}
    for name, state, tiling in (("fp32", "fp32", None), ("fp32_tiles", "fp32", (8, 1, 1)),
                                ("fp32x2", "fp32x2", None)):
        g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 4, state)
        g.set_problem(p)
        if tiling is not None:
            g.set_tiling(0, cells_per_thread=tiling[0])
            g.set_tiling(1, cells_per_thread=1, bins_per_group=tiling[1])
            g.set_tiling(2, cells_per_thread=1, bins_per_group=tiling[2])
        dt = g.time_step(0.0, 2, 0.9)
        g.run(dt, nsteps, 0)
        c, a = g.get_state()
        b = g.budget()
        g.release()
        res[name] = (c, a, b)
        print(f"{name}: dt {dt:.3f} s, {nsteps} steps done", flush=True)
    c32, a32, _ = res["fp32"]
    assert np.all(np.isfinite(a32))
    # concentration bits do not depend on the tile shapes
    assert np.array_equal(c32, res["fp32_tiles"][0])
    assert np.array_equal(a32, res["fp32_tiles"][1])
    cx2, ax2, bx2 = res["fp32x2"]
    e_c = _rel_l2_chunked(c32, cx2)
    e_a = np.linalg.norm(a32.astype(np.float64) - ax2) / np.linalg.norm(ax2.astype(np.float64))
    m0 = _initial_mass(p)
    closure = abs(bx2["mass_cells"] + bx2["ledger"].sum() - bx2["injected"] - m0) / m0
    print(f"fp32 vs fp32x2: rel L2 c {e_c:.2e}, deposit {e_a:.2e}; compensated closure {closure:.2e}", flush=True)
    assert e_c <= REL_L2 and e_a <= REL_L2
    assert closure <= 1e-12


def _initial_mass(p):
    """Mass of the initial field in the dX_b measure (uniform grid)."""
    g = p.grid
    vol = float(g["dX1_b"][1]) * float(g["dX2_b"][1]) * float(g["dX3_b"][1])
    tot = 0.0
    for b in range(p.nb):
        tot += float(np.sum(p.c0[2:-2, 2:-2, 2:-2, b], dtype=np.float64))
    return tot * vol


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_config5_full_length_lines_match_oracle(axis):
    """One sweep of lines of the configuration's full length (1024 cells in x,
    768 in y, 80 in z) with its winds, spacing and settling velocities."""
    set_num_threads(16)
    shape = [(1024, 4, 80, 20), (6, 768, 80, 20), (40, 30, 80, 20)][axis]
    print(f"case: config-5 lines along axis {axis}, slab {shape}", flush=True)
    kw = dict(C5_KW, sigma_cells=12.0)
    p = config1_wind_problem(*shape, seed=1234, **kw)
    # winds of the slab: the generator's profile is horizontally uniform, as in the full domain
    o = Oracle(p, "minmod", "rk")
    dt = o.time_step(p.u1, p.v1, p.w1, p.vs, cfl=0.9)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32", grid=p.grid)
    cr, cg = p.new_c(), p.new_c()
    ar, ag = p.new_acum(), p.new_acum()
    if axis == 0:
        u, _, _ = g.get_vcomps()
        u[...] = p.u1
        o.solve_x(dt, cr, p.k1, p.u1)
        g.solve_along_x(0.0, 0.0, dt, cg, p.k1)
    elif axis == 1:
        _, v, _ = g.get_vcomps()
        v[...] = p.v1
        o.solve_y(dt, cr, p.k2, p.v1)
        g.solve_along_y(0.0, 0.0, dt, cg, p.k2)
    else:
        o.solve_z(dt, cr, ar, p.k3, p.w1, p.w2, 0.0, p.vs)
        g.solve_along_z(0.0, 0.0, dt, cg, ag, p.k3, p.w1, p.w2, 0.0, p.vs)
    g.release()
    e = np.linalg.norm(cg - cr) / np.linalg.norm(cr)
    print(f"axis {axis}: rel L2 {e:.2e}", flush=True)
    assert e <= REL_L2
    if axis == 2:
        ea = np.linalg.norm(ag - ar) / np.linalg.norm(ar)
        print(f"deposit rel L2 {ea:.2e}", flush=True)
        assert ea <= REL_L2
