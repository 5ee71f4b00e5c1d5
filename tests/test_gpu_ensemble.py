"""Ensemble members of uneven sizes advanced by one batched launch per sweep
(fall3d_ads_ensemble_*), through the C ABI.

A member's result must be bit-identical to the same member run alone with
the same tile shape (the batched kernels run the same bodies on the same
arguments), and within the FP32 tolerances of the FP64 oracle. Each member
has its own time step (its own CFL), as separate task_Fall3d runs of a
FALL3D ensemble would.
"""
import numpy as np
import pytest

from fall3d_amd import ads
from fall3d_amd.problem import ensemble_members
from oracle.ads_oracle import Oracle, set_num_threads

pytestmark = pytest.mark.gpu


def _members(n=4, nb=3, seed=11):
    # small uneven members (x widths straddle the 64-lane segment limits)
    return ensemble_members(n, nb=nb, seed=seed, nxy=(20, 90), nzr=(12, 26))


def _fill(ens, specs):
    dts = []
    for g, sp in zip(ens.members, specs):
        p = sp.problem()
        g.set_problem(p)
        dts.append(g.time_step(0.0, 2, 0.9))
    return np.array(dts)


def _alone(sp, tiling, state, tm, dt, nsteps, graph=True):
    p = sp.problem()
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", tm, 8, state)
    g.set_problem(p)
    g.set_tiling(0, cells_per_thread=tiling[0], bins_per_group=p.nb)
    g.set_tiling(1, bins_per_group=tiling[1])
    g.set_tiling(2, bins_per_group=tiling[2])
    g.set_graph(graph)
    g.run(dt, nsteps, 0)
    c, acum = g.get_state()
    b = g.budget()
    g.release()
    return c, acum, b


@pytest.mark.parametrize("state,tm", [("fp32", "rk"), ("fp32x2", "rk"), ("fp32", "euler")])
def test_ensemble_members_bit_identical_to_single_runs(state, tm):
    print(f"case state={state} tm={tm}", flush=True)
    specs = _members()
    ens = ads.Ensemble([(s.nx, s.ny, s.nz) for s in specs], specs[0].nb, "minmod", tm, 8, state)
    dts = _fill(ens, specs)
    nsteps = 9
    ens.run(dts, nsteps, 0)
    ens.synchronize()
    tiling = ens.tiling()
    for m, sp in enumerate(specs):
        c_e, a_e = ens.members[m].get_state()
        b_e = ens.members[m].budget()
        c_s, a_s, b_s = _alone(sp, tiling, state, tm, dts[m], nsteps)
        print(f"member {m} {sp.nx}x{sp.ny}x{sp.nz}: max|dc|={np.abs(c_e - c_s).max():.3e}", flush=True)
        assert np.array_equal(c_e, c_s), m
        assert np.array_equal(a_e, a_s), m
        for k in ("mass_cells", "injected", "ground"):
            assert b_e[k] == b_s[k], (m, k)
        assert np.array_equal(b_e["ledger"], b_s["ledger"]), m
    ens.release()


def test_ensemble_two_wave_lines_bit_identical():
    """Two cells per lane: the wider members' x lines take two waves of a
    workgroup, the narrow one's a single wave, in the same batched launch;
    each member bit-identical to its single run with that tile shape."""
    specs = ensemble_members(3, nb=2, seed=8, nxy=(40, 200), nzr=(8, 12))  # nx 155, 92, 77
    ens = ads.Ensemble([(s.nx, s.ny, s.nz) for s in specs], specs[0].nb, "minmod", "rk", 8, "fp32")
    dts = _fill(ens, specs)
    ens.set_tiling(2, 0, 0)
    tiling = ens.tiling()
    assert tiling[0] == 2 and any(sp.nx // 2 + 1 > 64 for sp in specs) and any(sp.nx // 2 + 1 <= 64 for sp in specs)
    nsteps = 5
    ens.run(dts, nsteps, 0)
    ens.synchronize()
    for m, sp in enumerate(specs):
        c_e, a_e = ens.members[m].get_state()
        c_s, a_s, _ = _alone(sp, tiling, "fp32", "rk", dts[m], nsteps)
        assert np.array_equal(c_e, c_s), m
        assert np.array_equal(a_e, a_s), m
    ens.release()


def test_ensemble_matches_oracle():
    specs = _members(3, nb=2, seed=5)
    ens = ads.Ensemble([(s.nx, s.ny, s.nz) for s in specs], specs[0].nb, "minmod", "rk", 8, "fp32")
    dts = _fill(ens, specs)
    nsteps = 6
    for it in range(nsteps):
        ens.step(dts, it)
    ens.synchronize()
    set_num_threads(8)
    for m, sp in enumerate(specs):
        p = sp.problem()
        o = Oracle(p, "minmod", "rk")
        c, acum = p.new_c(), p.new_acum()
        for it in range(nsteps):
            o.step(dts[m], it, c, acum)
        cg, ag = ens.members[m].get_state()
        e2 = np.linalg.norm(cg - c) / np.linalg.norm(c)
        ea = np.linalg.norm(ag - acum) / max(np.linalg.norm(acum), 1e-300)
        print(f"member {m}: relL2 c={e2:.3e} acum={ea:.3e}", flush=True)
        assert e2 <= 1e-5 and ea <= 1e-5
    ens.release()


def test_ensemble_graph_equals_stepwise():
    specs = _members(3, nb=2, seed=8)
    dims = [(s.nx, s.ny, s.nz) for s in specs]
    out = []
    for graph in (True, False):
        ens = ads.Ensemble(dims, specs[0].nb, "minmod", "rk", 8, "fp32")
        dts = _fill(ens, specs)
        ens.set_graph(graph)
        ens.run(dts, 5, 1)  # odd first iteration: y before x
        ens.synchronize()
        out.append([ens.members[m].get_state()[0] for m in range(len(specs))])
        ens.release()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_ensemble_autotune_restores_state_and_keeps_results():
    specs = _members(4, nb=3, seed=21)
    dims = [(s.nx, s.ny, s.nz) for s in specs]
    ens = ads.Ensemble(dims, specs[0].nb, "minmod", "rk", 8, "fp32")
    dts = _fill(ens, specs)
    ens.run(dts, 4, 0)
    ens.synchronize()
    before = [ens.members[m].get_state() for m in range(len(specs))]
    bud = [ens.members[m].budget() for m in range(len(specs))]
    tiling, ms = ens.autotune(dts, reps=2)
    print(f"autotuned tiling {tiling}, sweep ms {ms}", flush=True)
    assert tiling[0] in (2, 3, 4, 6, 8, 12, 16) and 1 <= tiling[1] <= 8 and 1 <= tiling[2] <= 8
    assert all(t > 0 for t in ms)
    for m in range(len(specs)):
        c, a = ens.members[m].get_state()
        assert np.array_equal(c, before[m][0]) and np.array_equal(a, before[m][1])
        assert ens.members[m].budget()["mass_cells"] == bud[m]["mass_cells"]
    # a run under the tuned shape equals single runs under that shape
    ens.run(dts, 3, 4)
    ens.synchronize()
    for m, sp in enumerate(specs):
        c_s, _, _ = _alone(sp, tiling, "fp32", "rk", dts[m], 7)
        assert np.array_equal(ens.members[m].get_state()[0], c_s), m
    ens.release()


def test_single_domain_autotune_restores_state():
    from fall3d_amd.problem import operational_problem
    p = operational_problem(70, 50, 16, 5, seed=4)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
    g.set_problem(p)
    dt = g.time_step(0.0, 2, 0.9)
    g.run(dt, 3, 0)
    c0, a0 = g.get_state()
    b0 = g.budget()
    tiling, ms = g.autotune(dt, 0.0, 2)
    print(f"tiling {tiling} ms {ms}", flush=True)
    c1, a1 = g.get_state()
    assert np.array_equal(c0, c1) and np.array_equal(a0, a1)
    assert g.budget()["mass_cells"] == b0["mass_cells"]
    g.run(dt, 4, 3)
    c_t = g.get_state()[0]
    g.release()
    # the concentration does not depend on the tile shape
    h = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
    h.set_problem(p)
    h.run(dt, 7, 0)
    assert np.array_equal(h.get_state()[0], c_t)
    h.release()
