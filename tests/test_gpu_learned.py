"""The learned flux correction end to end on a small grid: the solver
generates (coarse face -> flux correction) rows from fine and coarse runs,
the network trains on the GPU, and the corrected coarse run is compared with
plain coarse KT, both judged by the fine run averaged onto the coarse grid.
"""
import numpy as np
import pytest

from fall3d_amd import ads, learned

pytestmark = pytest.mark.gpu

SIZE = (96, 96, 16, 2)  # fine grid; the coarse grid is 48 x 48 x 16
KW = {"dx": 5000.0, "dz": 1000.0, "nsources": 1, "puff": True}


def test_collected_rows_are_finite_and_consistent():
    net = ads.Network(capacity=400_000)
    n = learned.collect(net, [1], SIZE, nwarm=10, nsamples=2, gap=1, **KW)
    print(f"rows {n}", flush=True)
    assert n > 1000 and net.count() == n
    x, t = net.rows()
    xv = ads.bf16_value(x)
    assert np.isfinite(xv).all() and np.isfinite(t).all()
    # normalised stencils: the largest |c| of each row is 1; Courant numbers <= ~1
    assert np.allclose(np.abs(xv[:, :27]).max(axis=1), 1.0, atol=1e-2)
    assert (xv[:, 27] >= 0).all() and (xv[:, 27] <= 1.2).all() and (xv[:, 31] == 0).all()
    net.release()


def test_rows_match_the_cpu_restatement():
    """Every collected row: features bit-identical to oracle/learned_oracle.py
    on the coarse-grained state, targets equal to the restatement from the
    fine and coarse face fluxes."""
    from fall3d_amd.problem import config1_wind_problem
    from oracle.learned_oracle import FaceFeatures, targets

    size = (48, 40, 8, 2)
    kw = {"dx": 5000.0, "dz": 1000.0, "nsources": 1, "puff": True}
    fine, coarse = learned.make_pair(*size, 7, **kw)
    dt = learned.coarse_dt(fine, coarse)
    fine.run(0.5 * dt, 6, 0)
    ads.coarse_grain(fine, coarse)
    c0, _ = coarse.get_state()
    net = ads.Network(capacity=200_000)
    n = ads.learned_collect(fine, coarse, net, dt, 0.0, 3, 1e-12)
    assert n > 100
    x, t = net.rows()
    ref, den = net.refs()
    pc = config1_wind_problem(*size, seed=7, coarsen=2, **kw)
    ff = FaceFeatures(pc, c0, dt)
    for r in range(n):
        q = ref[r]
        bits, S, d = ff.face(q[0] & 1, q[0] >> 1, q[1], q[2], q[3])
        assert np.array_equal(bits, x[r]), (r, q)
        assert d == den[r], (r, q)
    tr = targets(ref, den, (fine.face_flux(0), fine.face_flux(1)), (coarse.face_flux(0), coarse.face_flux(1)))
    assert np.allclose(tr, t, rtol=1e-6, atol=1e-6 * np.abs(tr).max())
    net.release()
    fine.release()
    coarse.release()


def test_trained_correction_beats_plain_coarse_kt():
    net = ads.Network(capacity=3_000_000)
    n = learned.collect(net, range(6), SIZE, nwarm=16, nsamples=6, gap=2, **KW)
    loss = net.train(600, batch=16384, lr=2e-3, seed=3)
    print(f"rows {n}, loss {loss[:5].mean():.4e} -> {loss[-20:].mean():.4e}", flush=True)
    assert loss[-20:].mean() < loss[:5].mean()
    res = learned.evaluate(net, 100, SIZE, 24, **KW)
    print(f"held-out seed: {res}", flush=True)
    assert res["learned"]["load_l1"] < 0.8 * res["plain"]["load_l1"]
    assert res["learned"]["deposit_l1"] < res["plain"]["deposit_l1"]
    net.release()
