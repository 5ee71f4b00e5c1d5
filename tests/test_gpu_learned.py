"""The learned flux correction end to end on a small grid: the solver
generates (coarse face -> flux correction) rows from fine and coarse runs,
the network trains on the GPU, and the corrected coarse run is compared with
plain coarse KT, both measured against the fine run averaged onto the
coarse grid.
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


def test_collection_and_training_are_reproducible():
    """Rows are kept in a canonical face order (not the order in which waves
    appended them), so two collections of the same problem store the same
    rows in the same order, and training from the same seed gives the same
    loss curve bit for bit."""
    nets = [ads.Network(capacity=400_000, seed=4) for _ in range(2)]
    counts = [learned.collect(n, [1], SIZE, nwarm=10, nsamples=2, gap=1, **KW) for n in nets]
    print(f"rows {counts}", flush=True)
    assert counts[0] == counts[1] > 1000
    (x0, t0), (x1, t1) = nets[0].rows(), nets[1].rows()
    (r0, d0), (r1, d1) = nets[0].refs(), nets[1].refs()
    assert np.array_equal(x0, x1) and np.array_equal(t0, t1)
    assert np.array_equal(r0, r1) and np.array_equal(d0, d1)
    # each of the 2 collection calls stores its faces in strictly increasing
    # key order (bin, axis, k, cross, face): at most one descent, no repeats
    key = [tuple(q) for q in np.column_stack([r0[:, 0] >> 1, r0[:, 0] & 1, r0[:, 3], r0[:, 2], r0[:, 1]])]
    assert sum(key[i + 1] < key[i] for i in range(len(key) - 1)) <= 1
    assert all(key[i + 1] != key[i] for i in range(len(key) - 1))
    losses = [n.train(50, batch=8192, lr=1e-3, seed=7) for n in nets]
    assert np.array_equal(losses[0], losses[1])
    for n in nets:
        n.release()


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


def _corrections_cpu(prob, c0, dt, net, smin, gain):
    """Face corrections (dPhi per axis, device face layouts) of one learned
    step restated on the CPU: a 16-face tile along x is active when one of
    its faces has a cell above max(smin, 1e-30); every face of an active tile
    with stencil maximum S > smin gets gain sgn clamp(S dxp 0.01 y), y the
    network on the face's features, clamped to a quarter of the donor
    cell's mass; all other faces get none."""
    from oracle.learned_oracle import FaceFeatures

    f32 = np.float32
    nx, ny, nz, nb = prob.nx, prob.ny, prob.nz, prob.nb
    c = np.asarray(c0, dtype=np.float32)
    ff = FaceFeatures(prob, c, dt)
    thr = max(f32(smin), f32(1e-30))
    out = [np.zeros((nx + 3, ny, nz, nb), np.float32), np.zeros((nx, ny + 3, nz, nb), np.float32)]
    faces, bits, stats = [], [], {"tiles": 0, "active": 0}
    for axis in (0, 1):
        nf, rows = (nx - 1, [(tr, k, b) for b in range(nb) for k in range(1, nz + 1) for tr in range(1, ny + 1)]) \
            if axis == 0 else (nx, [(f, k, b) for b in range(nb) for k in range(1, nz + 1) for f in range(2, ny + 1)])
        for (r, k, b) in rows:
            for t0 in range(0, nf, 16):
                xs = range(t0, min(t0 + 16, nf))
                if axis == 0:  # face f = x + 2 between cells f - 1 and f of row j = r
                    fl = [(x + 2, r, (x + 1, r), (x + 2, r)) for x in xs]
                else:  # cell i = x + 1, face f = r between rows f - 1 and f
                    fl = [(r, x + 1, (x + 1, r - 1), (x + 1, r)) for x in xs]
                stats["tiles"] += 1
                if not any(max(abs(c[i0 + 1, j0 + 1, k + 1, b]), abs(c[i1 + 1, j1 + 1, k + 1, b])) > thr
                           for (_, _, (i0, j0), (i1, j1)) in fl):
                    continue
                stats["active"] += 1
                for (f, tr, L, R) in fl:
                    fb, S, den = ff.face(axis, b, f, tr, k)
                    if S > f32(smin):
                        faces.append((axis, f, tr, k, b, L, R, S, den))
                        bits.append(fb)
    y = net.forward(np.array(bits, dtype=np.uint16)) if bits else np.zeros(0, np.float32)
    for (axis, f, tr, k, b, L, R, S, den), yv in zip(faces, y):
        sgn = f32(1.0) if den > 0 else f32(-1.0)
        U, D = (L, R) if sgn > 0 else (R, L)
        dxp = f32(f32(0.5) / ff.hdxp[axis][f])
        F = f32(f32(f32(S * dxp) * f32(0.01)) * yv)
        du = f32(f32(1.0) / ff.idxb[axis][U[axis]])
        dd = f32(f32(1.0) / ff.idxb[axis][D[axis]])
        F = min(F, f32(f32(f32(0.25) * max(c[U[0] + 1, U[1] + 1, k + 1, b], f32(0.0))) * du))
        F = max(F, f32(f32(f32(-0.25) * max(c[D[0] + 1, D[1] + 1, k + 1, b], f32(0.0))) * dd))
        v = f32(f32(f32(gain) * sgn) * F)
        if axis == 0:
            out[0][f, tr - 1, k - 1, b] = v
        else:
            out[1][tr - 1, f, k - 1, b] = v
    return out, stats


@pytest.mark.parametrize("threshold", ["absolute", "relative"])
def test_learned_step_applies_the_network_corrections(threshold):
    """One learned step equals the plain step minus the divergence of the
    face corrections restated on the CPU (features from the oracle, y from
    the network's own forward pass, which is pinned against the MLP oracle
    separately). A ragged grid (tiles of 16 faces with partial last tiles),
    a threshold that leaves some tiles inactive (absolute, or relative to the
    field's largest |c| and computed on the device), and an output layer
    scaled so that part of the faces reach the positivity clamp."""
    from fall3d_amd.problem import config1_wind_problem

    p = config1_wind_problem(37, 29, 6, 2, seed=5, dx=5000.0, dz=1000.0, nsources=1, puff=True)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
    g.set_problem(p)
    dt = g.time_step(0.0, 2, 0.9)
    g.run(dt, 4, 0)
    c0, a0 = g.get_state()
    params = ads.Network.initial_params(11)
    params[-65:] *= 300.0  # w3, b3: outputs large enough to meet the clamp
    net = ads.Network(params=params)
    gain = 0.5
    if threshold == "absolute":
        smin = dev_smin = 1e-3 * float(np.abs(c0).max())
    else:  # the device's threshold: fp32(1e-3) * max |c| over the whole array, in FP32
        dev_smin = -1e-3
        smin = float(np.float32(1e-3) * np.float32(np.abs(c0).max()))
    res = []
    for learned in (False, True):
        g.set_state(c0, a0)
        g.set_learned(net if learned else None, smin=dev_smin, gain=gain)
        g.step(dt, 0)
        res.append(g.get_state()[0])
    g.release()
    dphi, stats = _corrections_cpu(p, c0, dt, net, smin, gain)
    print(f"tiles {stats}", flush=True)
    assert 0 < stats["active"] < stats["tiles"]
    nx, ny = p.nx, p.ny
    ix = [np.asarray(1.0 / np.asarray(p.grid[k], dtype=np.float64), dtype=np.float32) for k in ("dX1_b", "dX2_b")]
    dx, dy = dphi
    d = (f32d(dx[2:nx + 2] - dx[1:nx + 1]) * ix[0][1:nx + 1, None, None, None]
         + f32d(dy[:, 2:ny + 2] - dy[:, 1:ny + 1]) * ix[1][None, 1:ny + 1, None, None]).astype(np.float32)
    got = (res[1] - res[0])[2:nx + 2, 2:ny + 2, 2:p.nz + 2, :]
    assert np.abs(d).max() > 0
    err = np.abs(got + d).max()
    print(f"max |correction| {np.abs(d).max():.3e}, max err {err:.3e}", flush=True)
    assert err <= 1e-4 * np.abs(d).max() + 1e-6 * np.abs(res[0]).max()


def f32d(a):
    return np.asarray(a, dtype=np.float32)


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


@pytest.mark.parametrize("poison", ["nan_bias", "inf_weight", "none"])
def test_learned_step_falls_back_to_plain_kt(poison):
    """Plain KT is the learned step's fallback: a network whose output is not
    finite on an active face (a NaN output bias, an infinite first-layer
    weight) leaves the step bitwise equal to the plain KT step and counts one
    fallback; a finite network applies its corrections and counts none."""
    from fall3d_amd.problem import config1_wind_problem

    print(f"case poison={poison}", flush=True)
    p = config1_wind_problem(37, 29, 6, 2, seed=5, dx=5000.0, dz=1000.0, nsources=1, puff=True)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
    g.set_problem(p)
    dt = g.time_step(0.0, 2, 0.9)
    g.run(dt, 4, 0)
    c0, a0 = g.get_state()
    g.set_learned(None)
    g.step(dt, 4)
    plain, _ = g.get_state()
    params = ads.Network.initial_params(3)
    if poison == "nan_bias":
        params[-1] = np.nan
    elif poison == "inf_weight":
        params[5] = np.inf
    net = ads.Network(params=params)
    g.set_state(c0, a0)
    g.set_learned(net, smin=1e-12, gain=1.0)
    g.step(dt, 4)
    got, _ = g.get_state()
    fb = g.learned_fallbacks()
    health = g.check_state()
    g.release()
    net.release()
    print(f"fallbacks {fb}, health {health}", flush=True)
    assert health["nonfinite"] == 0
    if poison == "none":
        assert fb == 0 and not np.array_equal(got, plain)
    else:
        assert fb == 1
        assert np.array_equal(got, plain)


def test_check_state_counts_bad_cells():
    from fall3d_amd.problem import config1_wind_problem

    p = config1_wind_problem(24, 20, 8, 3, seed=2, dx=5000.0, dz=1000.0, nsources=1, puff=True)
    c = p.new_c()
    c[5, 6, 4, 1] = np.nan
    c[7, 3, 2, 0] = np.inf
    c[3, 3, 3, 2] = -2.5e-3
    c[10, 9, 5, 0] = -1e-9
    c[0, 5, 5, 0] = np.nan  # a ghost cell: not part of the interior
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
    g.set_problem(p)
    g.set_state(c)
    h = g.check_state()
    g.release()
    inner = c[2:-2, 2:-2, 2:-2, :].astype(np.float32).astype(np.float64)
    fin = np.isfinite(inner)
    assert h["nonfinite"] == int((~fin).sum()) == 2
    assert h["negative"] == int((inner[fin] < 0).sum()) == 2
    assert h["min"] == float(inner[fin].min())


def test_runs_in_reported_chunks_equal_one_run():
    """The evaluation's long runs go in chunks of an even number of steps
    with a progress line after each; the x/y order follows the iteration
    count, so the state equals that of one run bit for bit."""
    from fall3d_amd.problem import config1_wind_problem

    p = config1_wind_problem(48, 40, 8, 2, seed=3, dx=5000.0, dz=1000.0, nsources=1, puff=True)
    one, chunked = (ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32") for _ in range(2))
    for g in (one, chunked):
        g.set_problem(p)
    dt = one.time_step(0.0, 2, 0.9)
    one.run(dt, 30, 0)
    msgs = []
    learned._run_reporting(chunked, dt, 30, msgs.append, "fine", chunk=8)
    assert msgs == ["fine: 8/30 steps", "fine: 16/30 steps", "fine: 24/30 steps", "fine: 30/30 steps"]
    (c1, a1), (c2, a2) = one.get_state(), chunked.get_state()
    assert np.array_equal(c1, c2) and np.array_equal(a1, a2)
    one.release()
    chunked.release()
