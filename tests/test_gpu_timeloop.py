"""FALL3D's time loop around the ADS step, on the device (FP64 state),
against the oracle chain run on the same inputs.

Per meteo interval (Sources/mod_F3D.F90): the meteo update on the device
(F3D_update_meteo_term, with dry deposition), the CFL step at mid-interval,
then steps whose interpolation factor s_time advances every step
(F3D_time_step, :1437) with a point source; two intervals in a row. Mid-way a
restart snapshot is taken and a second context resumes from it. At the end
the output post-processing, the cuts and the budget log (F3D_end_time_step).

The oracle side reads the device's meteo fields back (they are checked
against the meteo restatement separately) and steps them with the FP64
oracle: the state, the deposit, the resumed run, the result variables and
the cuts must equal the oracle chain bit for bit.
"""
import numpy as np
import pytest

from fall3d_amd import ads
from fall3d_amd.output import budget_log, cut_variables
from fall3d_amd.problem import Problem, meteo_corner_problem
from oracle import output_oracle as O
from oracle.ads_oracle import Oracle

pytestmark = pytest.mark.gpu


def test_time_loop_two_meteo_intervals_restart_and_output():
    nx, ny, nz, nb = 22, 18, 10, 4
    nsub = 6  # steps per meteo interval
    diam = [5e-8, 2e-6, 3e-5, 2e-4]
    print("case time loop 22x18x10x4 fp64", flush=True)
    intervals = [meteo_corner_problem(nx, ny, nz, nb, seed=31 + q, map_v=1, modkh=2, modkv=1, modv=2,
                                      s_time=0.5, dry_deposition=True, diam=diam) for q in range(2)]
    g = intervals[0][0]
    s = np.zeros((nx, ny, nz, nb), order="F")
    s[nx // 2, ny // 2, 2:6, :] = 1e3  # a short eruption column, every bin
    a = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 8, "fp64", grid=g)
    a.set_source(s)
    a.set_state(None)
    cr = np.zeros((nx + 4, ny + 4, nz + 4, nb), order="F")
    ar = np.zeros((nx + 4, ny + 4, nb), order="F")
    blob, resume_at, it = None, None, 0
    for q, (_, corner, mg, opts) in enumerate(intervals):
        a.update_meteo(corner, mg, opts)
        m = a.get_meteo()
        p = Problem(nx, ny, nz, nb, g, m["u1"], m["u2"], m["v1"], m["v2"], m["w1"], m["w2"], m["k1"], m["k2"],
                    m["k3"], m["vs"], cr, s, name=f"interval {q}")
        o = Oracle(p, "minmod", "rk")
        dt = a.time_step(0.5, 2, 0.9)  # the CFL step at mid-interval
        for k in range(nsub):
            st = (k + 0.5) / nsub      # s_time advances every step
            a.step(dt, it, stime=st)
            o.step(dt, it, cr, ar, stime=st)
            it += 1
            if q == 0 and k == nsub // 2:
                blob, resume_at = a.snapshot(it * dt, it), (q, k + 1, it)
    cg, ag = a.get_state()
    print(f"  {np.count_nonzero(cg != cr)} state values differ, {np.count_nonzero(ag != ar)} deposit values",
          flush=True)
    assert np.count_nonzero(cg != cr) == 0 and np.count_nonzero(ag != ar) == 0
    assert ag.max() > 0.0 and cg.max() > 0.0

    # a second context resumes from the mid-way snapshot with the same inputs
    b = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 8, "fp64", grid=g)
    b.set_source(s)
    q0, k0, it = resume_at
    b.update_meteo(*intervals[0][1:])
    t, iiter = b.restore(blob)
    assert iiter == it
    for q in range(q0, 2):
        if q > q0:
            b.update_meteo(*intervals[q][1:])
        dt = b.time_step(0.5, 2, 0.9)
        for k in range(k0 if q == q0 else 0, nsub):
            b.step(dt, it, stime=(k + 0.5) / nsub)
            it += 1
    cb, ab = b.get_state()
    assert np.array_equal(cb, cg) and np.array_equal(ab, ag)
    b.release()

    # output at the end of the run: result variables, cuts, budget log
    zc = intervals[1][2]["z_c"]
    got = a.postprocess(z_c=zc)
    ref = O.postprocess(cr, ar, z_c=zc)
    for key in got:
        assert np.array_equal(got[key], ref[key]), key
    sigma = np.linspace(0.0, 1.0, nz + 1)
    X3max = float(zc[:, :, -1].max())
    cuts = cut_variables(a, zc, sigma, X3max, "tephra", x_cut=[3.5], z_cut=[500.0, 2500.0], fl_cut=[50.0])
    rc = O.cuts(ref["con"], zc, sigma, X3max, x_cut=[3.5], z_cut=[500.0, 2500.0], fl_cut=[50.0])
    assert np.array_equal(cuts["tephra_con_yz"], rc["cut_x"]) and np.array_equal(cuts["tephra_fl"], rc["cut_fl"])
    bud = a.budget()
    assert bud["mass_volume"] == pytest.approx(o.mass_volume(cr), rel=1e-13)
    log = budget_log(bud, it, dt, it * dt, 0.0, 2 * nsub * dt, "end", bud["injected"])
    print(log, flush=True)
    assert "Mass inside  the domain" in log and "Injected mass" in log
    a.release()
