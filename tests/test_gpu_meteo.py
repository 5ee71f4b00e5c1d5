"""fall3d_ads_update_meteo (F3D_update_meteo_term on the device) against the
CPU restatement (oracle/meteo_oracle.py), through the C ABI.

Tolerances: the device forms every value in FP64 with the reference's
statements; staggering, averaging, CMAQ, constant diffusion and the sigma
corrections use only +, -, *, / and sqrt (correctly rounded on both sides)
and agree to 1e-14 relative; RAMS, the similarity profile and the drag laws
call pow/log10, whose last-bit rounding differs between the device and the
host libm, and the drag laws iterate to a relative tolerance of 1e-6 (the
reference's rtol): winds and diffusion to 1e-12, settling velocities to
1e-5. The FP32 states store the same values rounded to FP32 (1e-6).
"""
import numpy as np
import pytest

from fall3d_amd import ads
from fall3d_amd.problem import meteo_corner_problem
from oracle import meteo_oracle as M
from oracle.ads_oracle import Oracle

pytestmark = pytest.mark.gpu

CASES = [  # (map_v, map_h, modkh, modkv, modv)
    (0, 0, 1, 1, 1), (0, 0, 2, 3, 2), (1, 1, 4, 3, 3), (2, 0, 2, 1, 4),
    (2, 1, 4, 3, 5), (1, 0, 1, 3, 6), (0, 1, 2, 3, 7),
]


def _rel(a, b):
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.where(np.abs(b) > 0, d, np.abs(a))))


@pytest.mark.parametrize("state", ["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES, ids=[f"mv{c[0]}mh{c[1]}kh{c[2]}kv{c[3]}v{c[4]}" for c in CASES])
def test_update_meteo_matches_restatement(case, state):
    map_v, map_h, modkh, modkv, modv = case
    print(f"case meteo {case} {state}", flush=True)
    nx, ny, nz, nb = 23, 17, 11, 5
    g, corner, mg, opts = meteo_corner_problem(nx, ny, nz, nb, seed=sum(case), map_v=map_v, map_h=map_h,
                                               modkh=modkh, modkv=modkv, modv=modv, s_time=0.37)
    ref = M.update_meteo(nx, ny, nz, nb, g["dX1_p"], g["dX2_p"], corner, mg, opts)
    assert ref["converged"]
    a = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 8, state, grid=g)
    a.update_meteo(corner, mg, opts)
    got = a.get_meteo()
    exact_rtol = 1e-14 if state == "fp64" else 1e-6
    for key in ("u1", "u2", "v1", "v2", "w1", "w2", "k1", "k2", "k3", "vs"):
        tol = exact_rtol
        if key in ("k1", "k2", "k3") and (modkh == 4 or modkv == 3) and state == "fp64":
            tol = 1e-12
        if key == "vs":
            tol = 1e-5
        e = _rel(got[key], ref[key])
        print(f"  {key}: max rel {e:.2e}", flush=True)
        assert e <= tol, key
    a.release()


@pytest.mark.parametrize("state", ["fp64", "fp32"])
@pytest.mark.parametrize("modv", [2, 6])
def test_dry_deposition_matches_restatement(modv, state):
    """Dry deposition velocities added at the two lowest faces (bins of all
    four Feng modes, one above 100 um, an aerosol bin; land-use codes 0-25 on
    a sigma map): the device's v_s against the restatement; exp and pow give
    last-bit differences only, the drag law iterates to 1e-6."""
    print(f"case meteo dry deposition modv={modv} {state}", flush=True)
    nx, ny, nz, nb = 23, 17, 11, 5
    g, corner, mg, opts = meteo_corner_problem(nx, ny, nz, nb, seed=21 + modv, map_v=1, modkh=2, modkv=1, modv=modv,
                                               s_time=0.37, dry_deposition=True,
                                               diam=[5e-8, 1e-6, 5e-6, 5e-5, 2e-4])
    ref = M.update_meteo(nx, ny, nz, nb, g["dX1_p"], g["dX2_p"], corner, mg, opts)
    base = M.update_meteo(nx, ny, nz, nb, g["dX1_p"], g["dX2_p"], corner, mg, dict(opts, dry_deposition=0))
    assert ref["converged"] and np.all(ref["vs"][:, :, :2, :4] > base["vs"][:, :, :2, :4])
    a = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 8, state, grid=g)
    a.update_meteo(corner, mg, opts)
    got = a.get_meteo()["vs"]
    e = _rel(got, ref["vs"])
    print(f"  vs: max rel {e:.2e}", flush=True)
    assert e <= 1e-5
    # the deposition part alone (the settling part cancels), to FP64 rounding in the FP64 state
    if state == "fp64":
        a.update_meteo(corner, mg, dict(opts, dry_deposition=0))
        got0 = a.get_meteo()["vs"]
        e2 = _rel(got[:, :, :2, :4] - got0[:, :, :2, :4], ref["vs"][:, :, :2, :4] - base["vs"][:, :, :2, :4])
        print(f"  deposition part: max rel {e2:.2e}", flush=True)
        assert e2 <= 1e-12
    a.release()


def test_step_after_update_matches_oracle_on_the_same_fields():
    """The updated fields drive the step: with the FP64 state, steps on the
    device meteo equal the oracle's steps on the fields read back."""
    nx, ny, nz, nb = 20, 16, 12, 4
    g, corner, mg, opts = meteo_corner_problem(nx, ny, nz, nb, seed=9, map_v=1, modkh=2, modkv=3, modv=2)
    a = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 8, "fp64", grid=g)
    a.update_meteo(corner, mg, opts)
    m = a.get_meteo()
    from fall3d_amd.problem import Problem
    rng = np.random.default_rng(2)
    c0 = np.asfortranarray(rng.uniform(0.0, 1.0, (nx + 4, ny + 4, nz + 4, nb)))
    p = Problem(nx, ny, nz, nb, g, m["u1"], m["u2"], m["v1"], m["v2"], m["w1"], m["w2"], m["k1"], m["k2"],
                m["k3"], m["vs"], c0, None, name="meteo-update")
    o = Oracle(p, "minmod", "rk")
    stime = 0.25
    dt = a.time_step(stime, 2, 0.8)
    u = np.asfortranarray((1 - stime) * p.u1 + stime * p.u2)
    v = np.asfortranarray((1 - stime) * p.v1 + stime * p.v2)
    w = np.asfortranarray((1 - stime) * p.w1 + stime * p.w2)
    assert dt == pytest.approx(o.time_step(u, v, w, p.vs, cfl=0.8), rel=1e-6)
    a.set_state(c0)
    cr, ar = p.new_c(), p.new_acum()
    for it in range(3):
        o.step(dt, it, cr, ar, stime=stime)
    a.run(dt, 3, 0, stime=stime)
    cg, ag = a.get_state()
    assert np.count_nonzero(cg != cr) == 0
    assert np.count_nonzero(ag != ar) == 0
    a.release()


def test_update_meteo_rejects_bad_options():
    nx, ny, nz, nb = 8, 6, 5, 3
    g, corner, mg, opts = meteo_corner_problem(nx, ny, nz, nb, seed=1)
    a = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 8, "fp32", grid=g)
    for bad in ({"modkh": 3}, {"modkv": 2}, {"modv": 8}, {"map_v": 3}, {"map_h": 2}):
        with pytest.raises(ads.Fall3dAdsError):
            a.update_meteo(corner, mg, dict(opts, **bad))
    with pytest.raises(ads.Fall3dAdsError):
        a.get_meteo()  # no meteo yet
    no_lu = {k: v for k, v in corner.items() if k != "lusec"}
    with pytest.raises(ads.Fall3dAdsError, match="dry deposition needs"):
        a.update_meteo(no_lu, mg, dict(opts, dry_deposition=1))
    a.update_meteo(no_lu, mg, opts)  # without dry deposition the land use is not read
    a.release()
