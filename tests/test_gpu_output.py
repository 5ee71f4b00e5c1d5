"""fall3d_ads_postprocess (nc_IO_out_res's result variables) on the GPU
against the CPU restatement (oracle/output_oracle.py) of the same state.

The device reads the state it holds; the restatement reads that state as
get_state returns it (FP64, or FP32 widened exactly, hi + lo for the
compensated state, which is the sum the device forms). Both evaluate the
same statements in the same order, so every output must be equal bit for
bit, for the three states, with map factors, terrain-following corner
heights, two species, PM thresholds, an aerosol species (no PM loads) and
Dobson units.
"""
import numpy as np
import pytest

from fall3d_amd import ads
from fall3d_amd.output import budget_log, result_variables
from fall3d_amd.problem import operational_problem
from oracle import output_oracle as O

pytestmark = pytest.mark.gpu


def _zc(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    h = rng.uniform(0.0, 800.0, (nx + 1, ny + 1))
    lev = np.cumsum(np.concatenate([[0.0], rng.uniform(200.0, 400.0, nz)]))
    top = lev[-1]
    return np.asfortranarray(h[:, :, None] + lev[None, None, :] * (1.0 - h[:, :, None] / (2 * top)))


@pytest.mark.parametrize("state", ["fp32", "fp32x2", "fp64"])
def test_postprocess_matches_restatement(state):
    print(f"case postprocess {state}", flush=True)
    p = operational_problem(48, 36, 16, 5, seed=3)
    nx, ny, nz, nb = p.nx, p.ny, p.nz, p.nb
    a = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 8, state)
    a.set_problem(p)
    dt = a.time_step(0.0, 2, 0.9)
    a.run(dt, 12, 0)
    c, acum = a.get_state()
    rng = np.random.default_rng(7)
    hm1, hm2, hm3 = rng.uniform(0.7, 1.0, ny), rng.uniform(0.9, 1.1, ny + 4), rng.uniform(0.85, 1.0, (nx, ny))
    zc = _zc(nx, ny, nz, 5)
    spe = np.array([0, 0, 1, 0, 1], dtype=np.int32)
    diam = np.array([1e-6, 4e-6, 9e-6, 2e-5, 1e-4])
    cases = [dict(species=0, pm_um=(2.5, 10.0)), dict(species=1, so2_du=True, pm_um=(2.5,)),
             dict(species=0, species_category=2, pm_um=(10.0,))]
    for kw in cases:
        got = a.postprocess(z_c=zc, hm1=hm1, hm2=hm2, hm3=hm3, bin_species=spe, bin_diam=diam, **kw)
        ref = O.postprocess(c, acum, z_c=zc, hm1=hm1, hm2=hm2, hm3=hm3, bin_species=spe, bin_diam=diam, **kw)
        for k in got:
            diff = np.count_nonzero(got[k] != ref[k])
            if diff:
                print(f"  {k}: {diff} differ, max {np.abs(got[k] - ref[k]).max():.3e}", flush=True)
            assert diff == 0, (kw, k)
        assert ref["cloud_top"].max() > 0.0 and ref["col_mass"].max() > 0.0
    names = result_variables(a, "tephra", z_c=zc)
    assert "tephra_col_mass" in names and names["tephra_grn_load"].shape == (nx + 1, ny + 1)
    print(budget_log(a.budget(), 12, dt, 12 * dt, 0.0, 100 * dt, "t", a.budget()["injected"]), flush=True)
    a.release()


@pytest.mark.parametrize("state", ["fp32", "fp64"])
def test_cuts_match_restatement(state):
    print(f"case postprocess cuts {state}", flush=True)
    p = operational_problem(40, 30, 14, 4, seed=9)
    nx, ny, nz, nb = p.nx, p.ny, p.nz, p.nb
    a = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 8, state)
    a.set_problem(p)
    dt = a.time_step(0.0, 2, 0.9)
    a.run(dt, 10, 0)
    c, acum = a.get_state()
    zc = _zc(nx, ny, nz, 11)
    X3max = float(zc[:, :, -1].max())
    sigma = np.linspace(0.0, 1.0, nz + 1) ** 1.3
    spe = np.array([0, 1, 0, 0], dtype=np.int32)
    kw = dict(lonmin=-3.0, dlon=0.5, latmin=10.0, dlat=0.25, x_cut=[-2.9, 0.0, 7.75, 99.0], y_cut=[10.1, 13.2],
              z_cut=[50.0, 1500.0, 4000.0, 1e6], fl_cut=[20.0, 100.0, 330.0])
    got = a.postprocess_cuts(zc, sigma, X3max, bin_species=spe, species=0, **kw)
    con = O.postprocess(c, acum, z_c=zc, bin_species=spe, species=0)["con"]
    ref = O.cuts(con, zc, sigma, X3max, **kw)
    for k in ref:
        assert got[k].shape == ref[k].shape, k
        diff = np.count_nonzero(got[k] != ref[k])
        if diff:
            print(f"  {k}: {diff} differ, max {np.abs(got[k] - ref[k]).max():.3e}", flush=True)
        assert diff == 0, k
    assert ref["cut_z"][:, :, 1].max() > 0.0 and ref["cut_fl"].max() > 0.0
    none = a.postprocess_cuts(zc, sigma, X3max, x_cut=[0.5])  # the other kinds may be left out
    assert none["cut_x"].shape == (ny + 1, nz + 1, 1) and none["cut_z"].size == 0
    a.release()


def test_cuts_reject_missing_inputs():
    p = operational_problem(16, 12, 6, 2, seed=1)
    a = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, "fp32")
    a.set_problem(p)
    zc = _zc(p.nx, p.ny, p.nz, 2)
    with pytest.raises(ads.Fall3dAdsError, match="z_c"):
        a.postprocess_cuts(None, np.linspace(0.0, 1.0, p.nz + 1), 5000.0, z_cut=[100.0])
    with pytest.raises(ValueError):
        a.postprocess_cuts(zc, np.linspace(0.0, 1.0, p.nz), 5000.0, z_cut=[100.0])  # nz sigma levels: one short
    a.release()
