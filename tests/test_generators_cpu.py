"""CPU checks of the host-side generators and entry points that the GPU
workloads rely on: ensemble members (sizes, perturbation rule, Latin
hypercube), fine/coarse problem pairs of the learned correction, and the
bench's refusal of more than one GPU."""
import os
import subprocess
import sys

import numpy as np

from fall3d_amd.problem import config1_wind_problem, ensemble_members, latin_hypercube

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_latin_hypercube_has_one_sample_per_stratum():
    rng = np.random.default_rng(0)
    n, d = 37, 5
    x = latin_hypercube(n, d, rng)
    assert x.shape == (n, d) and (x >= -1).all() and (x <= 1).all()
    for k in range(d):
        strata = np.floor((x[:, k] + 1.0) / 2.0 * n).astype(int)
        assert sorted(strata) == list(range(n))


def test_ensemble_members_follow_the_perturbation_rule():
    ms = ensemble_members(64, nb=6, seed=3)
    assert len(ms) == 64
    for m in ms:
        assert 96 <= m.nx <= 256 and 96 <= m.ny <= 256 and 40 <= m.nz <= 60 and m.nb == 6
        assert 15000.0 * 0.7 - 1e-9 <= m.column_top <= 15000.0 * 1.3 + 1e-9
        assert np.all(np.abs(m.amp_factors - 1.0) <= 0.2 + 1e-12)
        assert m.mass_rate > 0
    # the first member is the unperturbed run (Sources/mod_Ensemble.F90:774-781)
    m0 = ms[0]
    assert m0.column_top == 15000.0 and m0.mass_rate == 1.0e6 and np.all(m0.amp_factors == 1.0)
    # log-normal mass rates: log-factors have a spread near sigma = 0.5
    lf = np.log([m.mass_rate / 1.0e6 for m in ms[1:]])
    assert 0.3 < lf.std() < 0.7
    # same seed, same members
    again = ensemble_members(64, nb=6, seed=3)
    assert all(a.nx == b.nx and a.column_top == b.column_top for a, b in zip(ms, again))


def test_member_problem_has_the_requested_source_mass():
    m = ensemble_members(4, nb=3, seed=9, nxy=(20, 30), nzr=(12, 14))[2]
    p = m.problem()
    assert p.s.sum() > 0
    assert abs(p.s.sum() - m.mass_rate) <= 1e-9 * m.mass_rate or m.column_top < 150.0


def test_coarse_problem_is_the_2x2_mean_of_the_fine_one():
    f = config1_wind_problem(32, 24, 10, 2, seed=5)
    c = config1_wind_problem(32, 24, 10, 2, seed=5, coarsen=2)
    assert (c.nx, c.ny, c.nz) == (16, 12, 10)
    fi = f.c0[2:34, 2:26]
    mean = fi.reshape(2, 16, 2, 12, 14, 2, order="F").mean(axis=(0, 2))
    assert np.allclose(c.c0[2:18, 2:14], mean, rtol=1e-14, atol=0)
    # same total source rate, same winds along the shared faces
    assert abs(f.s.sum() - c.s.sum()) <= 1e-9 * f.s.sum()
    assert np.allclose(c.u1[::1, :, :], c.u1[0:1, 0:1, :])  # horizontally uniform profile
    assert np.allclose(f.u1[0, 0, :], c.u1[0, 0, :])
    assert c.grid["dX1_b"][0] == 2 * f.grid["dX1_b"][0]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
