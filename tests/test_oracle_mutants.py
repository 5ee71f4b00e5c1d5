"""The oracle pins must catch plausible mistakes.

Each case plants one mistake in a copy of oracle/ads_oracle.c (on the CPU,
in a temporary directory), builds it, and runs the pin tests against that
build (FALL3D_ORACLE_LIB). A pin suite that still passes on a planted
mistake would pin nothing for that convention, so every case must make the
suite fail with an assertion. The comparison with the second restatement
(oracle/ads_literal.py) is not run here: only the closed forms and
invariants count.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "ads_oracle.c")
PINS = [os.path.join(ROOT, "tests", "test_oracle_pins.py"),
        os.path.join(ROOT, "tests", "test_oracle_pins_boundaries.py"),
        os.path.join(ROOT, "tests", "test_oracle_pins_bcmodes.py")]

# (name, text in the oracle, planted replacement, occurrences to replace)
MUTANTS = [
    ("inflow ghost copies the first cell",
     "if (in) line[0] = 2.0 * 0.0 - line[1];", "if (in) line[0] = line[1];", 1),
    ("inflow test not strict",
     "mode == 0 ? (ubeg > 0.0)", "mode == 0 ? (ubeg >= 0.0)", 1),
    ("last-face inflow sign flipped",
     "mode == 0 ? (uend < 0.0)", "mode == 0 ? (uend > 0.0)", 1),
    ("dirichlet ghosts follow the velocity",
     "mode == 0 ? (ubeg > 0.0) : (mode == 1)", "mode <= 1 ? (ubeg > 0.0) : 0", 1),
    ("zero-gradient rule reflects",
     "mode == 0 ? (uend < 0.0) : (mode == 1)", "mode == 0 ? (uend < 0.0) : (mode >= 1)", 1),
    ("second ghost not a copy of the first",
     "    line[-1] = line[0];\n", "    line[-1] = line[1];\n", 1),
    ("reference accumulator with F - P",
     "flux_ref[0] = Fe[0] + Pe[0];", "flux_ref[0] = Fe[0] - Pe[0];", 2),
    ("RK ledger weight of stage 3",
     "g0 += (Fe[0] - Pe[0]) / 3.0; g1 += (Fe[1] - Pe[1]) / 3.0;\n  for (int i = 1; i <= n; ++i) {\n    c0[i] = line[i] + dt",
     "g0 += (Fe[0] - Pe[0]) / 6.0; g1 += (Fe[1] - Pe[1]) / 3.0;\n  for (int i = 1; i <= n; ++i) {\n    c0[i] = line[i] + dt",
     1),
    ("diffusive flux over dxp(i) instead of dxp(i-1)",
     "(c[i] - c[i - 1]) / dxp[i - 1];", "(c[i] - c[i - 1]) / dxp[i];", 1),
    ("RHS divided by dxp",
     "rhs[i] = -1.0 * (F_p - F_m) / dxb[i] + (P_p - P_m) / dxb[i];",
     "rhs[i] = -1.0 * (F_p - F_m) / dxp[i] + (P_p - P_m) / dxp[i];", 1),
    ("RHS divided by the next cell's size",
     "rhs[i] = -1.0 * (F_p - F_m) / dxb[i] + (P_p - P_m) / dxb[i];",
     "rhs[i] = -1.0 * (F_p - F_m) / dxb[i + 1] + (P_p - P_m) / dxb[i + 1];", 1),
    ("face K not averaged",
     "double P_m = 0.5 * (k[i] + k[i - 1])", "double P_m = (k[i])", 1),
    ("upwind flux sign",
     "double F_m = 0.5 * (F_ml + F_mr) - 0.5 * fabs(u[i])", "double F_m = 0.5 * (F_ml + F_mr) + 0.5 * fabs(u[i])", 1),
    ("periodic second ghost off by one",
     "if (u[1] > 0.0) { c[0] = c[n]; c[-1] = c[n - 1];", "if (u[1] > 0.0) { c[0] = c[n]; c[-1] = c[n];", 1),
    ("periodic wrap on outflow",
     "if (u[1] > 0.0) { c[0] = c[n];", "if (u[1] >= 0.0) { c[0] = c[n];", 1),
    ("Dirichlet ghost is the boundary value",
     "c[0] = 2.0 * bvalue[0] - c[1]; c[-1] = c[0];\n  } else if (bcode[0] == 2)",
     "c[0] = bvalue[0]; c[-1] = c[0];\n  } else if (bcode[0] == 2)", 1),
    ("source without the terrain factor",
     "double vol = dX * dY * dZ * (Hm1 * Hm2 * Hm3);", "double vol = dX * dY * dZ * (Hm1 * Hm2);", 1),
    ("source over the vertical cell size",
     "    double dZ = dx3p[k];\n    for (int j = 1; j <= ny; ++j) {\n      double dY = dx2p[j];\n      double Hm1",
     "    double dZ = grid[5][k];\n    for (int j = 1; j <= ny; ++j) {\n      double dY = dx2p[j];\n      double Hm1", 1),
    ("CFL on the east face",
     "2.0 * K1_(i, j, k) / (dx * dx) + fabs(U_(i, j, k)) / dx +",
     "2.0 * K1_(i, j, k) / (dx * dx) + fabs(U_(i + 1, j, k)) / dx +", 1),
    ("minmod EPSILON branch gives 0",
     "      if (lim == LIM_MINMOD) {\n        r[i] = 1.0;", "      if (lim == LIM_MINMOD) {\n        r[i] = 0.0;", 1),
]


@pytest.fixture(scope="module")
def gcc():
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    return cc


@pytest.mark.parametrize("name,old,new,count", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_catch_planted_mistake(tmp_path, gcc, name, old, new, count):
    src = open(SRC).read()
    assert src.count(old) >= count, f"mutation site of '{name}' not found in the oracle"
    mutated = src.replace(old, new, count)
    cfile = tmp_path / "ads_oracle_mutant.c"
    cfile.write_text(mutated)
    so = tmp_path / "libads_oracle_mutant.so"
    subprocess.run([gcc, "-O1", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(so), str(cfile), "-lm"],
                   check=True, capture_output=True)
    env = dict(os.environ, FALL3D_ORACLE_LIB=str(so))
    # the comparison with the literal restatement is left out: the planted
    # mistake must be caught by what the mathematics fixes
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        "-k", "not literal_restatement", *PINS],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    tail = r.stdout[-2000:]
    assert r.returncode == 1 and " failed" in tail, f"pins did not catch '{name}':\n{tail}"
