"""CPU restatement of the ensemble statistics of Sources/task_PosEns.F90
(test infrastructure only: imported by tests/, never by the product path).

x: (nmem, npoints) float64 member fields. Each function cites the
reference lines it restates; ranks follow postp_rank_vector's ascending
order (Sources/mod_Postp.F90:181-318, heapsort into growing order).
"""
from __future__ import annotations

import numpy as np


def mean(x):
    """task_PosEns.F90:321-329: sum over members in member order / nens."""
    s = np.zeros(x.shape[1])
    for m in range(x.shape[0]):
        s = s + x[m]
    return s / x.shape[0]


def logmean(x):
    """:334-343: 10**(sum log10 / nens)."""
    s = np.zeros(x.shape[1])
    with np.errstate(divide="ignore"):
        for m in range(x.shape[0]):
            s = s + np.log10(x[m])
    return 10.0 ** (s / x.shape[0])


def stddev(x):
    """:348-365 divide a sum of squared deviations by nens-1; the sum is
    accumulated here over all members (the reference's loop assigns)."""
    mu = mean(x)
    q = np.zeros(x.shape[1])
    for m in range(x.shape[0]):
        q = q + (x[m] - mu) * (x[m] - mu)
    return np.sqrt(q / (x.shape[0] - 1)) if x.shape[0] > 1 else np.zeros(x.shape[1])


def _ranked(x):
    return np.sort(x, axis=0)  # ascending, as postp_rank_vector


def median(x):
    """:369-383: ens_rank(int(nens/2)), 1-based."""
    n = max(x.shape[0] // 2, 1)
    return _ranked(x)[n - 1]


def percentiles(x, pv):
    """:415-437: n = int(val/100*nens) clamped to 1..nens; ens_rank(n)."""
    r = _ranked(x)
    nens = x.shape[0]
    out = []
    for p in pv:
        n = min(max(int(p / 100.0 * nens), 1), nens)
        out.append(r[n - 1])
    return np.array(out)


def probabilities(x, th):
    """:388-410: 100 * #(x >= th) / nens."""
    nens = x.shape[0]
    return np.array([100.0 * (x >= t).sum(axis=0) / nens for t in th])
