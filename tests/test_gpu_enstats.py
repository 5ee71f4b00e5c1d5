"""fall3d_ens_statistics (the statistics of Sources/task_PosEns.F90) on the
GPU against the CPU restatement: mean, standard deviation and log-mean to
FP64 rounding, median, percentiles and exceedance probabilities exactly
(they select member values or count them), for member counts handled by
the register sorting network (<= 64) and by the LDS sort (> 64)."""
import numpy as np
import pytest

from fall3d_amd import ads
from oracle import enstats_oracle as E

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nmem", [2, 5, 8, 13, 64, 100, 300])
def test_statistics_match_restatement(nmem):
    import torch

    print(f"case nmem={nmem}", flush=True)
    rng = np.random.default_rng(nmem)
    npts = 3001
    x = rng.lognormal(0.0, 1.0, (nmem, npts))
    x[:, :10] = 0.75  # ties
    x[: nmem // 2, 10:20] = 2.0
    pv = [5.0, 25.0, 50.0, 90.0, 100.0]
    tv = [0.5, 1.0, 2.0, 10.0]
    r = ads.ensemble_statistics(torch.from_numpy(x).cuda(), pv, tv)
    got = {k: (v.cpu().numpy() if v is not None else None) for k, v in r.items()}
    np.testing.assert_allclose(got["mean"], E.mean(x), rtol=1e-14, atol=0)
    np.testing.assert_allclose(got["stddev"], E.stddev(x), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(got["logmean"], E.logmean(x), rtol=1e-13, atol=0)
    np.testing.assert_array_equal(got["median"], E.median(x))
    np.testing.assert_array_equal(got["per"], E.percentiles(x, pv))
    np.testing.assert_array_equal(got["prob"], E.probabilities(x, tv))
