"""CPU-side checks of the C ABI: the library builds/loads and exports every
entry point include/fall3d_ads.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

import pytest

from fall3d_amd import ads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fall3d_ads.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fall3d_(?:ads|mlp|ens)_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_binding_table():
    assert _declared() == sorted(ads.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(ads.LIB_PATH):
        pytest.skip("library not built (run python -m fall3d_amd.build)")
    lib = ctypes.CDLL(ads.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_no_gpu_initialize_fails_loudly(monkeypatch):
    """Without a visible HIP device the library refuses to create a context
    (there is no CPU fallback)."""
    if not os.path.exists(ads.LIB_PATH):
        pytest.skip("library not built")
    import torch  # noqa: F401  (device count only; does not initialise HIP)
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(ads.Fall3dAdsError):
        ads.ADS(8, 8, 8, 1)


def test_fortran_shim_declares_every_module_ads_procedure():
    """The Fortran replacement MODULE ADS exposes the reference's PUBLIC names
    (Sources/mod_ADS.F90:48-55)."""
    src = open(os.path.join(ROOT, "fall3d_amd", "fortran", "mod_ADS_hip.F90")).read()
    for name in ("ADS_initialize", "ADS_get_vcomps", "ADS_update_horizontal_vcomps", "ADS_release",
                 "ADS_solve_along_x", "ADS_solve_along_y", "ADS_solve_along_z", "ADS_solve_1D"):
        assert re.search(r"PUBLIC\s*::\s*" + name + r"\b", src), name
        assert re.search(r"subroutine\s+" + name + r"\s*\(", src, re.I), name


def test_strided_arrays_are_copied_or_refused():
    """A strided view (1-D included) never reaches the library as is: read-only
    inputs are copied to contiguous memory, writable arguments are refused."""
    import numpy as np

    obj = ads.ADS.__new__(ads.ADS)
    obj.dtype = np.float64
    base = np.arange(20, dtype=np.float64)
    view = base[::2]
    out = obj._arr(view, (10,), "x")
    assert out.flags.f_contiguous and np.array_equal(out, view)
    with pytest.raises(ValueError):
        obj._arr(view, (10,), "x", writable=True)
    c2 = np.zeros((6, 5), order="C")
    assert obj._arr(c2, (6, 5), "y").flags.f_contiguous


def test_initialize_rejects_grids_past_the_32bit_buffer_limit():
    if not os.path.exists(ads.LIB_PATH):
        pytest.skip("library not built")
    with pytest.raises(ads.Fall3dAdsError, match="2 GiB"):
        ads.ADS(4096, 4096, 40, 1)  # one bin of c: 4128 (padded row) * 4100 * 44 * 4 B > 2 GiB


@pytest.mark.parametrize("dims", [(1, 4, 4, 1), (4, 1, 4, 1), (4, 4, 1, 1), (4, 4, 4, 0)])
def test_initialize_rejects_axes_below_two_cells(dims):
    if not os.path.exists(ads.LIB_PATH):
        pytest.skip("library not built")
    with pytest.raises(ads.Fall3dAdsError, match=">= 2 cells"):
        ads.ADS(*dims)
