"""The two known-traffic copies that calibrate the HBM byte counters
(fall3d_ads_copy_probe: 16-byte buffer accesses, fall3d_ads_copy_probe64:
8-byte accesses) copy exactly what they are given, odd tails included, and
touch nothing past it."""
import ctypes

import pytest

from fall3d_amd import ads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 7, 4096 + 3, (1 << 20) + 5])
def test_copy_probes_copy_exactly(n):
    import torch

    print(f"case: copy probes n={n}", flush=True)
    lib = ads.load_library()
    g = torch.Generator(device="cpu").manual_seed(n)
    src = torch.randn(2 * n + 2, generator=g).to("cuda:0")
    dst = torch.full((2 * n + 2,), -7.0, device="cuda:0")
    assert lib.fall3d_ads_copy_probe(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                     ctypes.c_size_t(n), None) == 0
    torch.cuda.synchronize()
    assert torch.equal(dst[:n], src[:n]) and bool((dst[n:] == -7.0).all())
    # the same storage as doubles: n doubles = the first 2n floats
    dst.fill_(-7.0)
    s64, d64 = src.view(torch.float64), dst.view(torch.float64)
    assert lib.fall3d_ads_copy_probe64(ctypes.c_void_p(s64.data_ptr()), ctypes.c_void_p(d64.data_ptr()),
                                       ctypes.c_size_t(n), None) == 0
    torch.cuda.synchronize()
    assert torch.equal(dst[:2 * n], src[:2 * n]) and bool((dst[2 * n:] == -7.0).all())
