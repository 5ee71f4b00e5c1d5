"""Pins of the FP64 oracle of the flux-correction network (CPU only).

Each check is one a plausible slip would fail: the GELU odd-part identity and
its limits, GELU's derivative and every parameter gradient against central
finite differences, BF16 rounding against torch's conversion, and the first
Adam step's closed form.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_gelu_identities():
    x = np.linspace(-8, 8, 401)
    # gelu(x) - gelu(-x) = x exactly for the tanh form
    assert np.allclose(mo.gelu(x) - mo.gelu(-x), x, rtol=0, atol=1e-14)
    assert mo.gelu(np.array(0.0)) == 0.0
    assert abs(mo.gelu(np.array(10.0)) - 10.0) < 1e-12
    assert abs(mo.gelu(np.array(-10.0))) < 1e-12


def test_gelu_grad_finite_difference():
    x = np.linspace(-5, 5, 101)
    h = 1e-6
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    assert np.allclose(mo.gelu_grad(x), fd, rtol=1e-7, atol=1e-9)


def test_parameter_gradients_finite_difference():
    rng = np.random.default_rng(3)
    p = mo.init_params(1)
    p["b1"] = rng.standard_normal(mo.NH) * 0.1
    p["b2"] = rng.standard_normal(mo.NH) * 0.1
    p["b3"] = np.asarray(0.3)
    x = rng.standard_normal((5, mo.NIN))
    t = rng.standard_normal(5)
    _, g = mo.loss_and_grad(p, x, t)
    v = mo.flatten(p)
    gv = mo.flatten(g)
    h = 1e-6
    idx = rng.choice(v.size, 120, replace=False)
    idx = np.concatenate([idx, [v.size - 1]])  # b3
    for i in idx:
        vp, vm = v.copy(), v.copy()
        vp[i] += h
        vm[i] -= h
        lp, _ = mo.loss_and_grad(mo.unflatten(vp), x, t)
        lm, _ = mo.loss_and_grad(mo.unflatten(vm), x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gv[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, gv[i])


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-6, 6, 10000),
                        [1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0]])
    ref = torch.tensor(a, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16_round(a), ref)


def test_adam_first_step_closed_form():
    g = np.array([0.5, -2.0, 1e-3])
    v = np.array([1.0, 1.0, 1.0])
    v1, m, s = mo.adam_step(v, g, np.zeros(3), np.zeros(3), 1, lr=0.01, eps=0.0)
    # bias-corrected first step moves every weight by lr against the sign of g
    assert np.allclose(v1, v - 0.01 * np.sign(g), rtol=0, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(s, 0.001 * g * g)
