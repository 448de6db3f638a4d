"""Pins of the surrogate-MLP oracle (oracle/mlp_oracle.py) on the CPU: BF16
rounding against torch's conversion, the FP64 network against torch's GELU
and autograd, and the analytic gradient against finite differences."""
import numpy as np
import pytest

import mlp_oracle as M

# PyTorch is imported inside the tests: a GPU run deselects them, and its
# HIP runtime must not load beside the library's in that process.


def test_bf16_matches_torch_round_to_nearest_even():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 20000),
        np.array([0.0, -0.0, 1.0, -1.0, 1.00390625, 1.01171875, 3.0e38, -3.0e38, 1e-40,
                  np.inf, -np.inf], np.float32)])
    # ties: 1 + 2^-8 rounds to even (1.0), 1 + 3*2^-8 rounds up (1 + 2^-6 ... even)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    got = M.bf16(x)
    np.testing.assert_array_equal(got, ref)
    assert M.bf16(np.array([1.00390625], np.float32))[0] == 1.0
    assert np.isnan(M.bf16(np.array([np.nan], np.float32))[0])


def test_gelu_closed_forms():
    import torch
    assert M.gelu(np.array(0.0)) == 0.0
    x = np.array([-30.0, 30.0])
    np.testing.assert_allclose(M.gelu(x), [0.0, 30.0], atol=1e-12)
    xs = np.linspace(-6, 6, 241)
    ref = torch.nn.functional.gelu(torch.from_numpy(xs), approximate="tanh").numpy()
    np.testing.assert_allclose(M.gelu(xs), ref, rtol=1e-12, atol=1e-14)
    h = 1e-6
    num = (M.gelu(xs + h) - M.gelu(xs - h)) / (2 * h)
    np.testing.assert_allclose(M.gelu_grad(xs), num, rtol=1e-7, atol=1e-8)


def _torch_loss_and_grad(flat, X, t, n_in, hidden):
    import torch
    p = torch.tensor(flat, dtype=torch.float64, requires_grad=True)
    parts, o = [], 0
    for s in [(hidden, n_in), (hidden,), (hidden, hidden), (hidden,), (hidden, hidden),
              (hidden,), (hidden,), (1,)]:
        n = int(np.prod(s))
        parts.append(p[o:o + n].reshape(s))
        o += n
    W1, b1, W2, b2, W3, b3, w4, b4 = parts
    h = torch.from_numpy(X)
    for W, b in ((W1, b1), (W2, b2), (W3, b3)):
        h = torch.nn.functional.gelu(torch.nn.functional.linear(h, W, b), approximate="tanh")
    y = h @ w4 + b4
    loss = ((y - torch.from_numpy(t)) ** 2).sum() / (2 * len(t))
    loss.backward()
    return float(loss.detach()), p.grad.numpy(), y.detach().numpy()


@pytest.mark.parametrize("n_in,hidden,n", [(7, 5, 11), (20, 16, 33)])
def test_fp64_network_matches_torch_autograd(n_in, hidden, n):
    rng = np.random.default_rng(n)
    flat = M.init_params(n_in, hidden, seed=n)
    X = rng.standard_normal((n, n_in))
    t = rng.standard_normal(n)
    loss, g = M.backward(flat, X, t, n_in, hidden)
    tl, tg, ty = _torch_loss_and_grad(flat, X, t, n_in, hidden)
    y, _ = M.forward(flat, X, n_in, hidden)
    np.testing.assert_allclose(y, ty, rtol=1e-12, atol=1e-12)
    assert abs(loss - tl) <= 1e-12 * max(1.0, abs(tl))
    np.testing.assert_allclose(g, tg, rtol=1e-10, atol=1e-12)


def test_gradient_matches_central_differences():
    n_in, hidden, n = 6, 4, 9
    rng = np.random.default_rng(3)
    flat = M.init_params(n_in, hidden, seed=3)
    X = rng.standard_normal((n, n_in))
    t = rng.standard_normal(n)
    _, g = M.backward(flat, X, t, n_in, hidden)
    h = 1e-6
    num = np.zeros_like(flat)
    for i in range(flat.size):
        e = np.zeros_like(flat)
        e[i] = h
        lp, _ = M.backward(flat + e, X, t, n_in, hidden)
        lm, _ = M.backward(flat - e, X, t, n_in, hidden)
        num[i] = (lp - lm) / (2 * h)
    np.testing.assert_allclose(g, num, rtol=1e-6, atol=1e-9)


def test_bf16_mode_is_a_perturbation_of_fp64():
    n_in, hidden, n = 24, 64, 50
    rng = np.random.default_rng(5)
    flat = M.init_params(n_in, hidden, seed=5)
    X = rng.standard_normal((n, n_in))
    t = rng.standard_normal(n)
    l64, g64 = M.backward(flat, X, t, n_in, hidden)
    l16, g16 = M.backward(flat, X, t, n_in, hidden, bf16_mode=True)
    # BF16 keeps 8 significant bits: relative perturbations of order 2^-8.
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 0.05
    assert abs(l16 - l64) / l64 < 0.05
    assert np.linalg.norm(g16 - g64) > 0  # the rounding is really applied
