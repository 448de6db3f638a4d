"""NumPy restatement of the fitness-surrogate MLP -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module, as the checker of the GPU network
(emt6-ro_amd/csrc/mlp.cpp, mlp_kernels.hip).  The surrogate has no counterpart
in the reference (its GA example loads an external forecasting model,
python_ga/example_run_model.py:24-184), so the network is defined here from
the mathematics:

  z_l = h_{l-1} W_l^T + b_l,  h_l = gelu(z_l)  (l = 1..3, h_0 = x)
  y = h_3 . w4 + b4,  loss = sum (y - t)^2 / (2 n)
  gelu(x) = x/2 (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))

`forward` / `backward` work in FP64.  With ``bf16=True`` they round where the
GPU network stores BF16 -- the input, the hidden weights, each z and h, and
each hidden-layer gradient dz -- and keep everything else in FP64, so the GPU
result differs from it only by FP32 accumulation order (and the rare BF16
rounding that order flips).

Pinned by the tests: `backward` against central finite differences of
`forward` (FP64), the FP64 network against torch.nn.functional (gelu
approximate='tanh') with autograd, and `bf16` against torch's FP32->BF16
conversion (round to nearest even).
"""
import numpy as np

SQRT_2_OVER_PI = 0.7978845608028654


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = np.isnan(a)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(SQRT_2_OVER_PI * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(SQRT_2_OVER_PI * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * SQRT_2_OVER_PI * (1.0 + 3 * 0.044715 * x * x)


def unflatten(flat, n_in, hidden):
    """Flat order W1 [H, in], b1, W2, b2, W3, b3, w4 [H], b4 [1]."""
    p = np.asarray(flat, np.float64)
    shapes = [(hidden, n_in), (hidden,), (hidden, hidden), (hidden,), (hidden, hidden),
              (hidden,), (hidden,), (1,)]
    out, o = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(p[o:o + n].reshape(s))
        o += n
    if o != p.size:
        raise ValueError("parameter count mismatch")
    return out


def flatten(parts):
    return np.concatenate([np.asarray(q, np.float64).reshape(-1) for q in parts])


def forward(flat, X, n_in, hidden, bf16_mode=False):
    """Returns (y, cache)."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unflatten(flat, n_in, hidden)
    rnd = bf16 if bf16_mode else (lambda v: np.asarray(v, np.float64))
    Ws = [rnd(W1), rnd(W2), rnd(W3)]
    bs = [b1, b2, b3]
    h = rnd(X)
    hs, zs = [h], []
    for W, b in zip(Ws, bs):
        z = rnd(h @ W.T + b)
        h = rnd(gelu(z))
        zs.append(z)
        hs.append(h)
    y = h @ w4 + b4[0]
    return y, {"Ws": Ws, "w4": w4, "hs": hs, "zs": zs, "rnd": rnd}


def backward(flat, X, t, n_in, hidden, bf16_mode=False):
    """(loss, flat gradient) of loss = sum (y - t)^2 / (2 n)."""
    y, c = forward(flat, X, n_in, hidden, bf16_mode)
    n = len(y)
    d = y - np.asarray(t, np.float64)
    loss = float((d * d).sum() / (2 * n))
    dy = d / n
    rnd, hs, zs, Ws = c["rnd"], c["hs"], c["zs"], c["Ws"]
    gw4 = hs[3].T @ dy
    gb4 = np.array([dy.sum()])
    dz = rnd(dy[:, None] * c["w4"][None, :] * gelu_grad(zs[2]))
    grads = [None] * 6
    for l in (2, 1, 0):
        grads[2 * l] = dz.T @ hs[l]
        grads[2 * l + 1] = dz.sum(axis=0)
        if l > 0:
            dz = rnd((dz @ Ws[l]) * gelu_grad(zs[l - 1]))
    return loss, flatten(grads + [gw4, gb4])


def init_params(n_in, hidden, seed):
    """U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for every weight and bias (any
    stream of such values serves the tests)."""
    rng = np.random.Generator(np.random.PCG64(int(seed)))
    parts = []
    for fan, shape in ((n_in, (hidden, n_in)), (n_in, (hidden,)), (hidden, (hidden, hidden)),
                       (hidden, (hidden,)), (hidden, (hidden, hidden)), (hidden, (hidden,)),
                       (hidden, (hidden,)), (hidden, (1,))):
        b = 1.0 / np.sqrt(fan)
        parts.append(rng.uniform(-b, b, size=shape))
    return flatten(parts)
