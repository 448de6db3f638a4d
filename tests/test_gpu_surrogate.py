"""The fitness-surrogate MLP on the GPU (emt6-ro_amd/csrc/mlp_kernels.hip,
mlp.cpp) against the NumPy oracle (oracle/mlp_oracle.py): forward values,
gradients, one Adam step, and training.

Tolerances: against the oracle's BF16-rounding mode the GPU differs only by
FP32 accumulation order and the BF16 roundings that order flips, so the
relative L2 error of outputs and of each gradient tensor is held to 1e-2;
against the plain FP64 network the BF16 storage itself (8 significant bits)
is the error, held to 5e-2 relative L2 and a gradient cosine above 0.995."""
import sys

import numpy as np
import pytest

import mlp_oracle as M

pytestmark = pytest.mark.gpu

N_IN, HID = 152, 512


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-30))


def split(flat, n_in, hidden):
    return M.unflatten(flat, n_in, hidden)


@pytest.fixture(scope="module")
def S():
    import emt6ro.surrogate as S
    return S


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, N_IN)).astype(np.float32)
    t = (0.5 * X[:, 0] + 0.3 * X[:, 1] * X[:, 2] + np.sin(X[:, 3]) +
         0.2 * X[:, :8].sum(axis=1)).astype(np.float32)
    return X, t


def test_params_round_trip_and_count(S):
    log("[surrogate] params")
    net = S.MLP(N_IN, HID, seed=1)
    assert net.num_params == 604161
    p = net.get_params()
    W1, b1, W2, b2, W3, b3, w4, b4 = split(p, N_IN, HID)
    assert np.abs(W1).max() <= 1 / np.sqrt(N_IN) and np.abs(W2).max() <= 1 / np.sqrt(HID)
    q = M.init_params(N_IN, HID, seed=9).astype(np.float32)
    net.set_params(q)
    np.testing.assert_array_equal(net.get_params(), q)
    with pytest.raises(ValueError):
        net.set_params(q[:-1])
    with pytest.raises(RuntimeError, match="multiple of 64"):
        S.MLP(N_IN, 100)
    X = np.zeros((8, N_IN), np.float32)
    for call in (lambda: net.grads(X, np.zeros(7, np.float32)),
                 lambda: net.set_data(X, np.zeros(9, np.float32)),
                 lambda: net.predict(np.zeros((8, N_IN + 1), np.float32))):
        with pytest.raises(ValueError):
            call()


@pytest.mark.parametrize("n", [1, 300, 4096 + 17])
def test_forward_matches_oracle(S, n):
    log(f"[surrogate] forward n={n}")
    net = S.MLP(N_IN, HID, seed=2)
    p = net.get_params().astype(np.float64)
    X, _ = _data(n, n)
    y = net.predict(X)
    y16, _ = M.forward(p, X.astype(np.float64), N_IN, HID, bf16_mode=True)
    y64, _ = M.forward(p, X.astype(np.float64), N_IN, HID)
    e16, e64 = rel(y, y16), rel(y, y64)
    log(f"  rel L2 vs bf16 oracle {e16:.2e}, vs fp64 {e64:.2e}")
    assert e16 < 1e-2
    assert e64 < 5e-2


@pytest.mark.parametrize("n", [700, 768])
def test_gradients_match_oracle(S, n):
    """n = 700 runs the 64 x 64 GEMM tiles, n = 768 (all dimensions multiples
    of 8) the 128 x 128 tiles with transposed operand loads and split K."""
    log(f"[surrogate] gradients n={n}")
    net = S.MLP(N_IN, HID, seed=3)
    p = net.get_params().astype(np.float64)
    X, t = _data(n, 7)
    loss, g = net.grads(X, t)
    l16, g16 = M.backward(p, X.astype(np.float64), t.astype(np.float64), N_IN, HID,
                          bf16_mode=True)
    l64, g64 = M.backward(p, X.astype(np.float64), t.astype(np.float64), N_IN, HID)
    names = ["W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4"]
    for name, a, b, c in zip(names, split(g, N_IN, HID), split(g16, N_IN, HID),
                             split(g64, N_IN, HID)):
        e16, e64 = rel(a, b), rel(a, c)
        cos = float(np.dot(a.ravel(), c.ravel()) /
                    max(np.linalg.norm(a) * np.linalg.norm(c), 1e-30))
        log(f"  {name}: rel L2 vs bf16 oracle {e16:.2e}, vs fp64 {e64:.2e}, cos {cos:.5f}")
        assert e16 < 1e-2, name
        assert e64 < 5e-2 and cos > 0.995, name
    assert abs(loss - l16) / l16 < 1e-3
    assert abs(loss - l64) / l64 < 2e-2


def test_one_adam_step_moves_each_weight_by_lr_against_its_gradient(S):
    """With a data set of one repeated row every minibatch is that row, so the
    first Adam step is w - lr * g / (|g| + eps) (bias-corrected moments equal
    g and g^2): each weight with a clear gradient moves by -lr sign(g)."""
    log("[surrogate] adam step")
    net = S.MLP(N_IN, HID, seed=4)
    p0 = net.get_params().astype(np.float64)
    X, t = _data(1, 11)
    Xr = np.repeat(X, 64, axis=0)
    tr = np.repeat(t, 64)
    _, g16 = M.backward(p0, X.astype(np.float64), t.astype(np.float64), N_IN, HID,
                        bf16_mode=True)
    net.set_data(Xr, tr)
    lr = 1e-3
    losses = net.train(1, batch=32, lr=lr, seed=5)
    y0, _ = M.forward(p0, X.astype(np.float64), N_IN, HID, bf16_mode=True)
    assert abs(losses[0] - (y0[0] - t[0]) ** 2) <= 1e-2 * (y0[0] - t[0]) ** 2
    d = net.get_params().astype(np.float64) - p0
    big = np.abs(g16) > 1e-3 * np.abs(g16).max()
    log(f"  {int(big.sum())} weights with |g| > 1e-3 max|g|")
    assert big.sum() > 1000
    np.testing.assert_allclose(d[big], -lr * np.sign(g16[big]), rtol=2e-3, atol=2e-7)
    assert np.abs(d).max() <= lr * (1 + 1e-3)


def test_training_reduces_loss_and_generalises(S):
    log("[surrogate] training")
    X, t = _data(20000, 13)
    Xh, th = _data(2000, 14)
    net = S.MLP(N_IN, HID, seed=6)
    net.set_data(X, t)
    losses = net.train(600, batch=1024, lr=1e-3, seed=1)
    first, last = losses[:20].mean(), losses[-20:].mean()
    yh = net.predict(Xh)
    mae = float(np.abs(yh - th).mean())
    base = float(np.abs(th - t.mean()).mean())
    log(f"  loss {first:.4f} -> {last:.4f}; held-out MAE {mae:.4f} (predict-the-mean {base:.4f})")
    assert np.isfinite(losses).all()
    # An FP32 PyTorch network of the same shape, trained the same way, ends
    # near loss 0.002 and held-out MAE 0.1 (predict-the-mean 1.0).
    assert last < 0.05 * first
    assert mae < 0.25 * base


def test_graph_replayed_training_equals_direct_launches(S):
    """With set_train_graph(g), train() replays a captured graph of g steps
    and enqueues the remainder directly; with the graph off (the default)
    every kernel is launched directly.  Same seeds, same results, bit for
    bit, over three calls with eager work (predict) and host synchronisation
    between them; the graph is replayed several times per call."""
    log("[surrogate] graph vs direct")
    X, t = _data(6000, 21)
    nets = [S.MLP(N_IN, HID, seed=8), S.MLP(N_IN, HID, seed=8)]
    for n in nets:
        n.set_data(X, t)
    for steps, g, seed in ((37, 8, 1), (20, 8, 2), (11, 3, 3)):
        nets[0].set_train_graph(g)
        la = nets[0].train(steps, batch=1024, lr=1e-3, seed=seed)
        lb = nets[1].train(steps, batch=1024, lr=1e-3, seed=seed)
        np.testing.assert_array_equal(la, lb)
        np.testing.assert_array_equal(nets[0].get_params(), nets[1].get_params())
        np.testing.assert_array_equal(nets[0].predict(X[:300]), nets[1].predict(X[:300]))
    assert np.isfinite(la).all()


def test_adam_clock_carries_across_training_calls(S):
    """With a one-row data set every minibatch is the same, so 20 steps in one
    call and 10 + 10 steps in two calls must give the same weights bit for
    bit: the second call continues the bias corrections at step 11."""
    log("[surrogate] adam clock")
    X, t = _data(1, 23)
    Xr, tr = np.repeat(X, 64, axis=0), np.repeat(t, 64)
    a, b = S.MLP(N_IN, HID, seed=9), S.MLP(N_IN, HID, seed=9)
    for n in (a, b):
        n.set_data(Xr, tr)
    la = a.train(20, batch=32, lr=1e-3, seed=4)
    lb = np.concatenate([b.train(10, batch=32, lr=1e-3, seed=4),
                         b.train(10, batch=32, lr=1e-3, seed=7)])
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(a.get_params(), b.get_params())
    # After set_params the clock starts again at step 1: a network that has
    # taken 20 steps matches a fresh one given the same parameters.
    p = a.get_params()
    c = S.MLP(N_IN, HID, seed=3)
    c.set_data(Xr, tr)
    for n in (a, c):
        n.set_params(p)
    np.testing.assert_array_equal(a.train(3, batch=32, lr=1e-3, seed=4),
                                  c.train(3, batch=32, lr=1e-3, seed=4))
    np.testing.assert_array_equal(a.get_params(), c.get_params())


def test_train_many_equals_training_one_by_one(S):
    """train_many interleaves the steps of several networks on their own
    streams (directly, then as replayed graphs of 4 steps); each network
    must end exactly as training it alone with its seed leaves it, over
    three calls (the Adam clocks carry on)."""
    log("[surrogate] train_many")
    X, t = _data(5000, 31)
    a = [S.MLP(N_IN, HID, seed=20 + i) for i in range(3)]
    b = [S.MLP(N_IN, HID, seed=20 + i) for i in range(3)]
    for n in a + b:
        n.set_data(X, t)
    for steps, base, g in ((25, 100, 0), (9, 200, 0), (10, 300, 4)):
        b[0].set_train_graph(g)
        la = np.stack([n.train(steps, batch=512, lr=1e-3, seed=base + i) for i, n in enumerate(a)])
        lb = S.train_many(b, steps, batch=512, lr=1e-3, seeds=[base, base + 1, base + 2])
        np.testing.assert_array_equal(la, lb)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x.get_params(), y.get_params())
    assert not np.array_equal(a[0].get_params(), a[1].get_params())
    with pytest.raises(Exception):
        S.train_many([b[0], b[0]], 2, batch=64, seeds=[1, 2])
    with pytest.raises(ValueError):
        S.train_many(b, 2, batch=64, seeds=[1])


def test_train_many_at_full_size(S):
    """The GA's ensemble shape: 5 networks of N_FEATURES-512-512-512-1 at
    batch 4096 (64 x 64 tiles, split-K weight gradients) trained together
    end exactly as trained one by one; so do 8 networks and a list mixing
    two shapes."""
    log("[surrogate] train_many full size")
    rng = np.random.default_rng(5)
    X = rng.standard_normal((20000, max(S.N_FEATURES, N_IN + 40))).astype(np.float32)
    t = (X[:, 0] + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    for shapes, batch in (([(S.N_FEATURES, 512)] * 5, 4096), ([(N_IN, HID)] * 8, 256),
                          ([(N_IN, HID), (N_IN + 40, 256), (N_IN, HID)], 256)):
        a = [S.MLP(k, hid, seed=40 + i) for i, (k, hid) in enumerate(shapes)]
        b = [S.MLP(k, hid, seed=40 + i) for i, (k, hid) in enumerate(shapes)]
        for n, (k, _) in zip(a + b, shapes + shapes):
            n.set_data(np.ascontiguousarray(X[:, :k]), t)
        la = np.stack([n.train(3, batch=batch, lr=1e-3, seed=7 + i) for i, n in enumerate(a)])
        lb = S.train_many(b, 3, batch=batch, lr=1e-3, seeds=[7 + i for i in range(len(b))])
        np.testing.assert_array_equal(la, lb)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x.get_params(), y.get_params())


def test_surrogate_standardises_and_errors(S):
    log("[surrogate] wrapper")
    rng = np.random.default_rng(0)
    X = rng.random((3000, S.N_FEATURES)).astype(np.float32) * 2.0
    y = 500.0 + 100.0 * X[:, 0] - 50.0 * X[:, S.N_FEATURES - 1]
    s = S.Surrogate(seed=1, members=2)
    losses = s.fit(X, y, steps=400, batch=512, lr=1e-3, seed=2)
    assert losses[-10:].mean() < losses[:10].mean()
    pred = s.predict(X[:500])
    assert np.abs(pred - y[:500]).mean() < 0.1 * np.abs(y[:500] - y.mean()).mean() + 5.0
    # The ensemble predicts its members' mean.
    xn = s._nx(X[:50])
    members = [net.predict(xn).astype(np.float64) * s.y_sd + s.y_mu for net in s.nets]
    assert np.allclose(s.predict(X[:50]), np.mean(members, axis=0), rtol=1e-6, atol=1e-6)
    assert not np.array_equal(members[0], members[1])
    with pytest.raises(ValueError):
        s.net.predict(np.zeros((2, 3), np.float32))
    with pytest.raises(ValueError):
        S.Surrogate(members=0)
