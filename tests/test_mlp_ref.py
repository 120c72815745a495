"""The CPU reference of the perceptron's gradient pinned by central finite
differences of its own loss (unquantized FP64), on a tiny padded network,
and the parameter layout helpers."""
import numpy as np

from tiger_nga_amd import mlp

from . import mlp_ref


def test_layout_and_padding():
    shapes = mlp.layer_shapes(38, 37, 64, 2)
    assert shapes == [(64, 64, 64, 38), (64, 64, 64, 64), (64, 64, 37, 64)]
    flat = mlp.init_params(38, 37, 64, 2, seed=1)
    assert flat.size == sum(fo * fi + fo for fo, fi, _, _ in shapes)
    (W0, b0), _, (W2, b2) = mlp.unflatten(flat, 38, 37, 64, 2)
    assert np.all(W0[:, 38:] == 0) and np.all(W2[37:] == 0) and np.all(b2 == 0)
    assert np.abs(W0[:, :38]).max() > 0


def test_reference_gradient_matches_finite_differences():
    nin, nout, hidden, layers = 5, 3, 64, 2
    flat = mlp.init_params(nin, nout, hidden, layers, seed=3).astype(np.float64)
    rng = np.random.default_rng(4)
    # nonzero biases so that every term of the gradient is exercised
    for W, b in mlp.unflatten(flat, nin, nout, hidden, layers):
        b[:] = rng.standard_normal(b.size) * 0.1
    flat[-64 + nout:] = 0.0
    x = np.zeros((6, 64))
    x[:, :nin] = rng.standard_normal((6, nin))
    tgt = rng.standard_normal((6, nout))

    def loss_of(f):
        return mlp_ref.loss_and_grad(mlp.unflatten(f, nin, nout, hidden, layers), x, tgt, nout, quantize=False)[0]
    _, grads = mlp_ref.loss_and_grad(mlp.unflatten(flat, nin, nout, hidden, layers), x, tgt, nout, quantize=False)
    g = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    idx = np.flatnonzero(np.abs(g) > 1e-6)
    pick = rng.choice(idx, 40, replace=False)
    h = 1e-6
    for i in pick:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        fd = (loss_of(fp) - loss_of(fm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5e-3], dtype=np.float64)
    r = mlp_ref.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0
    assert abs(r[4] + 2.5e-3) < 2.5e-3 * 2 ** -8
