"""Optimiser step rules and the fit loop (optim/*.py of the reference) on the
host: closed-form first steps (which pin the reference's first-step
conventions), convergence on a quadratic, Polyak averaging, fixed indices."""
import numpy as np
import pytest

from spdepy_amd.optim import Adam, AdaDelta, Adagrad, Optimize, RMSprop, SGD


def test_adam_first_step_closed_form():
    # m = g, v = g^2 at t = 1, then bias correction by 1 - beta^t:
    # step = lr * (g / 0.1) / (|g| / sqrt(0.001) + eps) = lr * sqrt(0.001) / 0.1 * sign(g)
    g = np.array([2.0, -0.5, 1e-3])
    x = Adam()(np.zeros(3), 0.0, g, 0.1)
    expect = -0.1 * (g / 0.1) / (np.abs(g) / np.sqrt(0.001) + 1e-8)
    assert np.allclose(x, expect, rtol=1e-12)
    assert np.allclose(np.abs(x[:2]), 0.1 * np.sqrt(0.001) / 0.1, rtol=1e-6)


def test_rmsprop_sgd_adagrad_first_steps():
    g = np.array([3.0, -4.0, 0.0])
    # RMSprop: gf2 = g^2 with zeros replaced by 1, dx = -lr g / |g|
    assert np.allclose(RMSprop()(np.zeros(3), 0, g, 0.5), [-0.5, 0.5, 0.0])
    # SGD: dx = -lr g, then momentum 0.9
    s = SGD()
    x1 = s(np.zeros(3), 0, g, 0.1)
    x2 = s(x1, 0, g, 0.1)
    assert np.allclose(x1, -0.1 * g) and np.allclose(x2, x1 + (0.9 * (-0.1 * g) - 0.1 * g))
    # Adagrad: -lr g / (|g| + 1e-8)
    assert np.allclose(Adagrad()(np.zeros(3), 0, g, 1.0), [-1.0, 1.0, 0.0], atol=1e-8)
    # AdaDelta: dx = -sqrt(eps) / sqrt(0.1 g^2 + eps) g
    d = AdaDelta()(np.zeros(3), 0, g, 1.0)
    assert np.allclose(d, -np.sqrt(1e-8) / np.sqrt(0.1 * g ** 2 + 1e-8) * g)


@pytest.mark.parametrize("step,lr", [("adam", 0.05), ("rmsprop", 0.02), ("sgd", 0.02), ("adagrad", 0.5)])
def test_fit_converges_on_a_quadratic(step, lr):
    A = np.diag([1.0, 3.0, 10.0])
    b = np.array([1.0, -2.0, 0.5])
    xs = np.linalg.solve(A, b)

    def fun(x):
        return 0.5 * x @ A @ x - b @ x, A @ x - b

    res = Optimize(fun).fit(x0=np.zeros(3), stepType=step, lr=lr, max_steps=3000, pol=50)
    assert np.allclose(res["x"], xs, atol=5e-3), (step, res["x"], xs)


def test_polyak_average_and_fixed_indices():
    seen = []

    def fun(x):
        seen.append(x.copy())
        return float(x @ x), 2 * x

    opt = Optimize(fun)
    res = opt.fit(x0=np.array([1.0, 2.0]), stepType="sgd", lr=np.full(6, 0.1), pol=3, fix=[1])
    assert opt.opt_steps == 6
    assert np.allclose(np.array(opt.histX)[:, 1], 2.0)              # index 1 stays fixed
    assert np.allclose(res["x"], np.array(opt.histX)[-3:].mean(0))  # Polyak over the last 3


def test_fit_writes_the_result_vector(tmp_path):
    """Optimize.fit(end=...) saves the Polyak-averaged vector as <end>.npy, the
    format of the reference's fits/*.npy files (optim/__init__.py:55-62)."""
    from spdepy_amd.optim import Optimize
    target = np.array([1.0, -2.0, 0.5])

    def fun(x):
        return float(((x - target) ** 2).sum()), 2 * (x - target)
    opt = Optimize(fun)
    stem = str(tmp_path / "fit_bc1")
    res = opt.fit(x0=np.zeros(3), lr=0.1, max_steps=50, pol=5, end=stem)
    saved = np.load(stem + ".npy", allow_pickle=False)
    assert saved.dtype == np.float64 and np.array_equal(saved, res["x"])


def test_oracle_adam_first_steps_closed_form():
    """The oracle's Adam loop on f(x) = 1/2 |x|^2: the first step moves each
    coordinate by lr * g / (|g| (1 - b1) / sqrt(1 - b2) + eps) with m = g,
    v = g^2; it matches the product rule over a few steps."""
    from oracle import optim as O
    from spdepy_amd.optim import Adam
    fun = lambda x: (0.5 * float(x @ x), x.copy())
    x0 = np.array([1.0, -2.0, 0.5])
    xs, fs, gs, xp = O.adam_fit(fun, x0, 0.1, 5, pol=2)
    g = x0
    first = x0 - 0.1 * (g / 0.1) / (np.abs(g) / np.sqrt(1 - 0.999) + 1e-8)
    assert np.allclose(xs[0], first, rtol=1e-14)
    step = Adam()
    x = x0.copy()
    for k in range(5):
        x = step(x, None, x.copy(), 0.1)
        assert np.allclose(x, xs[k], rtol=1e-14)
    assert np.allclose(xp, xs[-2:].mean(0))
