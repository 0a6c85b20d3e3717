"""CPU restatement of the reference's optimisation loop with the Adam rule
(test infrastructure only): ``optim/__init__.py:26-62`` (x <- step(x, f,
jac, lr) per evaluation, Polyak average of the last `pol` iterates) and
``optim/adam.py:12-26`` (the first step seeds m = g, v = g^2; bias
corrections by 1 - beta^t)."""
import numpy as np


def adam_fit(fun, x0, lr, steps, pol=10, beta1=0.9, beta2=0.999, eps=1e-8):
    """Returns (iterates, values, gradients, Polyak average)."""
    x = np.asarray(x0, dtype=np.float64).copy()
    m = v = None
    xs, fs, gs = [], [], []
    for t in range(1, steps + 1):
        f, g = fun(x)
        fs.append(f)
        gs.append(g)
        if m is None:
            m, v = g.copy(), g ** 2
        else:
            m = beta1 * m + (1 - beta1) * g
            v = beta2 * v + (1 - beta2) * g ** 2
        mh = m / (1 - beta1 ** t)
        vh = v / (1 - beta2 ** t)
        x = x - lr * mh / (np.sqrt(vh) + eps)
        xs.append(x.copy())
    return np.array(xs), np.array(fs), np.array(gs), np.array(xs)[-pol:].mean(0)
