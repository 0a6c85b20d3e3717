"""Optimisation loop of the drop-in API (berild/spdepy ``src/spdepy/optim``):
``Optimize.fit`` iterates ``x <- step(x, f, jac, lr)`` over the model's
``logLike(x)`` (one GPU evaluation per step) and returns the Polyak average
of the last ``pol`` iterates.  The step rules keep the reference's update
formulas, including their first-step conventions (Adam and RMSprop seed
their moment estimates with the first gradient itself,
``adam.py:13-23``, ``rmsprop.py:11-18``), so a fit follows the same
trajectory for the same gradients."""
import numpy as np


class Adam:
    """optim/adam.py:3-41."""

    def __init__(self, beta1=None, beta2=None, epsilon=None):
        self.beta1 = 0.9 if beta1 is None else beta1
        self.beta2 = 0.999 if beta2 is None else beta2
        self.epsilon = 1e-8 if epsilon is None else epsilon
        self.reset()

    def __call__(self, x, f, gf, lr):
        self.t += 1
        if self.m is None:
            self.m, self.v = gf, gf ** 2
        else:
            self.m = self.beta1 * self.m + (1 - self.beta1) * gf
            self.v = self.beta2 * self.v + (1 - self.beta2) * gf ** 2
        mhat = self.m / (1 - self.beta1 ** self.t)
        vhat = self.v / (1 - self.beta2 ** self.t)
        return x - lr * mhat / (np.sqrt(vhat) + self.epsilon)

    def settings(self, **kw):
        for k in ("beta1", "beta2", "epsilon"):
            if kw.get(k) is not None:
                setattr(self, k, kw[k])
        self.reset()

    def reset(self):
        self.m = self.v = None
        self.t = 0

    def printInit(self):
        pass


class RMSprop:
    """optim/rmsprop.py:3-35."""

    def __init__(self, decay=None, memory=None):
        self.decay = 0.0 if decay is None else decay
        self.memory = 0.9 if memory is None else memory
        self.reset()

    def __call__(self, x, f, gf, lr):
        if self.gf2 is None:
            self.gf2 = gf ** 2
        else:
            self.gf2 = self.memory * self.gf2 + (1 - self.memory) * gf ** 2
        if hasattr(self.gf2, "__len__"):
            self.gf2 = np.where(self.gf2 == 0, 1.0, self.gf2)
        self.dx = self.decay * self.dx - lr * gf / np.sqrt(self.gf2)
        return x + self.dx

    def settings(self, **kw):
        for k in ("decay", "memory"):
            if kw.get(k) is not None:
                setattr(self, k, kw[k])
        self.reset()

    def reset(self):
        self.gf2 = None
        self.dx = 0

    def printInit(self):
        pass


class SGD:
    """optim/sgd.py:3-20 (momentum form)."""

    def __init__(self, decay=None):
        self.decay = 0.9 if decay is None else decay
        self.dx = 0

    def __call__(self, x, f, gf, lr):
        self.dx = self.decay * self.dx - lr * gf
        return x + self.dx

    def settings(self, **kw):
        if kw.get("decay") is not None:
            self.decay = kw["decay"]

    def printInit(self):
        pass


class Adagrad:
    """optim/ada_grad.py:3-30."""

    def __init__(self, epsilon=1e-8):
        self.epsilon = epsilon
        self.reset()

    def __call__(self, x, f, gf, lr):
        self.g2 = gf ** 2 if self.g2 is None else self.g2 + gf ** 2
        self.dx = -lr * gf / (np.sqrt(self.g2) + 1e-8)
        return x + self.dx

    def settings(self, **kw):
        if kw.get("epsilon") is not None:
            self.epsilon = kw["epsilon"]
        self.reset()

    def reset(self):
        self.g2 = None

    def printInit(self):
        pass


class AdaDelta:
    """optim/ada_delta.py:3-40 (lr is not used by this rule)."""

    def __init__(self, rho=None, epsilon=None):
        self.rho = 0.9 if rho is None else rho
        self.epsilon = 1e-8 if epsilon is None else epsilon
        self.reset()

    def __call__(self, x, f, gf, lr):
        if self.egf2 is None:
            self.egf2 = (1 - self.rho) * gf ** 2
            self.dx = -np.sqrt(self.epsilon) / np.sqrt(self.egf2 + self.epsilon) * gf
            self.edx2 = (1 - self.rho) * self.dx ** 2
        else:
            self.egf2 = self.rho * self.egf2 + (1 - self.rho) * gf ** 2
            self.dx = -np.sqrt(self.edx2 + self.epsilon) / np.sqrt(self.egf2 + self.epsilon) * gf
            self.edx2 = self.rho * self.edx2 + (1 - self.rho) * self.dx ** 2
        return x + self.dx

    def settings(self, **kw):
        for k in ("rho", "epsilon"):
            if kw.get(k) is not None:
                setattr(self, k, kw[k])
        self.reset()

    def reset(self):
        self.egf2 = self.dx = self.edx2 = None

    def printInit(self):
        pass


_STEPS = {"adam": Adam, "rmsprop": RMSprop, "sgd": SGD, "adagrad": Adagrad, "adadelta": AdaDelta}


class Optimize:
    """optim/__init__.py:8-196: fit(x0=, fun=, stepType=, pol=, verbose=,
    lr= (scalar or per-step array), max_steps=, fix= (indices kept fixed),
    end= (file stem for the result))."""

    def __init__(self, fun=None):
        self.verbose = False
        self.histX, self.histF, self.histJac = [], [], []
        self.x = self.f = self.jac = None
        self.lr = 0.1
        self.stepType = "adam"
        self.step = Adam()
        self.fun = fun
        self.opt_steps = 0
        self.pol = 10
        self.max_steps = None
        self.fix = None
        self.printPar = lambda x: ""

    def setFun(self, fun):
        self.fun = fun

    def setStep(self, stepType):
        if stepType not in _STEPS:
            raise ValueError("Step type not recognized")
        self.stepType = stepType
        self.step = _STEPS[stepType]()

    def settings(self, **kw):
        self.opt_steps = 0
        if kw.get("x0") is not None:
            self.x = np.asarray(kw["x0"], dtype=np.float64)
        if kw.get("fun") is not None:
            self.setFun(kw["fun"])
        self.setStep(kw["stepType"] if kw.get("stepType") is not None else self.stepType)
        if kw.get("pol") is not None:
            self.pol = kw["pol"]
        if kw.get("verbose") is not None:
            self.verbose = kw["verbose"]
            if self.verbose and kw.get("print") is not None:
                self.printPar = kw["print"]
        if kw.get("max_steps") is not None and self.max_steps is None:
            self.max_steps = kw["max_steps"]
        if kw.get("lr") is not None:
            self.lr = kw["lr"]
            if hasattr(self.lr, "__len__") and self.max_steps is None:
                self.max_steps = len(self.lr)
        self.fix = kw.get("fix")
        self.truth = kw.get("truth")

    def stop(self):
        if self.max_steps is None:
            raise ValueError("Optimize.fit needs max_steps or a per-step learning-rate array")
        return self.opt_steps >= self.max_steps

    def polyak(self):
        return np.array(self.histX).T[:, -self.pol:].mean(axis=1)

    def fit(self, **kw):
        assert self.fun is not None or kw.get("fun") is not None
        self.settings(**kw)
        self.step.settings(**kw)
        self.step.printInit()
        while not self.stop():
            self.f, self.jac = self.fun(self.x)
            if self.verbose:
                print("# %3.0d" % self.opt_steps, "| fun = %2.4f" % self.f, self.printPar(self.x))
            self.histF.append(self.f)
            self.histJac.append(self.jac)
            lr = self.lr[self.opt_steps] if hasattr(self.lr, "__len__") else self.lr
            keep = self.x[self.fix] if self.fix is not None else None
            self.x = self.step(self.x, self.f, self.jac, lr)
            if self.fix is not None:
                self.x[self.fix] = keep
            self.histX.append(self.x)
            self.opt_steps += 1
        self.x = self.polyak()
        res = {"method": self.stepType, "x": self.x, "fun": self.f, "jac": self.jac}
        if kw.get("end") is not None:
            np.save(kw["end"] + ".npy", res["x"])
        return res
