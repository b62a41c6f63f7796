"""santoku.tsetlin.surrogate -- learned proposals for the hyper-parameter search.

An FP32 MLP (12 -> 3x64 ReLU -> 1, tm_mlp_* in include/tsetlin_amd.h, trained
by the engine's own HIP forward/backward/Adam kernels) predicts a trial's dev
accuracy from its hyper-parameters.  Each round the proposer fits it on every
trial outcome so far, scores a large pool of candidates drawn from the search's
own samplers, and proposes the best-scoring ones plus a share of plain draws.
Plain random search stays the fallback (too little history, or a fit that is
worse than predicting the mean on held-out trials) and the judge (the target
accuracy comes from a plain search).  No reference counterpart: optimize.lua
proposes from its samplers only (optimize.lua:391-516).
"""
import ctypes
import math

import numpy as np

from . import TsetlinError
from . import capi as _c
from . import optimize as opt

INPUTS, HIDDEN, NPARAMS = 12, 64, 9217


def init_params(seed):
    """He-normal weights, zero biases, in the tm_mlp parameter layout."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal((HIDDEN, INPUTS)) * math.sqrt(2.0 / INPUTS), np.zeros(HIDDEN),
             rng.standard_normal((HIDDEN, HIDDEN)) * math.sqrt(2.0 / HIDDEN), np.zeros(HIDDEN),
             rng.standard_normal((HIDDEN, HIDDEN)) * math.sqrt(2.0 / HIDDEN), np.zeros(HIDDEN),
             rng.standard_normal(HIDDEN) * math.sqrt(1.0 / HIDDEN), np.zeros(1)]
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def _lib():
    L = _c.load_library()
    if not getattr(L, "_mlp_bound", False):
        fp = ctypes.POINTER(ctypes.c_float)
        vp, u32 = ctypes.c_void_p, ctypes.c_uint32
        sigs = {
            "tm_mlp_create": (ctypes.c_int, [fp, ctypes.POINTER(vp)]),
            "tm_mlp_train": (ctypes.c_int, [vp, fp, fp, u32, u32, ctypes.c_float, fp]),
            "tm_mlp_gradient": (ctypes.c_int, [vp, fp, fp, u32, fp, fp]),
            "tm_mlp_predict": (ctypes.c_int, [vp, fp, u32, fp]),
            "tm_mlp_params": (ctypes.c_int, [vp, fp]),
            "tm_mlp_last_error": (ctypes.c_char_p, []),
            "tm_mlp_destroy": (None, [vp]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        L._mlp_bound = True
    return L


def _f32(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class MLP:
    """The surrogate network on the GPU."""

    def __init__(self, params=None, seed=0):
        self._L = _lib()
        p, pp = _f32(init_params(seed) if params is None else params)
        h = ctypes.c_void_p()
        self._check(self._L.tm_mlp_create(pp, ctypes.byref(h)))
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise TsetlinError("%s: %s" % (_c.ERRORS.get(rc, rc), self._L.tm_mlp_last_error().decode()))

    def train(self, X, Y, steps, lr=3e-3):
        x, xp = _f32(X)
        y, yp = _f32(Y)
        loss = ctypes.c_float()
        self._check(self._L.tm_mlp_train(self._h, xp, yp, y.shape[0], steps, lr, ctypes.byref(loss)))
        return loss.value

    def gradient(self, X, Y):
        x, xp = _f32(X)
        y, yp = _f32(Y)
        g, gp = _f32(np.zeros(NPARAMS))
        loss = ctypes.c_float()
        self._check(self._L.tm_mlp_gradient(self._h, xp, yp, y.shape[0], gp, ctypes.byref(loss)))
        return g, loss.value

    def predict(self, X):
        x, xp = _f32(X)
        out, op = _f32(np.zeros(x.shape[0]))
        self._check(self._L.tm_mlp_predict(self._h, xp, x.shape[0], op))
        return out

    def params(self):
        out, op = _f32(np.zeros(NPARAMS))
        self._check(self._L.tm_mlp_params(self._h, op))
        return out

    def destroy(self):
        if self._h:
            self._L.tm_mlp_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def effective_target(params):
    """The machine's target for these options (capi.c:480-483)."""
    c = int(math.ceil(params["clauses"] / 8.0) * 8)
    t = params.get("target", -1)
    if t is None or t < 0:
        return int(math.sqrt(c / 2.0) * params["clause_tolerance"])
    if t >= 1:
        return int(math.ceil(t))
    return int(math.ceil(max(1.0, float(np.float32(c * t)))))


def features(params, ctx):
    """The 12 surrogate inputs: log clauses, log specificity, target/clauses,
    log target, log tolerance, log maximum, include_bits, epochs, feature
    density, log classes, log features, best dev accuracy so far."""
    c = max(8.0, float(params["clauses"]))
    t = max(1, effective_target(params))
    return [
        math.log2(c) / 12.0,
        math.log(max(1e-3, float(params["specificity"]))) / 4.0,
        t / c,
        math.log2(t) / 12.0,
        math.log2(max(1.0, float(params.get("clause_tolerance", 1)))) / 7.0,
        math.log2(max(1.0, float(params.get("clause_maximum", 1)))) / 7.0,
        float(params.get("include_bits", 1)) / 4.0,
        float(ctx.get("epochs", 10)) / 20.0,
        float(ctx.get("density", 0.5)),
        math.log(max(2.0, float(ctx.get("classes", 2)))) / 5.0,
        math.log2(max(2.0, float(ctx.get("features", 2)))) / 16.0,
        float(ctx.get("best", 0.0)),
    ]


class Proposer:
    """propose(samplers, n, history) for optimize.search_batched."""

    def __init__(self, ctx, param_names=opt.PARAM_NAMES, seed=0, pool=4096, explore=0.25, min_history=48,
                 steps=300, lr=3e-3, key=opt.make_key):
        self.ctx, self.names, self.pool, self.explore = dict(ctx), param_names, pool, explore
        self.min_history, self.steps, self.lr, self.key = min_history, steps, lr, key
        self.rng = np.random.default_rng(seed)
        self.seed = seed
        self.last = {"mode": "plain"}

    def _fit(self, history):
        best = max(s for _, s in history)
        ctx = dict(self.ctx, best=best)
        X = np.array([features(p, ctx) for p, _ in history], np.float32)
        y = np.array([s for _, s in history], np.float32)
        mu, sd = float(y.mean()), float(y.std() + 1e-6)
        Y = (y - mu) / sd
        # held-out check: fit on 80%, score the rest
        idx = self.rng.permutation(len(y))
        cut = max(1, int(0.8 * len(y)))
        net = MLP(seed=self.seed + len(y))
        net.train(X[idx[:cut]], Y[idx[:cut]], self.steps, self.lr)
        pred = net.predict(X[idx[cut:]])
        resid = ((pred - Y[idx[cut:]]) ** 2).sum()
        total = ((Y[idx[cut:]] - Y[idx[:cut]].mean()) ** 2).sum() + 1e-9
        r2 = 1.0 - resid / total
        net.destroy()
        net = MLP(seed=self.seed + len(y))
        loss = net.train(X, Y, self.steps, self.lr)
        return net, ctx, r2, loss

    def __call__(self, samplers, n, history):
        plain = [opt.sample_params(samplers, self.names) for _ in range(n)]
        if len(history) < self.min_history:
            self.last = {"mode": "plain", "reason": "history %d < %d" % (len(history), self.min_history)}
            return plain
        net, ctx, r2, loss = self._fit(history)
        if not r2 > 0.0:
            net.destroy()
            self.last = {"mode": "plain", "reason": "held-out R2 %.3f" % r2}
            return plain
        # candidate pool: draws around the search centre plus global draws
        cand = [opt.sample_params(samplers, self.names) for _ in range(self.pool // 2)]
        for _ in range(self.pool - len(cand)):
            p = {}
            for name in self.names:
                s = samplers.get(name)
                if s is not None:
                    p[name] = s.sample(None) if s.type == "range" else s.sample()
            cand.append(p)
        X = np.array([features(opt._swap_tolerance(dict(p)), ctx) for p in cand], np.float32)
        score = net.predict(X)
        net.destroy()
        seen = {self.key(p) for p, _ in history}
        n_model = n - int(round(self.explore * n))
        picks = []
        for i in np.argsort(-score):
            k = self.key(cand[i])
            if k in seen:
                continue
            seen.add(k)
            picks.append(cand[i])
            if len(picks) >= n_model:
                break
        n_picked = len(picks)
        picks += plain[: n - len(picks)]
        # the first n_model proposals are the network's picks, the rest plain draws
        self.last = {"mode": "surrogate", "r2": float(r2), "loss": float(loss), "n_model": n_picked}
        return picks
