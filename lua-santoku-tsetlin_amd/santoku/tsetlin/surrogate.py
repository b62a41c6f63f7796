"""santoku.tsetlin.surrogate -- learned proposals for the hyper-parameter search.

An FP32 MLP (12 -> 3x64 ReLU -> 1, tm_mlp_* in include/tsetlin_amd.h, trained
by the engine's own HIP forward/backward/Adam kernels) predicts a trial's dev
accuracy from its hyper-parameters.  Each round the proposer fits an ensemble
of them on every trial outcome so far, scores a pool of candidates drawn over
the search's own space and refined around the best of them, and proposes a
spread-out set of the best-scoring ones plus a share of plain draws.
Plain random search stays the fallback (too little history, or fits that
predict held-out trials worse than their mean) and the judge (the target
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
            "tm_mlp_reset": (ctypes.c_int, [vp, fp]),
            "tm_mlp_train": (ctypes.c_int, [vp, fp, fp, u32, u32, ctypes.c_float, fp]),
            "tm_mlp_train_many": (ctypes.c_int, [ctypes.POINTER(vp), u32, ctypes.POINTER(fp), ctypes.POINTER(fp),
                                                 ctypes.POINTER(u32), u32, ctypes.c_float, fp]),
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

    def reset(self, params=None, seed=0):
        """Start again from `params` (or a fresh He-normal init from `seed`)
        with zeroed Adam state, on the same device buffers."""
        p, pp = _f32(init_params(seed) if params is None else params)
        self._check(self._L.tm_mlp_reset(self._h, pp))

    def train(self, X, Y, steps, lr=3e-3):
        x, xp = _f32(X)
        y, yp = _f32(Y)
        loss = ctypes.c_float()
        self._check(self._L.tm_mlp_train(self._h, xp, yp, y.shape[0], steps, lr, ctypes.byref(loss)))
        return loss.value

    @staticmethod
    def train_many(nets, Xs, Ys, steps, lr=3e-3):
        """net.train(X, Y, steps, lr) for every (net, X, Y), the fits side by
        side on the networks' own streams; the networks must be distinct.
        Returns each network's last-step loss."""
        k = len(nets)
        if not (len(Xs) == len(Ys) == k):
            raise ValueError("train_many: %d networks, %d inputs, %d targets" % (k, len(Xs), len(Ys)))
        if k == 0:
            return []
        L = nets[0]._L
        xs = [_f32(X) for X in Xs]
        ys = [_f32(Y) for Y in Ys]
        for (x, _), (y, _) in zip(xs, ys):
            if x.ndim != 2 or x.shape[1] != INPUTS or y.shape != (x.shape[0],):
                raise ValueError("train_many: X must be n x %d and Y n" % INPUTS)
        fp = ctypes.POINTER(ctypes.c_float)
        hs = (ctypes.c_void_p * k)(*[n._h.value for n in nets])
        xp = (fp * k)(*[p for _, p in xs])
        yp = (fp * k)(*[p for _, p in ys])
        ns = (ctypes.c_uint32 * k)(*[y.shape[0] for y, _ in ys])
        loss = (ctypes.c_float * k)()
        nets[0]._check(L.tm_mlp_train_many(hs, k, xp, yp, ns, steps, lr, loss))
        return list(loss)

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
    t = params.get("target")
    return int(_effective_target(np.array([float(params["clauses"])]), np.array([-1.0 if t is None else float(t)]),
                                 np.array([float(params["clause_tolerance"])]))[0])


def _effective_target(clauses, target, tolerance):
    """effective_target over arrays of options (clauses, target, tolerance)."""
    c = np.ceil(clauses / 8.0) * 8
    t = np.where(np.isnan(target), -1.0, target)
    frac = np.ceil(np.maximum(1.0, (c * t).astype(np.float32).astype(np.float64)))
    out = np.where(t < 0, np.floor(np.sqrt(c / 2.0) * tolerance), np.where(t >= 1, np.ceil(t), frac))
    return out.astype(np.int64)


_FEATURE_OPTIONS = (("clauses", 8.0), ("specificity", 1.0), ("target", -1.0), ("clause_tolerance", 1.0),
                    ("clause_maximum", 1.0), ("include_bits", 1.0))


def _option_cols(params_list):
    """One float array per option over a list of settings (unset -> default)."""
    return {k: np.array([float(p.get(k, d) if p.get(k, d) is not None else d) for p in params_list])
            for k, d in _FEATURE_OPTIONS}


def features(params, ctx):
    """The 12 surrogate inputs: log clauses, log specificity, target/clauses,
    log target, log tolerance, log maximum, include_bits, epochs, feature
    density, log classes, log features, best dev accuracy so far."""
    return [float(v) for v in features_batch(_option_cols([params]), ctx)[0]]


def features_batch(cols, ctx):
    """features() of a pool of candidates, one array per option: an (n, 12)
    float32 matrix."""
    c = np.maximum(8.0, cols["clauses"])
    t = np.maximum(1, _effective_target(cols["clauses"], cols["target"], cols["clause_tolerance"])).astype(np.float64)
    n = c.shape[0]
    f = np.empty((n, INPUTS), np.float64)
    f[:, 0] = np.log2(c) / 12.0
    f[:, 1] = np.log(np.maximum(1e-3, cols["specificity"])) / 4.0
    f[:, 2] = t / c
    f[:, 3] = np.log2(t) / 12.0
    f[:, 4] = np.log2(np.maximum(1.0, cols["clause_tolerance"])) / 7.0
    f[:, 5] = np.log2(np.maximum(1.0, cols["clause_maximum"])) / 7.0
    f[:, 6] = cols["include_bits"] / 4.0
    f[:, 7] = float(ctx.get("epochs", 10)) / 20.0
    f[:, 8] = float(ctx.get("density", 0.5))
    f[:, 9] = math.log(max(2.0, float(ctx.get("classes", 2)))) / 5.0
    f[:, 10] = math.log2(max(2.0, float(ctx.get("features", 2)))) / 16.0
    f[:, 11] = float(ctx.get("best", 0.0))
    return f.astype(np.float32)


class Space:
    """The search space of a set of samplers as arrays: candidates are drawn,
    moved and decoded a whole pool at a time, with each range's bounds,
    log scale and rounding as its sampler applies them (optimize.lua:187-258)."""

    def __init__(self, samplers, names, rng):
        self.rng = rng
        self.ranges, self.lists, self.fixed = {}, {}, {}
        for name in names:
            s = samplers.get(name)
            if s is None:
                continue
            if s.type == "range" and getattr(s, "bounds", None):
                self.ranges[name] = (s, s.bounds)
            elif s.type == "list" and getattr(s, "values", None) and \
                    all(isinstance(v, (int, float)) and not isinstance(v, bool) for v in s.values):
                self.lists[name] = s.values
            else:
                # a fixed option, or one the pool cannot vary numerically:
                # held at one draw for the whole pool
                self.fixed[name] = s.sample()

    def _numeric_fixed(self):
        """The fixed options that enter the pool's arrays (others only
        reach the proposed options)."""
        return [(k, v) for k, v in self.fixed.items() if isinstance(v, (int, float)) and not isinstance(v, bool)]

    @staticmethod
    def _span(b):
        lo, hi = b["min"] + b["shift"], b["max"] + b["shift"]
        return (math.log(lo), math.log(hi) - math.log(lo)) if b["log"] else (b["min"], b["max"] - b["min"])

    def _decode(self, b, x):
        """Values (after the log map) to the sampler's rounded values."""
        x = np.minimum(np.maximum(x, b["min"]), b["max"])
        if b["pow2"]:
            x = 2.0 ** np.floor(np.log(x) / math.log(2) + 0.5)
        elif b["round"]:
            x = np.floor(x / b["round"] + 0.5) * b["round"]
        elif b["int"]:
            x = np.floor(x + 0.5)
        return np.minimum(np.maximum(x, b["min"]), b["max"])

    def _from_unit(self, b, u):
        base, span = self._span(b)
        x = u * span + base
        return self._decode(b, np.exp(x) - b["shift"] if b["log"] else x)

    def _to_unit(self, b, v):
        base, span = self._span(b)
        x = np.log(v + b["shift"]) if b["log"] else v
        return (x - base) / span if span > 0 else np.zeros_like(v)

    def draw(self, n, centred):
        """n candidates: uniform over each range (as a sampler without a
        centre), or (centred) jittered around each sampler's centre with the
        sampler's current jitter and reflection."""
        cols = {}
        for name, (s, b) in self.ranges.items():
            if centred and s.center is not None:
                c = math.log(s.center + b["shift"]) if b["log"] else float(s.center)
                x = self.rng.normal(c, s.state["jitter"], n)
                if b["log"]:
                    x = np.exp(x) - b["shift"]
                x = np.where(x > b["max"], 2 * b["max"] - x, x)
                x = np.where(x < b["min"], 2 * b["min"] - x, x)
                cols[name] = self._decode(b, x)
            else:
                cols[name] = self._from_unit(b, self.rng.random(n))
        for name, vals in self.lists.items():
            cols[name] = np.array(vals, np.float64)[self.rng.integers(len(vals), size=n)]
        for name, v in self._numeric_fixed():
            cols[name] = np.full(n, float(v))
        return cols

    def move(self, cols, idx, per, scales=(0.03, 0.1, 0.25), snap=0.15, redraw=0.25):
        """`per` local moves from each candidate in idx: every range coordinate
        moves by a normal step (one of `scales`, in units of the range) and
        reflects at the bounds, or, with probability `snap`, goes to the
        nearer bound; a list option is drawn again with probability `redraw`."""
        src = np.repeat(np.asarray(idx), per)
        m = src.shape[0]
        out = {}
        sc = np.array(scales)[self.rng.integers(len(scales), size=m)]
        for name, (s, b) in self.ranges.items():
            u = self._to_unit(b, cols[name][src])
            u = u + sc * self.rng.standard_normal(m)
            u = np.where(u > 1, 2 - u, u)
            u = np.where(u < 0, -u, u)
            u = np.clip(u, 0.0, 1.0)
            snapped = self.rng.random(m) < snap
            u = np.where(snapped, np.round(u), u)
            out[name] = self._from_unit(b, u)
        for name, vals in self.lists.items():
            again = self.rng.random(m) < redraw
            out[name] = np.where(again, np.array(vals, np.float64)[self.rng.integers(len(vals), size=m)],
                                 cols[name][src])
        for name, v in self._numeric_fixed():
            out[name] = np.full(m, float(v))
        return out

    def unit(self, cols):
        """Each candidate's coordinates in the unit box (log ranges on their
        log scale; a list option as its index over the list's length)."""
        parts = [self._to_unit(b, cols[name]) for name, (s, b) in self.ranges.items()]
        for name, vals in self.lists.items():
            v = np.array(vals, np.float64)
            parts.append(np.argmin(np.abs(cols[name][:, None] - v[None, :]), 1) / max(1, len(vals) - 1))
        n = next(iter(cols.values())).shape[0]
        return np.stack(parts, 1) if parts else np.zeros((n, 0))

    def params(self, cols, i):
        """Candidate i as an options dict (integral ranges as ints)."""
        p = {}
        for name, (s, b) in self.ranges.items():
            v = float(cols[name][i])
            p[name] = int(v) if (b["int"] or b["round"]) and float(v).is_integer() else v
        for name in self.lists:
            v = float(cols[name][i])
            p[name] = next((x for x in self.lists[name] if float(x) == v), v)
        for name, v in self.fixed.items():
            p[name] = v
        return p


def _swapped(cols):
    """optimize._swap_tolerance over a pool: tolerance and maximum swapped
    where the tolerance is the larger."""
    out = dict(cols)
    if "clause_tolerance" in cols and "clause_maximum" in cols:
        a, b = cols["clause_tolerance"], cols["clause_maximum"]
        sw = (a > 0) & (b > 0) & (a > b)
        out["clause_tolerance"], out["clause_maximum"] = np.where(sw, b, a), np.where(sw, a, b)
    return out


def _with_defaults(cols, n):
    full = dict(cols)
    for k, d in (("clauses", 8.0), ("specificity", 1.0), ("target", -1.0), ("clause_tolerance", 1.0),
                 ("clause_maximum", 1.0), ("include_bits", 1.0)):
        if k not in full:
            full[k] = np.full(n, d)
    return full


def _concat(a, b):
    return {k: np.concatenate([a[k], b[k]]) for k in a}


class Proposer:
    """propose(samplers, n, history) for optimize.search_batched.

    The first round is a share `initial` of draws spread uniformly over the
    whole space, the rest plain draws, so the first fit sees more than the
    search's starting neighbourhood.  From `min_history` trials on,
    `ensemble` networks are fitted to every trial so far (the first on all of
    them, the others on bootstrap resamples), and candidates are scored by the
    ensemble's mean prediction less `kappa` times its spread, so settings where
    the fits disagree (far from the trials seen) are discounted.  The pool:
    half drawn around the search's sampler centres, half uniform over the
    space; then `refine` times, `pool` local moves from the `refine_top`
    best-scoring candidates (a share `snap` of the coordinates pushed onto the
    nearer bound, where the best settings of a space often lie) are scored and
    added.  The network's picks are the best-scoring candidates not tried
    before: a share `exploit` of them in score order, the rest keeping a
    distance `spread` (unit box) from each other; a share `explore` of the
    round stays plain draws.  With `anneal`, the spacing shrinks with the
    search's own jitter.  (A trial's score is one noisy draw: trials of
    nearly the same setting differ by several points, and the search keeps
    the best one.  `exploit` repeats trials near the predicted optimum for
    that reason.)"""

    def __init__(self, ctx, param_names=opt.PARAM_NAMES, seed=0, pool=4096, explore=0.25, min_history=48,
                 steps=300, lr=3e-3, key=opt.make_key, refine=2, refine_top=64, ensemble=4, kappa=0.5, snap=0.15,
                 spread=0.2, initial=0.5, folds=5, anneal=False, exploit=0.0):
        self.ctx, self.names, self.pool, self.explore = dict(ctx), param_names, pool, explore
        self.min_history, self.steps, self.lr, self.key = min_history, steps, lr, key
        self.refine, self.refine_top = refine, refine_top
        self.ensemble, self.kappa, self.snap, self.spread = max(1, ensemble), kappa, snap, spread
        self.initial, self.folds = initial, max(2, folds)
        self.anneal, self.exploit = anneal, exploit
        self.rng = np.random.default_rng(seed)
        self.seed = seed
        self.last = {"mode": "plain"}
        self._nets = []

    def _net(self, i, seed):
        """Network i of the proposer's pool, started from a fresh init."""
        while len(self._nets) <= i:
            self._nets.append(MLP(seed=seed))
        self._nets[i].reset(seed=seed)
        return self._nets[i]

    def close(self):
        for net in self._nets:
            net.destroy()
        self._nets = []

    def _fit(self, history):
        best = max(s for _, s in history)
        ctx = dict(self.ctx, best=best)
        X = features_batch(_option_cols([p for p, _ in history]), ctx)
        y = np.array([s for _, s in history], np.float32)
        mu, sd = float(y.mean()), float(y.std() + 1e-6)
        Y = (y - mu) / sd
        # held-out check: out-of-fold predictions of `folds` fits, each
        # against the mean of its own training part
        idx = self.rng.permutation(len(y))
        folds = []
        for f in range(self.folds):
            test = idx[f::self.folds]
            train = np.setdiff1d(idx, test)
            if test.size and train.size:
                folds.append((f, train, test))
        sels = [np.arange(len(y)) if e == 0 else self.rng.integers(len(y), size=len(y))
                for e in range(self.ensemble)]
        # every fit at once: the fold fits first, then the ensemble members
        fold_nets = [self._net(i, self.seed + len(y) + f) for i, (f, _, _) in enumerate(folds)]
        nets = [self._net(len(folds) + e, self.seed + len(y) + 7919 * e) for e in range(self.ensemble)]
        losses = MLP.train_many(fold_nets + nets, [X[tr] for _, tr, _ in folds] + [X[sel] for sel in sels],
                                [Y[tr] for _, tr, _ in folds] + [Y[sel] for sel in sels], self.steps, self.lr)
        resid = total = 0.0
        for net, (_, train, test) in zip(fold_nets, folds):
            resid += float(((net.predict(X[test]) - Y[test]) ** 2).sum())
            total += float(((Y[test] - Y[train].mean()) ** 2).sum())
        r2 = 1.0 - resid / (total + 1e-9)
        return nets, ctx, r2, losses[len(folds)]

    def __call__(self, samplers, n, history):
        plain = [opt.sample_params(samplers, self.names) for _ in range(n)]
        if len(history) < self.min_history:
            self.last = {"mode": "plain", "reason": "history %d < %d" % (len(history), self.min_history)}
            if not history and self.initial > 0:
                # the first round: a share `initial` drawn uniformly over the
                # whole space, so the first fit sees more than the search's
                # starting neighbourhood
                space = Space(samplers, self.names, self.rng)
                k = int(round(self.initial * n))
                cols = space.draw(k, False)
                self.last["initial"] = k
                return [space.params(cols, i) for i in range(k)] + plain[: n - k]
            return plain
        nets, ctx, r2, loss = self._fit(history)
        if not r2 > 0.0:
            self.last = {"mode": "plain", "reason": "held-out R2 %.3f" % r2}
            return plain
        space = Space(samplers, self.names, self.rng)

        def score(cols):
            m = next(iter(cols.values())).shape[0]
            F = features_batch(_with_defaults(_swapped(cols), m), ctx)
            P = np.stack([net.predict(F) for net in nets])
            return P.mean(0) - (self.kappa * P.std(0) if len(nets) > 1 else 0.0)

        half = self.pool // 2
        cols = _concat(space.draw(half, True), space.draw(self.pool - half, False))
        sc = score(cols)
        for _ in range(self.refine):
            top = np.argsort(-sc)[: self.refine_top]
            moved = space.move(cols, top, max(1, self.pool // max(1, len(top))), snap=self.snap)
            cols = _concat(cols, moved)
            sc = np.concatenate([sc, score(moved)])
        seen = {self.key(p) for p, _ in history}
        n_model = n - int(round(self.explore * n))
        n_exploit = int(round(self.exploit * n_model))
        spread = self.spread
        if self.anneal:
            # the search's own step size (its jitter, narrowed by the 1/5
            # success rule) relative to its start scales the picks' spacing
            for smp in samplers.values():
                st = getattr(smp, "state", None)
                if st and st.get("jitter0"):
                    spread *= st["jitter"] / st["jitter0"]
                    break
        picks = []
        U = space.unit(cols) if spread > 0 else None
        # every candidate's distance to the nearest spaced pick so far
        near = np.full(sc.shape[0], np.inf) if U is not None else None
        for i in np.argsort(-sc, kind="stable"):
            # past the first `n_exploit` picks, a pick keeps at least `spread`
            # (Euclidean, unit box) from the round's other spaced picks, so
            # the round covers more than one optimum
            spaced = len(picks) >= n_exploit
            if spaced and near is not None and near[i] < spread:
                continue
            p = space.params(cols, i)
            k = self.key(opt._swap_tolerance(dict(p)))
            if k in seen:
                continue
            seen.add(k)
            picks.append(p)
            if spaced and near is not None:
                np.minimum(near, np.sqrt(((U - U[i]) ** 2).sum(1)), out=near)
            if len(picks) >= n_model:
                break
        n_picked = len(picks)
        picks += plain[: n - len(picks)]
        # the first n_model proposals are the network's picks, the rest plain draws
        self.last = {"mode": "surrogate", "r2": float(r2), "loss": float(loss), "n_model": n_picked,
                     "pool": int(sc.shape[0]), "spread": spread}
        return picks
