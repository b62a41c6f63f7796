"""santoku.tsetlin.optimize -- hyper-parameter search around the engine.

Restates the TM half of lib/santoku/tsetlin/optimize.lua:
  build_sampler / sample_params            optimize.lua:187-288
  recenter / adapt_jitter / success_factor optimize.lua:359-389
  search (adaptive random search)          optimize.lua:391-516
  train_tm (early stopping + checkpoint)   optimize.lua:586-653
  optimize_tm / classifier / encoder       optimize.lua:657-844
The landmark/threshold branch of optimize_tm (hlth, itq; optimize.lua:697-735)
belongs to the spectral pipeline and is not part of this engine.

Random numbers come from a numpy Generator (`rng`), standing in for
santoku.random / santoku.num, whose sources are not available; the sequence
differs, the procedure is the same.  `search_batched` evaluates all the trials
of a round with one call (the batched trial engine), which the Lua driver runs
one by one.
"""
import math

import numpy as np

from . import evaluator  # noqa: F401  (re-exported for callers of the search)


def _round_to_pow2(x):
    return 2.0 ** math.floor(math.log(x) / math.log(2) + 0.5)


class Sampler:
    """One hyper-parameter: fixed, range (optionally log/int/pow2/round) or list."""

    def __init__(self, kind, center=None, sample=None, adapt=None):
        self.type = kind
        self.center = center
        self._sample = sample
        self._adapt = adapt

    def sample(self, center=None):
        return self._sample(center)

    def adapt(self, factor):
        if self._adapt:
            self._adapt(factor)


def build_sampler(spec, global_dev=None, rng=None):
    """optimize.lua:187-258."""
    rng = rng or np.random.default_rng()
    if spec is None:
        return None
    if isinstance(spec, (int, float, bool, str)):
        return Sampler("fixed", spec, lambda c=None: spec)
    if isinstance(spec, dict) and (spec.get("def") is not None or ("min" in spec and "max" in spec)):
        minv, maxv = spec.get("min"), spec.get("max")
        if minv is None or maxv is None:
            raise ValueError("range spec missing min|max")
        is_log, is_int, is_pow2 = bool(spec.get("log")), bool(spec.get("int")), bool(spec.get("pow2"))
        round_to = spec.get("round")
        shift = (1 - minv) if (is_log and minv <= 0) else 0
        smin, smax = minv + shift, maxv + shift
        span = (math.log(smax) - math.log(smin)) if is_log else (maxv - minv)
        j0 = (spec.get("dev") or global_dev or 1.0) * span
        state = {"jitter": j0, "jitter0": j0}

        def sample(center=None):
            if center is not None:
                c = math.log(center + shift) if is_log else center
                x = rng.normal(c, state["jitter"])        # fast_normal(c, jitter^2): variance jitter^2
                if is_log:
                    x = math.exp(x) - shift
            else:
                if is_log:
                    x = math.exp(rng.random() * span + math.log(smin)) - shift
                else:
                    x = rng.random() * span + minv
            if x > maxv:
                x = 2 * maxv - x
            if x < minv:
                x = 2 * minv - x
            x = min(max(x, minv), maxv)
            if is_pow2:
                x = _round_to_pow2(x)
            elif round_to:
                x = math.floor(x / round_to + 0.5) * round_to
            elif is_int:
                x = math.floor(x + 0.5)
            return min(max(x, minv), maxv)

        def adapt(factor):
            state["jitter"] *= factor

        s = Sampler("range", spec.get("def"), sample, adapt)
        s.state = state
        # the range's definition, for samplers of whole candidate pools
        s.bounds = {"min": minv, "max": maxv, "log": is_log, "int": is_int, "pow2": is_pow2, "round": round_to,
                    "shift": shift}
        return s
    if isinstance(spec, (list, tuple)) and len(spec) > 0:
        s = Sampler("list", spec[0], lambda c=None: spec[int(rng.integers(len(spec)))])
        s.values = list(spec)
        return s
    raise ValueError("Bad hyper-parameter specification: %r" % (spec,))


def build_samplers(args, param_names, global_dev=None, rng=None):
    return {p: build_sampler(args.get(p), global_dev, rng) for p in param_names}


def sample_params(samplers, param_names, base_cfg=None, use_exact_defaults=False):
    """optimize.lua:268-288."""
    p = dict(base_cfg or {})
    for name in param_names:
        s = samplers.get(name)
        if s is None:
            continue
        if use_exact_defaults and s.center is not None:
            p[name] = s.center
        elif s.type == "range":
            p[name] = s.sample(s.center)
        else:
            p[name] = s.sample()
    return p


def all_fixed(samplers):
    return all(s is None or s.type == "fixed" for s in samplers.values())


def recenter(samplers, param_names, best_params):
    for p in param_names:
        s = samplers.get(p)
        if s is not None and s.type == "range" and best_params.get(p) is not None:
            s.center = best_params[p]


def adapt_jitter(samplers, factor):
    for s in samplers.values():
        if s is not None:
            s.adapt(factor)


def success_factor(success_rate):
    """1/5th success rule (optimize.lua:381-389)."""
    if success_rate > 0.2:
        return 1.0
    if success_rate > 0.05:
        return 0.8
    return 0.5


def _round_end(samplers, param_names, best_params, improvements, samples):
    if best_params is not None:
        recenter(samplers, param_names, best_params)
    rate = improvements / samples if samples > 0 else 0.0
    factor = success_factor(rate)
    adapt_jitter(samplers, factor)
    return rate, factor


def search(param_names, samplers, trial_fn, rounds=3, trials=10, tolerance=1e-8, make_key=None, each=None,
           cleanup=None, skip_final=False):
    """Adaptive random search (optimize.lua:391-516).  trial_fn(params, info)
    returns (score, metrics, result)."""
    best_score, best_params, best_result, best_metrics = -math.inf, None, None, None
    if all_fixed(samplers) or rounds <= 0 or trials <= 0:
        best_params = sample_params(samplers, param_names, None, True)
        if skip_final:
            return None, best_params, None
        _, metrics, result = trial_fn(best_params, {"is_final": True})
        return result, best_params, metrics
    for r in range(1, rounds + 1):
        seen = set()
        round_best_score, round_best_params, improvements, samples = -math.inf, None, 0, 0
        for t in range(1, trials + 1):
            params = sample_params(samplers, param_names)
            if make_key is not None:
                key = make_key(params)
                if key in seen:
                    continue
                seen.add(key)
            score, metrics, result = trial_fn(params, {"round": r, "trial": t, "is_final": False})
            samples += 1
            if each:
                each({"event": "trial", "round": r, "trial": t, "params": params, "score": score, "metrics": metrics})
            if score > round_best_score:
                round_best_score, round_best_params = score, params
            if score > best_score + tolerance:
                improvements += 1
                if best_result is not None and cleanup:
                    cleanup(best_result)
                best_score, best_params, best_result, best_metrics = score, params, result, metrics
            elif result is not None and cleanup:
                cleanup(result)
        rate, factor = _round_end(samplers, param_names, best_params, improvements, samples)
        if each:
            each({"event": "round", "round": r, "round_best_score": round_best_score,
                  "round_best_params": round_best_params, "global_best_score": best_score,
                  "global_best_params": best_params, "success_rate": rate, "adapt_factor": factor})
    if not skip_final and best_params is not None:
        if best_result is not None and cleanup:
            cleanup(best_result)
        _, best_metrics, best_result = trial_fn(best_params, {"is_final": True})
    return best_result, best_params, best_metrics


def search_batched(param_names, samplers, batch_fn, rounds=3, trials=10, tolerance=1e-8, make_key=None, each=None,
                   propose=None):
    """search() with every trial of a round evaluated by one batch_fn(list of
    params) -> list of scores.  propose(samplers, n, history) may replace the
    sampler draws for a round (the learned surrogate); history holds
    (params, score) of every evaluated trial.  Returns (best_params,
    best_score, history)."""
    best_score, best_params = -math.inf, None
    history = []
    for r in range(1, rounds + 1):
        if propose is not None:
            cand = propose(samplers, trials, history)
        else:
            cand = [sample_params(samplers, param_names) for _ in range(trials)]
        seen, batch = set(), []
        for params in cand:
            if make_key is not None:
                key = make_key(params)
                if key in seen:
                    continue
                seen.add(key)
            batch.append(params)
        scores = batch_fn(batch, {"round": r}) if batch else []
        improvements = 0
        for params, score in zip(batch, scores):
            history.append((params, score))
            if each:
                each({"event": "trial", "round": r, "params": params, "score": score})
            if score > best_score + tolerance:
                improvements += 1
                best_score, best_params = score, params
        rate, factor = _round_end(samplers, param_names, best_params, improvements, len(batch))
        if each:
            each({"event": "round", "round": r, "global_best_score": best_score, "global_best_params": best_params,
                  "success_rate": rate, "adapt_factor": factor})
    return best_params, best_score, history


def train_tm(typ, machine, args, params, iterations, early_patience, tolerance, metric_fn, each_cb, info):
    """Train with a per-epoch metric, patience early stop and restore of the
    best epoch's actions (optimize.lua:586-653).  Returns (score, metrics)."""
    st = {"best": -math.inf, "last": -math.inf, "metrics": None, "since": 0, "ckpt": None}

    def on_epoch(epoch):
        res = metric_fn(machine)
        score, metrics = (res if isinstance(res, tuple) else (res, None))
        st["last"], st["metrics"] = score, metrics
        if each_cb is not None:
            if each_cb(machine, info.get("is_final"), metrics, params, epoch, info.get("round"), info.get("trial")) is False:
                return False
        if score > st["best"] + tolerance:
            st["best"], st["since"] = score, 0
            if early_patience:
                st["ckpt"] = machine.checkpoint()
        else:
            st["since"] += 1
        if early_patience and st["since"] >= early_patience:
            return False
        return True

    if typ == "encoder":
        machine.train(sentences=args["sentences"], codes=args["codes"], samples=args["samples"],
                      iterations=iterations, each=on_epoch)
    else:
        machine.train(samples=args["samples"], problems=args["problems"], solutions=args["solutions"],
                      iterations=iterations, each=on_epoch)
    if st["ckpt"] is not None:
        machine.restore(st["ckpt"])
        res = metric_fn(machine)
        st["last"], st["metrics"] = (res if isinstance(res, tuple) else (res, None))
    return st["last"], st["metrics"]


PARAM_NAMES = ("clauses", "clause_tolerance", "clause_maximum", "target", "specificity", "include_bits")


def make_key(p):
    """Trial dedup key (optimize.lua:781-793)."""
    return "%d|%d|%d|%d|%d|%d" % (p["clauses"], p["clause_tolerance"], p["clause_maximum"],
                                  math.floor(p["target"] * 1e6 + 0.5), math.floor(p["specificity"] * 1e6 + 0.5),
                                  p.get("include_bits") or 1)


def _swap_tolerance(params):
    if params.get("clause_tolerance") and params.get("clause_maximum") and \
            params["clause_tolerance"] > params["clause_maximum"]:
        params["clause_tolerance"], params["clause_maximum"] = params["clause_maximum"], params["clause_tolerance"]
    return params


def optimize_tm(args, typ, rng=None):
    """optimize.lua:657-836: search on one reusable machine (reconfigure per
    trial), then train a fresh machine with the best parameters."""
    from . import create
    rng = rng or np.random.default_rng(args.get("seed"))
    patience = args.get("search_patience", 10)
    final_patience = args.get("final_patience", 40)
    tolerance = args.get("search_tolerance", 1e-8)
    iters_search = args.get("search_iterations", 10)
    final_iters = args.get("final_iterations") or iters_search * 10
    global_dev = args.get("search_dev", 0.2)
    metric_fn = args["search_metric"]
    each_cb = args.get("each")
    samplers = build_samplers(args, PARAM_NAMES, global_dev, rng)
    base = {"classes": args.get("classes"), "features": args.get("features"), "negative": args.get("negative"),
            "hidden": args.get("hidden"), "visible": args.get("visible"), "seed": args.get("seed", 0)}
    base = {k: v for k, v in base.items() if v is not None}
    search_tm = None
    if not all_fixed(samplers):
        search_tm = create(typ, dict(base, clauses=8, clause_tolerance=8, clause_maximum=8, target=4,
                                     specificity=1000, include_bits=1, reusable=True))

    def trial_fn(params, info):
        _swap_tolerance(params)
        search_tm.reconfigure(dict(params, seed=args.get("seed", 0)))
        score, metrics = train_tm(typ, search_tm, args, params, iters_search, patience if patience > 0 else None,
                                  tolerance, metric_fn, each_cb, info)
        return score, metrics, None

    _, best_params, _ = search(PARAM_NAMES, samplers, trial_fn, rounds=args.get("search_rounds", 3),
                               trials=args.get("search_trials", 10), tolerance=tolerance, skip_final=True,
                               make_key=make_key)
    if search_tm is not None:
        search_tm.destroy()
    final_tm = create(typ, dict(base, **{k: best_params[k] for k in PARAM_NAMES if best_params.get(k) is not None}))
    _, final_metrics = train_tm(typ, final_tm, args, best_params, final_iters,
                                final_patience if final_patience > 0 else None, tolerance, metric_fn, each_cb,
                                {"is_final": True})
    return final_tm, final_metrics, best_params


def classifier(args):
    return optimize_tm(args, "classifier")


def encoder(args):
    return optimize_tm(args, "encoder")
