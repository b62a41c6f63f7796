"""Surrogate MLP kernels against the oracle, and what the trained network gains."""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as MO
from santoku.tsetlin import optimize as opt
from santoku.tsetlin import surrogate as S

pytestmark = pytest.mark.gpu


def data(seed, n=200):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 12)).astype(np.float32)
    Y = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    return X, Y


def test_gradient_matches_oracle():
    X, Y = data(1)
    p = S.init_params(5)
    net = S.MLP(p)
    g, loss = net.gradient(X, Y)
    gref = MO.gradient(p, X, Y)
    assert loss == pytest.approx(MO.loss(p, X, Y), rel=1e-4)
    # FP32 sums over 200 samples against the float64 restatement
    assert np.allclose(g, gref, rtol=2e-3, atol=2e-5), np.abs(g - gref).max()


def test_predict_matches_oracle():
    X, _ = data(2, 1000)
    p = S.init_params(6)
    net = S.MLP(p)
    out = net.predict(X)
    ref, _ = MO.forward(p, X)
    assert np.allclose(out, ref, rtol=1e-4, atol=1e-5)


def test_adam_steps_match_oracle_and_reduce_loss():
    X, Y = data(3)
    p = S.init_params(7)
    net = S.MLP(p)
    net.train(X, Y, 5, 1e-3)
    pref, _, _ = MO.adam(p, X, Y, 5, 1e-3)
    assert np.allclose(net.params(), pref, rtol=1e-3, atol=2e-5), np.abs(net.params() - pref).max()
    l0 = MO.loss(p, X, Y)
    net.train(X, Y, 400, 3e-3)
    assert MO.loss(net.params(), X, Y) < 0.2 * l0


def test_train_many_equals_separate_fits():
    """tm_mlp_train_many: networks fitted side by side on their own streams,
    each on its own sample count, end bit-equal to the same fits one after
    another; a network listed twice is refused."""
    sets = [data(20 + i, n) for i, n in enumerate((300, 77, 448, 32, 1))]
    inits = [S.init_params(40 + i) for i in range(len(sets))]
    one = []
    for p, (X, Y) in zip(inits, sets):
        net = S.MLP(p)
        loss = net.train(X, Y, 25, 2e-3)
        one.append((net.params(), loss))
        net.destroy()
    nets = [S.MLP(p) for p in inits]
    losses = S.MLP.train_many(nets, [X for X, _ in sets], [Y for _, Y in sets], 25, 2e-3)
    for net, (p1, l1), l in zip(nets, one, losses):
        assert np.array_equal(net.params(), p1)
        assert l == l1
    # a second call continues every network's Adam step count
    S.MLP.train_many(nets[:2], [sets[0][0], sets[1][0]], [sets[0][1], sets[1][1]], 5, 2e-3)
    ref = S.MLP(inits[0])
    ref.train(*sets[0], 25, 2e-3)
    ref.train(*sets[0], 5, 2e-3)
    assert np.array_equal(nets[0].params(), ref.params())
    with pytest.raises(S.TsetlinError):
        S.MLP.train_many([nets[0], nets[1], nets[0]], [sets[0][0]] * 3, [sets[0][1]] * 3, 1, 1e-3)
    for net in nets + [ref]:
        net.destroy()
    # more networks than one launch takes (16): each on its own stream
    many = [S.MLP(S.init_params(60 + i)) for i in range(17)]
    X, Y = data(30, 90)
    S.MLP.train_many(many, [X] * 17, [Y] * 17, 6, 2e-3)
    for i, net in enumerate(many):
        ref = S.MLP(S.init_params(60 + i))
        ref.train(X, Y, 6, 2e-3)
        assert np.array_equal(net.params(), ref.params()), i
        ref.destroy()
        net.destroy()


def test_reset_equals_a_fresh_network():
    """tm_mlp_reset: parameters, Adam moments and step count start again, so
    training after a reset equals training a new network (the slice sums
    are in a fixed order, so equal means bit-equal)."""
    X, Y = data(4, 300)
    p = S.init_params(9)
    a = S.MLP(p)
    a.train(X, Y, 20, 1e-3)
    a.reset(p)
    a.train(X, Y, 20, 1e-3)
    b = S.MLP(p)
    b.train(X, Y, 20, 1e-3)
    assert np.array_equal(a.params(), b.params())


SPACE = {"clauses": {"def": 512, "min": 64, "max": 4096, "log": True, "round": 8},
         "clause_tolerance": {"def": 8, "min": 1, "max": 64, "log": True, "int": True},
         "clause_maximum": {"def": 16, "min": 1, "max": 64, "log": True, "int": True},
         "target": [0.125, 0.25],
         "specificity": {"def": 10, "min": 2, "max": 40},
         "include_bits": 1}


def objective(p, rng):
    """A smooth stand-in for dev accuracy with a single optimum."""
    lc = math.log2(p["clauses"])
    ls = math.log(p["specificity"])
    lt = math.log2(max(1, min(p["clause_tolerance"], p["clause_maximum"])))
    v = 0.92 - 0.06 * ((lc - 10) / 2) ** 2 - 0.08 * ((ls - math.log(7)) / 0.6) ** 2 - 0.05 * ((lt - 2) / 1.5) ** 2
    v += 0.02 * (p["target"] == 0.25)
    return max(0.0, v + rng.normal(0, 0.004))


def run(seed, surrogate):
    rng = np.random.default_rng(seed)
    samplers = opt.build_samplers(SPACE, opt.PARAM_NAMES, 0.3, rng)
    noise = np.random.default_rng(seed + 100)
    fn = lambda batch, info: [objective(p, noise) for p in batch]
    prop = S.Proposer({"epochs": 10, "density": 0.5, "classes": 4, "features": 256}, seed=seed, pool=2048,
                      min_history=24) if surrogate else None
    _, best, hist = opt.search_batched(opt.PARAM_NAMES, samplers, fn, rounds=6, trials=32, propose=prop)
    late = [s for _, s in hist[96:]]
    return best, float(np.mean(late))


def test_surrogate_gains_over_plain_search():
    plain = [run(s, False) for s in range(3)]
    sur = [run(s, True) for s in range(3)]
    # trials proposed in rounds 4-6 score higher on average, and the best found is no worse
    assert np.mean([m for _, m in sur]) > np.mean([m for _, m in plain]) + 0.01
    assert np.mean([b for b, _ in sur]) >= np.mean([b for b, _ in plain]) - 0.002


# ---------------------------------------------------------------------------
# workload 5: the surrogate fitted on outcomes the engine generates (batched
# trials on workload 1's data, scored by the reference's search metric), against
# the plain search (optimize.lua:391-516) over the workload-4 space

def _engine_search(seed, surrogate, dev, rounds=5, trials=32, epochs=4):
    """A search over the workload-4 space scored by the engine's batched
    trials; returns the best score, and per surrogate round the mean scores of
    the network's picks and of the plain draws proposed beside them."""
    from santoku.tsetlin import trials as TR
    from test_batch_gpu import BASE, C4_SPACE
    d = dev
    fn = TR.batch_trial_fn(BASE, d["rows"], d["ys"], d["drows"], d["dys"], iterations=epochs, patience=2,
                           seed=seed * 1000)
    samplers = opt.build_samplers(C4_SPACE, opt.PARAM_NAMES, 0.3, np.random.default_rng(seed))
    prop = S.Proposer({"epochs": epochs, "density": 0.5, "classes": 4, "features": 256}, seed=seed,
                      min_history=48) if surrogate else None
    rounds_out, cur = [], []

    def each(ev):
        if ev["event"] == "trial":
            cur.append(ev["score"])
        elif ev["event"] == "round":
            last = prop.last if prop is not None else {"mode": "plain"}
            if last.get("mode") == "surrogate":
                k = last["n_model"]
                rounds_out.append((float(np.mean(cur[:k])), float(np.mean(cur[k:])) if cur[k:] else None))
            else:
                rounds_out.append(None)
            cur.clear()

    _, best, hist = opt.search_batched(opt.PARAM_NAMES, samplers, fn, rounds=rounds, trials=trials, propose=prop,
                                       make_key=opt.make_key, each=each)
    return best, rounds_out


@pytest.fixture(scope="module")
def workload1_dev():
    import torch
    from santoku.tsetlin import data as D
    d = D.xor_and_task(21)
    return {"rows": torch.from_numpy(d["train_rows"]).cuda(), "ys": torch.from_numpy(d["train_y"]).cuda(),
            "drows": torch.from_numpy(d["test_rows"]).cuda(), "dys": torch.from_numpy(d["test_y"]).cuda()}


def test_surrogate_on_engine_outcomes_gains_over_plain(workload1_dev):
    """The network, fitted on the engine's own trial outcomes, picks trials
    that score higher than the plain draws proposed beside them in the same
    rounds (same sampler state, same data), and the surrogate-guided search
    finds a better best than the plain search from the same seed."""
    plain = [_engine_search(s, False, workload1_dev) for s in range(3)]
    sur = [_engine_search(s, True, workload1_dev) for s in range(3)]
    gains = [m - p for _, rs in sur for r in rs if r is not None and r[1] is not None for m, p in [r]]
    print("plain best:", [round(b, 4) for b, _ in plain])
    print("surrogate best:", [round(b, 4) for b, _ in sur])
    print("surrogate rounds (picks, plain beside):", [rs for _, rs in sur])
    print("per-round gain of the picks: mean %.4f over %d rounds" % (np.mean(gains), len(gains)))
    # rounds 3-5 of 3 seeds, minus any that fell back to plain search
    assert len(gains) >= 6, sur
    assert np.mean(gains) > 0.01, gains
    assert np.mean([b for b, _ in sur]) > np.mean([b for b, _ in plain]) + 0.005, (plain, sur)
