"""Evaluator metrics and the search driver (host logic, CPU)."""
import math

import numpy as np
import pytest

from santoku.tsetlin import evaluator as ev
from santoku.tsetlin import optimize as opt


def test_class_accuracy_hand_table():
    # 3 classes; pairs (pred, true)
    pred = [0, 0, 1, 2, 2, 1, 0, 5]
    true = [0, 1, 1, 2, 0, 1, 0, 1]      # last sample ignored: class 5 out of range
    r = ev.class_accuracy(pred, true, 8, 3)
    # class 0: tp 2 (s0, s6), fp 1 (s1), fn 1 (s4) -> p 2/3, r 2/3
    # class 1: tp 2 (s2, s5), fp 0, fn 1 (s1) -> p 1, r 2/3
    # class 2: tp 1 (s3), fp 1 (s4), fn 0 -> p 1/2, r 1
    f = lambda p, q: 2 * p * q / (p + q)
    assert r["classes"][0]["precision"] == pytest.approx(2 / 3)
    assert r["classes"][1]["recall"] == pytest.approx(2 / 3)
    assert r["classes"][2]["f1"] == pytest.approx(f(0.5, 1.0))
    assert r["f1"] == pytest.approx((f(2 / 3, 2 / 3) + f(1, 2 / 3) + f(0.5, 1)) / 3)
    assert r["precision"] == pytest.approx((2 / 3 + 1 + 0.5) / 3)


def test_class_accuracy_empty_class_scores_zero():
    r = ev.class_accuracy([0, 0], [0, 0], 2, 3)
    assert r["classes"][1] == {"precision": 0.0, "recall": 0.0, "f1": 0.0}
    assert r["f1"] == pytest.approx(1 / 3)


def test_encoding_accuracy_hand_codes():
    # 4 samples x 10 bits (2 bytes)
    exp = np.zeros((4, 2), np.uint8)
    pred = exp.copy()
    pred[0, 0] = 0b1       # bit 0 wrong in sample 0
    pred[1, 0] = 0b1       # bit 0 wrong in sample 1
    pred[2, 1] = 0b10      # bit 9 wrong in sample 2
    pred[3, 1] = 0b100     # bit 10: padding, not counted
    r = ev.encoding_accuracy(pred, exp, 4, 10)
    assert r["bits"][0] == 0.5 and r["bits"][9] == 0.25 and sum(r["bits"]) == 0.75
    assert r["mean_hamming"] == pytest.approx(1 - 3 / 40)
    assert r["ber_min"] == 1.0 and r["ber_max"] == 0.5
    ber = np.array(r["bits"])
    assert r["ber_std"] == pytest.approx(math.sqrt(((ber - 3 / 40) ** 2).sum() / 9))


def test_range_sampler_reflects_rounds_and_clamps():
    rng = np.random.default_rng(0)
    s = opt.build_sampler({"def": 16, "min": 8, "max": 256, "round": 8}, 0.2, rng)
    xs = [s.sample(s.center) for _ in range(2000)]
    assert all(8 <= x <= 256 and x % 8 == 0 for x in xs)
    s = opt.build_sampler({"min": 2, "max": 40}, 5.0, rng)      # huge jitter: reflection keeps bounds
    xs = [s.sample(39.0) for _ in range(2000)]
    assert all(2 <= x <= 40 for x in xs)
    s = opt.build_sampler({"min": 64, "max": 4096, "log": True, "pow2": True}, 0.2, rng)
    xs = {s.sample() for _ in range(500)}
    assert xs <= {64.0, 128.0, 256.0, 512.0, 1024.0, 2048.0, 4096.0} and len(xs) >= 5
    s = opt.build_sampler({"min": 1, "max": 4, "int": True}, 0.2, rng)
    assert {s.sample() for _ in range(400)} == {1, 2, 3, 4}
    assert opt.build_sampler(7, rng=rng).sample() == 7
    assert opt.build_sampler([3, 5], rng=rng).sample() in (3, 5)
    with pytest.raises(ValueError):
        opt.build_sampler({"def": 3}, rng=rng)


def test_log_sampler_is_log_uniform():
    rng = np.random.default_rng(1)
    s = opt.build_sampler({"min": 64, "max": 4096, "log": True}, 0.2, rng)
    xs = np.array([s.sample() for _ in range(20000)])
    frac_low = (xs < 512).mean()          # log-uniform: P(x < 512) = ln(8)/ln(64) = 0.5
    assert abs(frac_low - 0.5) < 0.02


def test_success_factor_and_jitter_adaptation():
    assert opt.success_factor(0.3) == 1.0
    assert opt.success_factor(0.1) == 0.8
    assert opt.success_factor(0.01) == 0.5
    rng = np.random.default_rng(2)
    samplers = opt.build_samplers({"a": {"def": 10, "min": 0, "max": 100}}, ["a"], 0.1, rng)
    j0 = samplers["a"].state["jitter"]
    scores = iter([1.0, 0.5, 0.4, 0.3, 0.2])          # 1 improvement in 5 trials -> rate 0.2 -> 0.8
    res = opt.search(["a"], samplers, lambda p, i: (next(scores), None, None), rounds=1, trials=5, skip_final=True)
    assert samplers["a"].state["jitter"] == pytest.approx(0.8 * j0)
    assert samplers["a"].center == res[1]["a"]


def test_search_dedups_and_finals():
    rng = np.random.default_rng(3)
    samplers = opt.build_samplers({"a": [1, 2]}, ["a"], 0.1, rng)
    calls = []

    def trial(p, info):
        calls.append((p["a"], info.get("is_final")))
        return float(p["a"]), None, "m%d" % p["a"]

    result, best, _ = opt.search(["a"], samplers, trial, rounds=2, trials=6, make_key=lambda p: str(p["a"]))
    assert best["a"] == 2 and result == "m2"
    assert calls[-1] == (2, True)
    assert len(calls) <= 2 * 2 + 1                  # at most 2 distinct keys per round


def test_search_batched_uses_proposals():
    rng = np.random.default_rng(4)
    samplers = opt.build_samplers({"a": {"def": 5, "min": 0, "max": 10}}, ["a"], 0.2, rng)
    seen = []

    def propose(s, n, hist):
        seen.append(len(hist))
        return [{"a": float(i)} for i in range(n)]

    best, score, hist = opt.search_batched(["a"], samplers, lambda b, info: [p["a"] for p in b], rounds=2, trials=4,
                                           propose=propose)
    assert seen == [0, 4] and score == 3.0 and len(hist) == 8


# ---------------------------------------------------------------------------
# surrogate MLP oracle (CPU): analytic gradient vs central finite differences

def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp_oracle as MO
    rng = np.random.default_rng(8)
    p = MO.init(3).astype(np.float64)
    p += rng.standard_normal(p.size) * 0.01        # break exact zeros of the biases
    X = rng.standard_normal((40, 12))
    Y = rng.standard_normal(40)
    g = MO.gradient(p, X, Y)
    idx = rng.choice(p.size, 300, replace=False)
    eps = 1e-6
    for i in idx:
        d = np.zeros_like(p)
        d[i] = eps
        fd = (MO.loss(p + d, X, Y) - MO.loss(p - d, X, Y)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 + 1e-4 * abs(g[i]), i


def test_mlp_oracle_adam_reduces_loss():
    from oracle import mlp_oracle as MO
    rng = np.random.default_rng(9)
    X = rng.standard_normal((64, 12))
    Y = np.sin(X[:, 0]) + 0.5 * X[:, 1] ** 2
    p0 = MO.init(4)
    p1, _, _ = MO.adam(p0, X, Y, 200, 3e-3)
    assert MO.loss(p1, X, Y) < 0.3 * MO.loss(p0, X, Y)


def test_mlp_oracle_adam_first_step_closed_form():
    """From zero moments Adam's first step is lr * g / (|g| + eps) per
    parameter after both bias corrections (Kingma & Ba 2015, Alg. 1); a
    missing correction scales it by 0.1 or 1/sqrt(0.001).  The moments after
    it are 0.1 g and 0.001 g^2, which catches swapped betas (their corrected
    first step is the same)."""
    from oracle import mlp_oracle as MO
    rng = np.random.default_rng(10)
    X = rng.standard_normal((50, 12))
    Y = rng.standard_normal(50)
    p0 = MO.init(5).astype(np.float64) + rng.standard_normal(MO.NPARAMS) * 0.01
    g = MO.gradient(p0, X, Y)
    lr = 1e-3
    p1, m, v = MO.adam(p0, X, Y, 1, lr)
    assert np.allclose(p0 - p1, lr * g / (np.abs(g) + 1e-8), rtol=1e-9, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_mlp_oracle_adam_matches_torch():
    """Five steps of the oracle's Adam against torch.optim.Adam (float64, the
    same initial parameters and full-batch mean squared error)."""
    torch = pytest.importorskip("torch")
    from oracle import mlp_oracle as MO
    rng = np.random.default_rng(11)
    X = rng.standard_normal((60, 12))
    Y = np.sin(X[:, 0]) + 0.3 * X[:, 1]
    p0 = MO.init(6).astype(np.float64)
    w = MO.unpack(p0.copy())
    params = [torch.tensor(w[name], dtype=torch.float64, requires_grad=True) for name, _ in MO.SIZES]
    opt_ = torch.optim.Adam(params, lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    Xt, Yt = torch.tensor(X), torch.tensor(Y)
    for _ in range(5):
        W1, b1, W2, b2, W3, b3, W4, b4 = params
        a1 = torch.relu(Xt @ W1.T + b1)
        a2 = torch.relu(a1 @ W2.T + b2)
        a3 = torch.relu(a2 @ W3.T + b3)
        out = a3 @ W4 + b4[0]
        loss = torch.mean((out - Yt) ** 2)
        opt_.zero_grad()
        loss.backward()
        opt_.step()
    want = np.concatenate([q.detach().numpy().reshape(-1) for q in params])
    got, _, _ = MO.adam(p0, X, Y, 5, 3e-3)
    assert np.allclose(got, want, rtol=1e-10, atol=1e-12), np.abs(got - want).max()


def test_surrogate_features_and_target():
    from santoku.tsetlin import surrogate as S
    p = dict(clauses=2048, clause_tolerance=64, clause_maximum=64, target=0.125, specificity=10.0, include_bits=1)
    assert S.effective_target(p) == 256
    assert S.effective_target(dict(p, target=-1)) == int(np.sqrt(1024) * 64)
    f = S.features(p, {"epochs": 10, "density": 0.05, "classes": 20, "features": 4096, "best": 0.9})
    assert len(f) == 12 and f[2] == pytest.approx(0.125) and f[0] == pytest.approx(11 / 12)


SPACE = {"clauses": {"def": 512, "min": 64, "max": 4096, "log": True, "round": 8},
         "clause_tolerance": {"def": 8, "min": 1, "max": 64, "log": True, "int": True},
         "clause_maximum": {"def": 16, "min": 1, "max": 64, "log": True, "int": True},
         "target": [0.125, 0.25],
         "specificity": {"def": 10, "min": 2, "max": 40},
         "include_bits": 1}


def test_candidate_space_matches_samplers():
    """The proposer's pooled candidates: every draw and local move stays on
    the sampler's grid (bounds, rounding, list values); pooled features equal
    features() of each candidate after the tolerance swap; the centred pool
    follows the sampler's own distribution; moves reach the bounds."""
    from santoku.tsetlin import surrogate as S
    smp = opt.build_samplers(SPACE, opt.PARAM_NAMES, 0.3, np.random.default_rng(3))
    sp = S.Space(smp, opt.PARAM_NAMES, np.random.default_rng(1))
    ctx = {"epochs": 10, "density": 0.5, "classes": 4, "features": 256, "best": 0.6}
    cols = S._concat(sp.draw(2000, True), sp.draw(2000, False))
    moved = sp.move(cols, np.arange(64), 64)
    pool = S._concat(cols, moved)
    n = pool["clauses"].shape[0]
    F = S.features_batch(S._with_defaults(S._swapped(pool), n), ctx)
    for i in range(0, n, 5):
        p = sp.params(pool, i)
        assert isinstance(p["clauses"], int) and 64 <= p["clauses"] <= 4096 and p["clauses"] % 8 == 0, p
        assert isinstance(p["clause_tolerance"], int) and 1 <= p["clause_tolerance"] <= 64, p
        assert 2 <= p["specificity"] <= 40 and p["target"] in (0.125, 0.25) and p["include_bits"] == 1, p
        want = np.array(S.features(opt._swap_tolerance(dict(p)), ctx), np.float32)
        assert np.array_equal(want, F[i]), (p, want, F[i])
    ref = [opt.sample_params(smp, opt.PARAM_NAMES) for _ in range(4000)]
    for name in ("clauses", "specificity", "clause_tolerance"):
        a, b = np.array([r[name] for r in ref], float), cols[name][:2000]
        assert abs(np.median(a) - np.median(b)) <= 0.1 * np.median(a) + 1, name
    assert (moved["clauses"] == 4096).any() and (moved["specificity"] == 40).any()
    assert (moved["clause_tolerance"] == 1).any() and (moved["clauses"] == 64).any()


def test_candidate_space_keeps_non_numeric_options_fixed():
    """An option the pool cannot vary numerically (a list of strings, a
    boolean) is held at one draw and passed through to the proposals."""
    from santoku.tsetlin import surrogate as S
    space = {"clauses": {"def": 64, "min": 8, "max": 512, "log": True, "round": 8}, "target": [0.25, 0.5],
             "specificity": 4.0, "clause_tolerance": ["a", "b"], "include_bits": True}
    smp = opt.build_samplers(space, opt.PARAM_NAMES, 0.3, np.random.default_rng(0))
    sp = S.Space(smp, opt.PARAM_NAMES, np.random.default_rng(0))
    cols = sp.draw(8, False)
    moved = sp.move(cols, [0, 1], 4)
    for p in [sp.params(cols, i) for i in range(8)] + [sp.params(moved, i) for i in range(8)]:
        assert p["clause_tolerance"] in ("a", "b") and p["include_bits"] is True and p["specificity"] == 4.0
        assert 8 <= p["clauses"] <= 512 and p["target"] in (0.25, 0.5)
