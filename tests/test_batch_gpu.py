"""Batched search trials (GPU): every trial equals its solo run bit for bit."""
import numpy as np
import pytest

import santoku.tsetlin as tm
from oracle import oracle as O
from santoku.tsetlin import data as D
from santoku.tsetlin import trials as TR

pytestmark = pytest.mark.gpu

BASE = dict(classes=4, features=256, negative=-1.0)


def trial_opts(rng, k):
    out = []
    for i in range(k):
        c = int(2 ** rng.uniform(np.log2(8), np.log2(96)))
        out.append(dict(clauses=c, clause_tolerance=int(rng.integers(2, 16)), clause_maximum=int(rng.integers(8, 32)),
                        target=float(max(1, c // int(rng.choice([4, 8])))), specificity=float(rng.uniform(2, 40)),
                        seed=1000 + i))
    return out


def solo(opts, rows, y, epochs):
    m = tm.classifier(dict(BASE, **opts))
    m.train(samples=rows.shape[0], problems=rows, solutions=y, iterations=epochs)
    return m.checkpoint()


def test_batch_trials_bitexact_vs_solo():
    d = D.xor_and_task(11, n_train=1500, n_test=300)
    rng = np.random.default_rng(5)
    trials = trial_opts(rng, 7)
    b = TR.Batch(BASE, trials)
    b.train(1500, d["train_rows"], d["train_y"], 3)
    tune = b.tuning()
    assert tune["lanes"] * tune["words_per_lane"] >= 16
    for t, o in enumerate(trials):
        got = b.machine(t).checkpoint()
        assert np.array_equal(got, solo(o, d["train_rows"], d["train_y"], 3)), t
    # one trial against the oracle replay as well
    orc = O.Machine(O.params("classifier", **dict(BASE, **trials[2])))
    orc.train_replay(d["train_rows"], labels=d["train_y"], epochs=3)
    assert np.array_equal(b.machine(2).checkpoint(), orc.actions)
    assert np.array_equal(b.machine(2).predict(d["test_rows"], 300), orc.predict(d["test_rows"]))


def test_batch_freeze_equals_shorter_solo_and_accuracy():
    import torch
    d = D.xor_and_task(12, n_train=1200, n_test=400)
    trials = trial_opts(np.random.default_rng(6), 4)
    b = TR.Batch(BASE, trials)
    seen = []

    def each(epoch, active):
        seen.append((epoch, active.copy()))
        if epoch == 2:
            active[1] = 0          # trial 1 stops after epoch 2
        return epoch < 4

    b.train(1200, d["train_rows"], d["train_y"], 6, each)
    assert [e for e, _ in seen] == [1, 2, 3, 4]
    assert np.array_equal(b.machine(1).checkpoint(), solo(trials[1], d["train_rows"], d["train_y"], 2))
    assert np.array_equal(b.machine(0).checkpoint(), solo(trials[0], d["train_rows"], d["train_y"], 4))
    rows = torch.from_numpy(d["test_rows"]).cuda()
    ys = torch.from_numpy(d["test_y"]).cuda()
    acc = b.accuracy(rows, ys, 400)
    for t in range(4):
        pred = b.machine(t).predict(d["test_rows"], 400)
        assert acc[t] == pytest.approx((pred == d["test_y"]).mean())


def test_batch_device_inputs_and_search_round():
    import torch
    from santoku.tsetlin import optimize as opt
    d = D.xor_and_task(13, n_train=1000, n_test=300)
    rows = torch.from_numpy(d["train_rows"]).cuda()
    ys = torch.from_numpy(d["train_y"]).cuda()
    dev_rows = torch.from_numpy(d["test_rows"]).cuda()
    dev_y = torch.from_numpy(d["test_y"]).cuda()
    rng = np.random.default_rng(7)
    samplers = opt.build_samplers({"clauses": {"def": 32, "min": 8, "max": 64, "round": 8},
                                   "clause_tolerance": {"def": 4, "min": 1, "max": 8, "int": True},
                                   "clause_maximum": {"def": 8, "min": 4, "max": 16, "int": True},
                                   "target": {"def": 8, "min": 2, "max": 16, "int": True},
                                   "specificity": {"def": 8, "min": 2, "max": 20},
                                   "include_bits": 1}, opt.PARAM_NAMES, 0.2, rng)
    fn = TR.batch_trial_fn(BASE, rows, ys, dev_rows, dev_y, iterations=3, patience=2)
    best, score, hist = opt.search_batched(opt.PARAM_NAMES, samplers, fn, rounds=2, trials=6, make_key=opt.make_key)
    assert len(hist) >= 6 and 0.0 <= score <= 1.0 and best is not None
