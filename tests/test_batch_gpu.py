"""Batched search trials (GPU): every trial equals its solo run bit for bit."""
import numpy as np
import pytest

import santoku.tsetlin as tm
from oracle import oracle as O
from santoku.tsetlin import data as D
from santoku.tsetlin import trials as TR
from state import assert_same_state

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


def solo(opts, rows, y, epochs, base=BASE):
    m = tm.classifier(dict(base, reusable=True, **opts))
    m.train(samples=rows.shape[0], problems=rows, solutions=y, iterations=epochs)
    return m


def assert_same_machine(a, b, label=""):
    """Two engine machines hold the same actions and counters."""
    assert np.array_equal(a.checkpoint(), b.checkpoint()), "%s: actions differ" % label
    assert np.array_equal(a.counts(), b.counts()), "%s: count planes differ" % label


def test_batch_trials_bitexact_vs_solo():
    d = D.xor_and_task(11, n_train=1500, n_test=300)
    rng = np.random.default_rng(5)
    trials = trial_opts(rng, 7)
    b = TR.Batch(BASE, trials)
    b.train(1500, d["train_rows"], d["train_y"], 3)
    tune = b.tuning()
    assert tune["lanes"] * tune["words_per_lane"] >= 16
    for t, o in enumerate(trials):
        assert_same_machine(b.machine(t), solo(o, d["train_rows"], d["train_y"], 3), "trial %d" % t)
    # one trial against the oracle replay as well
    orc = O.Machine(O.params("classifier", **dict(BASE, **trials[2])))
    orc.train_replay(d["train_rows"], labels=d["train_y"], epochs=3)
    assert_same_state(b.machine(2), orc, label="trial 2")
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
    assert_same_machine(b.machine(1), solo(trials[1], d["train_rows"], d["train_y"], 2), "frozen trial")
    assert_same_machine(b.machine(0), solo(trials[0], d["train_rows"], d["train_y"], 4), "trial 0")
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


# ---------------------------------------------------------------------------
# workload 4: 64 trials on workload 1's data (8192 x 256 features, 4 classes),
# clauses log-uniform in 64..4096 per class, specificity uniform in 2..40,
# target clauses/8 or clauses/4; tolerance and maximum (not fixed by the
# workload) log-uniform in 1..64

C4_SPACE = {"clauses": {"def": 512, "min": 64, "max": 4096, "log": True, "round": 8},
            "clause_tolerance": {"def": 8, "min": 1, "max": 64, "log": True, "int": True},
            "clause_maximum": {"def": 16, "min": 1, "max": 64, "log": True, "int": True},
            "target": [0.125, 0.25],
            "specificity": {"def": 10, "min": 2, "max": 40},
            "include_bits": 1}


def c4_trials(seed, k=64):
    from santoku.tsetlin import optimize as opt
    samplers = opt.build_samplers(C4_SPACE, opt.PARAM_NAMES, 1.0, np.random.default_rng(seed))
    out = []
    for i in range(k):
        p = {n: (s.sample(None) if s.type == "range" else s.sample()) for n, s in samplers.items() if s is not None}
        out.append(dict(opt._swap_tolerance(p), seed=5000 + i))
    return out


def test_workload4_sixty_four_trials_bitexact():
    """All 64 trials of one batched launch per epoch equal their solo runs
    (actions and every count plane), and four of them equal the oracle's
    replay of chunks spread over their classes."""
    import torch
    d = D.xor_and_task(7)
    rows = torch.from_numpy(d["train_rows"]).cuda()
    ys = torch.from_numpy(d["train_y"]).cuda()
    trials = c4_trials(3)
    cl = [t["clauses"] for t in trials]
    assert min(cl) < 128 and max(cl) > 2048, cl
    b = TR.Batch(BASE, trials)
    b.train(8192, rows, ys, 3)
    for t, o in enumerate(trials):
        assert_same_machine(b.machine(t), solo(o, rows, ys, 3), "trial %d %s" % (t, o))
    order = np.argsort(cl)
    for t in (int(order[0]), int(order[-1]), int(order[20]), int(order[45])):
        o = dict(BASE, **trials[t])
        orc = O.Machine(O.params("classifier", **o))
        per = orc.p.clause_chunks
        chunks = sorted({0, per - 1, per + per // 2, 3 * per + per // 3, 4 * per - 1})
        orc.train_replay_chunks(d["train_rows"], chunks, labels=d["train_y"], epochs=3)
        assert_same_state(b.machine(t), orc, chunks=chunks, label="trial %d %s" % (t, o))


def test_batch_class_accuracy_is_the_evaluator():
    """The batch scorer's macro precision / recall / F1 equal
    evaluator.class_accuracy of each trial's predictions, double for double."""
    import torch
    from santoku.tsetlin import evaluator as E
    d = D.xor_and_task(8, n_train=2000, n_test=700)
    trials = trial_opts(np.random.default_rng(9), 5)
    b = TR.Batch(BASE, trials)
    b.train(2000, d["train_rows"], d["train_y"], 3)
    rows = torch.from_numpy(d["test_rows"]).cuda()
    ys = torch.from_numpy(d["test_y"]).cuda()
    got = b.class_accuracy(rows, ys, 700)
    for t in range(5):
        want = E.class_accuracy(b.machine(t).predict(d["test_rows"], 700), d["test_y"], 700, 4)
        assert (got[t, 0], got[t, 1], got[t, 2]) == (want["precision"], want["recall"], want["f1"]), t


def test_batch_trial_fn_restores_best_epoch():
    """optimize.lua:586-653 in the batched driver: with patience, a trial ends
    on its best epoch's actions (restored from the device checkpoint) and its
    result is that epoch's score; each such machine equals a solo run of that
    many epochs."""
    import torch
    from santoku.tsetlin import optimize as opt
    d = D.xor_and_task(14, n_train=3000, n_test=800)
    rows = torch.from_numpy(d["train_rows"]).cuda()
    ys = torch.from_numpy(d["train_y"]).cuda()
    drows = torch.from_numpy(d["test_rows"]).cuda()
    dys = torch.from_numpy(d["test_y"]).cuda()
    trials = trial_opts(np.random.default_rng(15), 6)
    params = [{k: v for k, v in t.items() if k != "seed"} for t in trials]
    fn = TR.batch_trial_fn(BASE, rows, ys, drows, dys, iterations=8, patience=2, seed=300)
    scores = fn(params, {"round": 1})
    b = fn.last_batch
    final = b.class_accuracy(drows, dys, 800)[:, 2]
    assert np.array_equal(np.array(scores), final)
    assert np.array_equal(final, fn.last_best)
    for t in range(6):
        e = int(fn.last_best_epoch[t])
        assert 1 <= e <= 8
        # the driver swaps an inverted tolerance / maximum (optimize.lua:731-733)
        ref = solo(dict(opt._swap_tolerance(dict(params[t])), seed=300 + t), d["train_rows"], d["train_y"], e)
        assert np.array_equal(b.machine(t).checkpoint(), ref.checkpoint()), (t, e)


# ---------------------------------------------------------------------------
# encoder trials (the encoder search, optimize.lua:842): a batch of encoders
# on one sentence/code set, every trial equal to its solo tm_train_encoder run

ENC_BASE = dict(classes=16, features=128)


def test_encoder_batch_bitexact_vs_solo():
    import torch
    d = D.projection_codes_task(21, n=600, visible=128, hidden=16)
    rng = np.random.default_rng(8)
    trials = trial_opts(rng, 5)
    trials[1]["specificity"] = 1.5          # many draws per event: repeat hits on the levels
    b = TR.Batch(ENC_BASE, trials, encoder=True)
    b.train(600, d["rows"], d["codes"], 3)
    for t, o in enumerate(trials):
        m = tm.encoder(dict(ENC_BASE, reusable=True, **o))
        m.train(samples=600, sentences=d["rows"], codes=d["codes"], iterations=3)
        assert_same_machine(b.machine(t), m, "encoder trial %d" % t)
        assert np.array_equal(b.machine(t).predict(d["rows"][:64], 64), m.predict(d["rows"][:64], 64))
    # device-resident sentences and codes, and an early stop through the callback
    rows = torch.from_numpy(d["rows"]).cuda()
    codes = torch.from_numpy(d["codes"]).cuda()
    b2 = TR.Batch(ENC_BASE, trials, encoder=True)

    def each(epoch, active):
        if epoch == 1:
            active[3] = 0
        return True

    b2.train(600, rows, codes, 3, each)
    for t, o in enumerate(trials):
        m = tm.encoder(dict(ENC_BASE, reusable=True, **o))
        m.train(samples=600, sentences=d["rows"], codes=d["codes"], iterations=1 if t == 3 else 3)
        assert_same_machine(b2.machine(t), m, "encoder trial %d (device inputs)" % t)
    # six state bits and include_bits 2 (another train object, another include test)
    base6 = dict(ENC_BASE, state=6)
    trials6 = [dict(o, include_bits=2) for o in trials[:3]]
    b3 = TR.Batch(base6, trials6, encoder=True)
    b3.train(600, d["rows"], d["codes"], 2)
    for t, o in enumerate(trials6):
        m = tm.encoder(dict(base6, reusable=True, **o))
        m.train(samples=600, sentences=d["rows"], codes=d["codes"], iterations=2)
        assert_same_machine(b3.machine(t), m, "encoder trial %d (6 state bits)" % t)


def test_batch_type_is_enforced():
    import torch
    from santoku.tsetlin import TsetlinError
    d = D.projection_codes_task(22, n=64, visible=128, hidden=16)
    b = TR.Batch(ENC_BASE, trial_opts(np.random.default_rng(9), 2), encoder=True)
    rows = torch.from_numpy(d["rows"]).cuda()
    labels = torch.zeros(64, dtype=torch.int64).cuda()
    with pytest.raises(TsetlinError):
        b.class_accuracy(rows, labels, 64)
    c = TR.Batch(BASE, trial_opts(np.random.default_rng(9), 2))
    c.encoder = True                        # a classifier batch called as an encoder one
    crow = torch.zeros((64, c.machine(0).info()["input_chunks"]), dtype=torch.uint8).cuda()
    ccodes = torch.zeros((64, c.machine(0).info()["class_chunks"]), dtype=torch.uint8).cuda()
    with pytest.raises(TsetlinError, match="classifier trials"):
        c.train(64, crow, ccodes, 1)


def test_encoder_trial_fn_scores_and_restores_best_epoch():
    """The encoder search's batch function: per-epoch mean_hamming of each
    trial's codes on the dev rows (the encoder specs' search_metric,
    encoder/mnist.lua:368-372), the best epoch restored; each trial's result
    equals the evaluator on a solo encoder trained that many epochs."""
    import torch
    from santoku.tsetlin import evaluator as EV
    from santoku.tsetlin import optimize as opt
    d = D.projection_codes_task(23, n=900, visible=128, hidden=16)
    rows = torch.from_numpy(d["rows"][:700]).cuda()
    codes = torch.from_numpy(d["codes"][:700]).cuda()
    drows = torch.from_numpy(d["rows"][700:]).cuda()
    dcodes = torch.from_numpy(d["codes"][700:]).cuda()
    trials = trial_opts(np.random.default_rng(16), 5)
    params = [{k: v for k, v in t.items() if k != "seed"} for t in trials]
    fn = TR.batch_trial_fn(ENC_BASE, rows, codes, drows, dcodes, iterations=6, patience=2, seed=500, encoder=True)
    scores = fn(params, {"round": 1})
    assert np.array_equal(np.array(scores), fn.last_best)
    for t in range(5):
        e = int(fn.last_best_epoch[t])
        assert 1 <= e <= 6
        m = tm.encoder(dict(ENC_BASE, reusable=True, **dict(opt._swap_tolerance(dict(params[t])), seed=500 + t)))
        m.train(samples=700, sentences=d["rows"][:700], codes=d["codes"][:700], iterations=e)
        assert np.array_equal(fn.last_batch.machine(t).checkpoint(), m.checkpoint()), (t, e)
        want = EV.encoding_accuracy(m.predict(d["rows"][700:], 200), d["codes"][700:], 200, 16)["mean_hamming"]
        assert scores[t] == want, (t, scores[t], want)


# ---------------------------------------------------------------------------
# individualized trials (the default of the reference's encoder and
# newsgroups specs): every class of every trial in one launch per epoch, each
# trial equal to its solo individualized machine

def _ragged(widths, n, rng):
    parts, offs, o = [], [], 0
    for k in widths:
        b = D.bits_to_cvec(rng.random((n, k)) < 0.3).reshape(-1)
        parts.append(b)
        offs.append(o)
        o += b.size
    return np.concatenate(parts), np.array(offs, np.int64)


@pytest.mark.parametrize("kind,widths", [("classifier", [5, 40, 64, 200]), ("encoder", [63, 9, 300, 17, 130])])
def test_individualized_batch_bitexact_vs_solo(kind, widths):
    import torch
    rng = np.random.default_rng(sum(widths))
    n, K = 240, len(widths)
    prob, offs = _ragged(widths, n, rng)
    feat = np.concatenate([[0], np.cumsum(widths)]).astype(np.int64)
    enc = kind == "encoder"
    tgt = (rng.integers(0, 256, (n, (K + 7) // 8), dtype=np.uint8) if enc
           else rng.integers(0, K, n).astype(np.int64))
    trials = trial_opts(np.random.default_rng(17), 4)
    trials[2]["specificity"] = 1.5
    base = dict(classes=K, features=1)
    b = TR.Batch(base, trials, encoder=enc, feat_offsets=feat)
    b.train(n, prob, tgt, 3, dim_offsets=offs)

    def solo_ind(o, epochs):
        m = tm.create(kind, dict(base, reusable=True, individualized=True, feat_offsets=feat, **o))
        if enc:
            m.train(samples=n, sentences=prob, codes=tgt, dim_offsets=offs, iterations=epochs)
        else:
            m.train(samples=n, problems=prob, solutions=tgt, dim_offsets=offs, iterations=epochs)
        return m

    for t, o in enumerate(trials):
        m = solo_ind(o, 3)
        assert_same_machine(b.machine(t), m, "%s ind trial %d" % (kind, t))
        assert np.array_equal(b.machine(t).predict(prob, offs, n), m.predict(prob, offs, n))
    # device inputs and a trial frozen after its first epoch
    b2 = TR.Batch(base, trials, encoder=enc, feat_offsets=feat)

    def each(epoch, active):
        if epoch == 1:
            active[1] = 0
        return True

    b2.train(n, torch.from_numpy(prob).cuda(), torch.from_numpy(tgt).cuda(), 3, each, dim_offsets=offs)
    for t, o in enumerate(trials):
        assert_same_machine(b2.machine(t), solo_ind(o, 1 if t == 1 else 3), "%s ind trial %d (device)" % (kind, t))


def test_individualized_batch_arguments():
    from santoku.tsetlin import TsetlinError
    rng = np.random.default_rng(3)
    prob, offs = _ragged([10, 20], 32, rng)
    feat = np.array([0, 10, 30], np.int64)
    b = TR.Batch(dict(classes=2, features=1), trial_opts(rng, 2), feat_offsets=feat)
    labels = rng.integers(0, 2, 32).astype(np.int64)
    with pytest.raises(TsetlinError):
        b.train(32, prob, labels, 1)                                    # no dim_offsets
    with pytest.raises(TsetlinError, match="bytes"):
        b.train(32, prob[:-1], labels, 1, dim_offsets=offs)             # problems too short for n rows
    c = TR.Batch(dict(classes=2, features=16), trial_opts(rng, 2))
    c.individualized = True                                            # a plain batch given dim_offsets
    with pytest.raises(TsetlinError, match="not individualized"):
        c.train(32, np.zeros((32, 4), np.uint8), labels, 1, dim_offsets=offs)


def test_individualized_trial_fn_restores_best_epoch():
    """The search's batch function over individualized classifier trials:
    per-epoch macro F1 through each trial's individualized predict on the dev
    buffer, the best epoch restored; each result equals the evaluator on a
    solo individualized machine trained to that epoch."""
    import torch
    from santoku.tsetlin import evaluator as EV
    from santoku.tsetlin import optimize as opt
    rng = np.random.default_rng(41)
    widths, n, nd = [12, 70, 33], 400, 150
    K = len(widths)
    prob, offs = _ragged(widths, n, rng)
    dprob, doffs = _ragged(widths, nd, rng)
    y = rng.integers(0, K, n).astype(np.int64)
    dy = rng.integers(0, K, nd).astype(np.int64)
    feat = np.concatenate([[0], np.cumsum(widths)]).astype(np.int64)
    base = dict(classes=K, features=1)
    params = [{k: v for k, v in t.items() if k != "seed"} for t in trial_opts(np.random.default_rng(42), 4)]
    fn = TR.batch_trial_fn(base, torch.from_numpy(prob).cuda(), torch.from_numpy(y).cuda(),
                           torch.from_numpy(dprob).cuda(), torch.from_numpy(dy).cuda(), iterations=6, patience=2,
                           seed=900, feat_offsets=feat, dim_offsets=offs, dev_dim_offsets=doffs)
    scores = fn(params, {"round": 1})
    assert np.array_equal(np.array(scores), fn.last_best)
    for t in range(4):
        e = int(fn.last_best_epoch[t])
        assert 1 <= e <= 6
        o = dict(opt._swap_tolerance(dict(params[t])), seed=900 + t)
        m = tm.create("classifier", dict(base, reusable=True, individualized=True, feat_offsets=feat, **o))
        m.train(samples=n, problems=prob, solutions=y, dim_offsets=offs, iterations=e)
        assert np.array_equal(fn.last_batch.machine(t).checkpoint(), m.checkpoint()), (t, e)
        want = EV.class_accuracy(m.predict(dprob, doffs, nd), dy, nd, K)["f1"]
        assert scores[t] == want, (t, scores[t], want)


def test_individualized_batch_class_accuracy_is_the_evaluator():
    """tm_batch_class_accuracy_ind_dev: every trial's macro precision, recall
    and F1 on an individualized dev buffer equal the evaluator on that
    trial's own individualized predictions."""
    import torch
    from santoku.tsetlin import evaluator as EV
    rng = np.random.default_rng(44)
    widths, n, nd = [20, 45, 8, 90], 300, 220
    K = len(widths)
    prob, offs = _ragged(widths, n, rng)
    dprob, doffs = _ragged(widths, nd, rng)
    y = rng.integers(0, K, n).astype(np.int64)
    dy = rng.integers(0, K, nd).astype(np.int64)
    feat = np.concatenate([[0], np.cumsum(widths)]).astype(np.int64)
    b = TR.Batch(dict(classes=K, features=1), trial_opts(np.random.default_rng(45), 5), feat_offsets=feat)
    b.train(n, prob, y, 2, dim_offsets=offs)
    got = b.class_accuracy(torch.from_numpy(dprob).cuda(), torch.from_numpy(dy).cuda(), nd, dim_offsets=doffs)
    for t in range(5):
        ev = EV.class_accuracy(b.machine(t).predict(dprob, doffs, nd), dy, nd, K)
        assert (got[t, 0], got[t, 1], got[t, 2]) == (ev["precision"], ev["recall"], ev["f1"]), t


def test_individualized_encoder_trial_fn():
    """The search's batch function over individualized encoder trials: each
    result is mean_hamming of the restored best epoch's codes, equal to the
    evaluator on a solo individualized encoder trained to that epoch."""
    import torch
    from santoku.tsetlin import evaluator as EV
    from santoku.tsetlin import optimize as opt
    rng = np.random.default_rng(47)
    widths, n, nd = [33, 7, 120, 64, 15, 40, 9, 70, 22], 260, 120
    K = len(widths)
    prob, offs = _ragged(widths, n, rng)
    dprob, doffs = _ragged(widths, nd, rng)
    codes = rng.integers(0, 256, (n, (K + 7) // 8), dtype=np.uint8)
    dcodes = rng.integers(0, 256, (nd, (K + 7) // 8), dtype=np.uint8)
    feat = np.concatenate([[0], np.cumsum(widths)]).astype(np.int64)
    base = dict(classes=K, features=1)
    params = [{k: v for k, v in t.items() if k != "seed"} for t in trial_opts(np.random.default_rng(48), 3)]
    fn = TR.batch_trial_fn(base, torch.from_numpy(prob).cuda(), torch.from_numpy(codes).cuda(),
                           torch.from_numpy(dprob).cuda(), torch.from_numpy(dcodes).cuda(), iterations=5, patience=2,
                           seed=700, encoder=True, feat_offsets=feat, dim_offsets=offs, dev_dim_offsets=doffs)
    scores = fn(params, {"round": 1})
    for t in range(3):
        e = int(fn.last_best_epoch[t])
        o = dict(opt._swap_tolerance(dict(params[t])), seed=700 + t)
        m = tm.create("encoder", dict(base, reusable=True, individualized=True, feat_offsets=feat, **o))
        m.train(samples=n, sentences=prob, codes=codes, dim_offsets=offs, iterations=e)
        assert np.array_equal(fn.last_batch.machine(t).checkpoint(), m.checkpoint()), (t, e)
        want = EV.encoding_accuracy(m.predict(dprob, doffs, nd), dcodes, nd, K)["mean_hamming"]
        assert scores[t] == want, (t, scores[t], want)
