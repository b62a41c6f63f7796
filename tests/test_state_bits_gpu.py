"""GPU parity at state bits other than 8 (the `state` option, capi.c:642;
counters of state_bits - 1 count planes plus the action plane, automata.h).

The oracle's counter ops and replay trainer are pinned at these state bits
against the counter model on the CPU (tests/test_oracle.py); here the
engine's training must equal the replay bit for bit, through the C ABI.
"""
import struct

import numpy as np
import pytest

import santoku.tsetlin as tm
from oracle import oracle as O
from santoku.tsetlin import data as D
from santoku.tsetlin import trials as TR
from state import assert_same_state

pytestmark = pytest.mark.gpu


def rows_for(features, n, rng, density=0.3):
    return D.bits_to_cvec(rng.random((n, features)) < density)


STATE_SHAPES = [
    # state, features, classes, clauses, kind, include_bits, n, epochs
    (2, 13, 3, 16, "classifier", 1, 200, 3),      # one count plane, masked tail byte
    (3, 63, 2, 16, "classifier", 2, 150, 3),      # 16-byte rows, include_bits 2
    (4, 100, 3, 8, "encoder", 1, 150, 2),         # 25-byte rows (repacked)
    (5, 256, 4, 16, "classifier", 3, 200, 2),     # 16 lanes
    (7, 1024, 2, 8, "classifier", 1, 100, 2),     # 64 lanes x 1 word
    (12, 2048, 3, 8, "encoder", 2, 60, 2),        # 64 x 2
    (16, 8192, 2, 8, "classifier", 1, 40, 1),     # 64 x 8, the widest row at 16 state bits
    (9, 4096, 2, 8, "classifier", 1, 60, 1),      # 64 x 4
]


@pytest.mark.parametrize("state,features,classes,clauses,kind,include_bits,n,epochs", STATE_SHAPES)
def test_train_bitexact_vs_replay_state_bits(state, features, classes, clauses, kind, include_bits, n, epochs):
    rng = np.random.default_rng(features * 3 + state)
    spec = max(2.0, features / 16.0)
    o = dict(classes=classes, features=features, clauses=clauses, clause_tolerance=6, clause_maximum=8,
             specificity=spec, target=8, include_bits=include_bits, seed=4321, state=state)
    eng = tm.create(kind, dict(o, reusable=True))
    orc = O.Machine(O.params(kind, **o))
    rows = rows_for(features, n, rng)
    if kind == "classifier":
        labels = rng.integers(0, classes, n).astype(np.int64)
        eng.train(samples=n, problems=rows, solutions=labels, iterations=epochs)
        orc.train_replay(rows, labels=labels, epochs=epochs)
    else:
        codes = rng.integers(0, 256, (n, (classes + 7) // 8), dtype=np.uint8)
        eng.train(samples=n, sentences=rows, codes=codes, iterations=epochs)
        orc.train_replay(rows, codes=codes, epochs=epochs)
    assert orc.stats[1] > 0 and orc.stats[3] > 0
    assert eng.info()["state_bits"] == state
    assert eng.info()["state_chunks"] == classes * orc.p.clauses * (state - 1) * orc.p.input_chunks
    assert_same_state(eng, orc, label="state %d" % state)
    test = rows_for(features, 150, rng)
    assert np.array_equal(eng.predict(test, 150), orc.predict(test))


def test_state_bits_persist_and_every_tile():
    # persisted header carries the state bits and counts size; every tile the
    # autotuner may pick at 6 state bits trains the same machine
    rng = np.random.default_rng(8)
    o = dict(classes=3, features=200, clauses=16, clause_tolerance=5, clause_maximum=6, specificity=4.0, target=6,
             seed=3, state=6)
    rows = rows_for(200, 160, rng)
    labels = rng.integers(0, 3, 160).astype(np.int64)
    orc = O.Machine(O.params("classifier", **o))
    orc.train_replay(rows, labels=labels, epochs=2)
    import os
    for tile in ("8x2", "16x1", "16x2", "32x1", "32x2", "64x1"):
        os.environ["TSETLIN_AMD_TILE"] = tile
        try:
            eng = tm.classifier(dict(o, reusable=True))
            eng.train(samples=160, problems=rows, solutions=labels, iterations=2)
        finally:
            del os.environ["TSETLIN_AMD_TILE"]
        assert_same_state(eng, orc, label=tile)
    blob = eng.persist()
    h = struct.unpack("<i13I", blob[:56])
    assert h[7] == 6 and h[12] == 3 * 16 * 5 * 50 and h[13] == 3 * 16 * 50
    loaded = tm.load(blob, True)
    assert loaded.info()["state_bits"] == 6
    assert np.array_equal(loaded.predict(rows, 160), orc.predict(rows))


def test_state_bits_individualized_and_batch():
    rng = np.random.default_rng(12)
    widths, n = [9, 70, 30], 150
    parts, offs, o_ = [], [], 0
    for k in widths:
        b = rows_for(k, n, rng).reshape(-1)
        parts.append(b)
        offs.append(o_)
        o_ += b.size
    prob, offs = np.concatenate(parts), np.array(offs, np.int64)
    o = dict(clauses=16, clause_tolerance=6, clause_maximum=8, specificity=3.0, target=8, seed=21, state=4)
    eng = tm.create("classifier", dict(o, classes=3, individualized=True, reusable=True,
                                       feat_offsets=np.concatenate([[0], np.cumsum(widths)])))
    orc = O.IndMachine(O.params("classifier", classes=3, features=1, **o), widths)
    labels = rng.integers(0, 3, n).astype(np.int64)
    eng.train(samples=n, problems=prob, solutions=labels, dim_offsets=offs, iterations=2)
    orc.train_replay(prob, offs, n, labels=labels, epochs=2)
    assert_same_state(eng, orc, label="individualized, 4 state bits")
    # batched trials at 11 state bits equal their solo machines
    rows = rows_for(40, 200, rng)
    y = rng.integers(0, 4, 200).astype(np.int64)
    base = dict(classes=4, features=40, state=11)
    trials = [dict(clauses=c, target=t, seed=s, clause_tolerance=5, clause_maximum=6, specificity=3.0)
              for c, t, s in ((8, 4, 1), (24, 6, 2), (16, 5, 3))]
    b = TR.Batch(base, trials)
    b.train(200, rows, y, 3)
    for i, t in enumerate(trials):
        ref = O.Machine(O.params("classifier", **dict(base, **t)))
        ref.train_replay(rows, labels=y, epochs=3)
        assert_same_state(b.machine(i), ref, label="batch trial %d" % i)


def test_state_bits_limits():
    base = dict(classes=2, clauses=8, clause_tolerance=3, clause_maximum=3, specificity=2.0)
    with pytest.raises(tm.TsetlinError, match="state bits above 16"):
        tm.classifier(dict(base, features=8, state=17))
    with pytest.raises(tm.TsetlinError, match="bits must be greater than 1"):
        tm.classifier(dict(base, features=8, state=1))
    # 16384 features need the 64 x 16 tile, built for up to 8 state bits
    tm.classifier(dict(base, features=16384, state=8))
    with pytest.raises(tm.TsetlinError, match="largest train tile for 9 state bits"):
        tm.classifier(dict(base, features=16384, state=9))


def test_batch_tuning_is_per_state_bits():
    """Two batches with the same row width, trial count and workgroup count,
    at 8 and then 16 state bits: the second must not reuse a tile tuned for
    the first (the widest 8-state-bit tile holds 16 words per lane, more than
    16 state bits fit), and both equal their solo machines."""
    rng = np.random.default_rng(21)
    rows = rows_for(4096, 48, rng)
    y = rng.integers(0, 2, 48).astype(np.int64)
    trials = [dict(clauses=8, target=4, seed=s, clause_tolerance=5, clause_maximum=6, specificity=400.0)
              for s in (1, 2)]
    for state in (8, 16):
        base = dict(classes=2, features=4096, state=state)
        b = TR.Batch(base, trials)
        b.train(48, rows, y, 2)
        tune = b.tuning()
        assert tune["words_per_lane"] * state <= 128, (state, tune)
        for i, t in enumerate(trials):
            ref = O.Machine(O.params("classifier", **dict(base, **t)))
            ref.train_replay(rows, labels=y, epochs=2)
            assert_same_state(b.machine(i), ref, label="state %d trial %d" % (state, i))
