"""Workload 2 at its full size on the GPU: 200,000 rows of 4096 features
(1024-byte rows), 20 classes x 2048 clauses, target 256, specificity 10,
tolerance = maximum = 64.

The oracle cannot replay the whole machine in test time, but chunks are
independent: it replays 16 chunks over all 200,000 rows for two epochs,
including chunks of the most and the least frequent class, and those must
match the engine's full-size epochs bit for bit, include plane and every
count plane.  Predict is checked against the oracle on a
sample of rows and for independence from how the rows are batched.
"""
import numpy as np
import pytest

import santoku.tsetlin as tm
from oracle import oracle as O
from santoku.tsetlin import data as D
from state import assert_same_state

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

OPTS = dict(classes=20, features=4096, clauses=2048, clause_tolerance=64, clause_maximum=64, specificity=10.0,
            target=256, seed=123)


@pytest.fixture(scope="module")
def trained():
    d = D.sparse_planted_task(7, n_train=200000, n_test=20000)
    eng = tm.classifier(dict(OPTS, reusable=True))
    rows = torch.from_numpy(d["train_rows"]).cuda()
    y = torch.from_numpy(d["train_y"]).cuda()
    eng.train(samples=200000, problems=rows, solutions=y, iterations=2)
    return d, eng


def test_full_size_epoch_matches_chunk_replay(trained):
    d, eng = trained
    y = d["train_y"]
    counts = np.bincount(y, minlength=20)
    big, small = int(np.argmax(counts)), int(np.argmin(counts))
    rng = np.random.default_rng(4)
    chunks = {big * 256 + 0, big * 256 + 255, small * 256 + 31, 0 * 256 + 7, 9 * 256 + 128, 17 * 256 + 64}
    chunks |= set(int(c) for c in rng.choice(20 * 256, 10, replace=False))
    chunks = sorted(chunks)
    orc = O.Machine(O.params("classifier", **OPTS))
    orc.train_replay_chunks(d["train_rows"], chunks, labels=y, epochs=2)
    assert orc.stats[1] > 0 and orc.stats[3] > 0 and orc.stats[4] > 0
    assert_same_state(eng, orc, chunks=chunks, label="workload 2, chunks %s" % chunks)


def test_full_size_predict(trained):
    d, eng = trained
    test = torch.from_numpy(d["test_rows"]).cuda()
    full = eng.predict(test, 20000)
    for s, e in ((0, 1), (1, 4097), (4097, 13000), (13000, 20000)):
        assert torch.equal(eng.predict(test[s:e].contiguous(), e - s), full[s:e])
    orc = O.Machine(O.params("classifier", **OPTS))
    orc.actions[:] = eng.checkpoint()
    pick = np.random.default_rng(0).choice(20000, 256, replace=False)
    want = orc.predict(d["test_rows"][pick])
    assert np.array_equal(full.cpu().numpy()[pick], want)
