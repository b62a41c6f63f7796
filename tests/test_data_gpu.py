"""Device packers and dev metrics (GPU) against their host restatements.

The packer's layout is bits_to_cvec's (santoku-matrix, doc/sth.md:85-91); the
metrics are evaluator/capi.c:119-203 and :497-574, whose doubles the device
path reproduces exactly (counts on the device, the reference's arithmetic
order on the host).
"""
import numpy as np
import pytest

import santoku.tsetlin as tm
from santoku.tsetlin import data as D
from santoku.tsetlin import evaluator as E

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def set_bits_of(x, rng):
    """Set-bit indices (sample * F + feature) of a bool matrix, shuffled, with duplicates."""
    n, F = x.shape
    idx = np.flatnonzero(x.reshape(-1)).astype(np.int64)
    dup = idx[rng.random(idx.size) < 0.1]
    idx = np.concatenate([idx, dup])
    rng.shuffle(idx)
    return idx


@pytest.mark.parametrize("features,n", [(1, 7), (3, 5), (4, 33), (8, 9), (37, 101), (256, 64), (4096, 40), (1000, 3)])
def test_bits_to_cvec_dev_matches_host(features, n):
    rng = np.random.default_rng(features + n)
    x = rng.random((n, features)) < 0.3
    want = D.bits_to_cvec(x)
    idx = set_bits_of(x, rng)
    bad = np.array([-1, -(2 ** 40), n * features, n * features + 5], np.int64)
    idx = np.concatenate([idx, bad])
    dev = torch.from_numpy(idx).cuda()
    rows, invalid = D.bits_to_cvec_dev(dev, n, features)
    assert invalid == bad.size
    assert np.array_equal(rows.cpu().numpy(), want)
    # an output that is not word-aligned goes through the staging buffer
    big = torch.full((n * want.shape[1] + 9,), 0x5A, dtype=torch.uint8, device="cuda")
    view = big[1:1 + n * want.shape[1]].view(n, want.shape[1])
    D.bits_to_cvec_dev(dev, n, features, out=view)
    got = big.cpu().numpy()
    assert np.array_equal(got[1:1 + want.size].reshape(want.shape), want)
    assert got[0] == 0x5A and (got[1 + want.size:] == 0x5A).all()


def test_bits_to_cvec_dev_empty_and_all():
    n, F = 17, 11
    rows, bad = D.bits_to_cvec_dev(torch.zeros(0, dtype=torch.int64, device="cuda"), n, F)
    assert bad == 0 and np.array_equal(rows.cpu().numpy(), D.bits_to_cvec(np.zeros((n, F), bool)))
    every = torch.arange(n * F, dtype=torch.int64, device="cuda")
    rows, bad = D.bits_to_cvec_dev(every, n, F)
    assert bad == 0 and np.array_equal(rows.cpu().numpy(), D.bits_to_cvec(np.ones((n, F), bool)))


@pytest.mark.parametrize("features,n", [(1, 5), (5, 9), (8, 16), (37, 70), (4096, 33), (12, 1)])
def test_flip_interleave_dev_matches_host(features, n):
    rng = np.random.default_rng(features * 3 + n)
    x = rng.random((n, features)) < 0.4
    raw = np.packbits(x, axis=1, bitorder="little")
    got = D.flip_interleave_dev(torch.from_numpy(raw).cuda(), n, features).cpu().numpy()
    assert np.array_equal(got, D.bits_to_cvec(x))


@pytest.mark.parametrize("classes,n", [(1, 50), (4, 1000), (20, 20000), (3000, 5000)])
def test_class_accuracy_dev_equals_host(classes, n):
    rng = np.random.default_rng(classes + n)
    y = rng.integers(0, classes, n).astype(np.int64)
    p = np.where(rng.random(n) < 0.6, y, rng.integers(0, classes, n)).astype(np.int64)
    # out-of-range values are skipped after the unsigned int conversion; 2^32 + c converts to c
    p[:5] = [-1, classes, 2 ** 32 + 0, -(2 ** 32) + min(1, classes - 1), classes + 7]
    y[5:8] = [-3, classes + 1, 2 ** 33]
    host = E.class_accuracy(p, y, n, classes)
    dev = E.class_accuracy(torch.from_numpy(p).cuda(), torch.from_numpy(y).cuda(), n, classes)
    assert dev == host


@pytest.mark.parametrize("hidden,n", [(1, 40), (7, 333), (256, 5000), (9000, 64)])
def test_encoding_accuracy_dev_equals_host(hidden, n):
    rng = np.random.default_rng(hidden + n)
    cc = (hidden + 7) // 8
    e = rng.integers(0, 256, (n, cc), dtype=np.uint8)
    flip = (rng.random((n, cc * 8)) < rng.uniform(0.01, 0.3, cc * 8)).astype(np.uint8)
    p = e ^ np.packbits(flip, axis=1, bitorder="little")
    host = E.encoding_accuracy(p, e, n, hidden)
    dev = E.encoding_accuracy(torch.from_numpy(p).cuda(), torch.from_numpy(e).cuda(), n, hidden)
    assert dev == host


def test_device_pipeline_end_to_end():
    # pack on the device, train and predict on device-resident rows, score on
    # the device; the host pipeline gives the same machine and the same score
    d = D.xor_and_task(4, n_train=3000, n_test=1000)
    x = d["x"]
    rng = np.random.default_rng(0)
    idx = torch.from_numpy(set_bits_of(x, rng)).cuda()
    rows, bad = D.bits_to_cvec_dev(idx, x.shape[0], x.shape[1])
    assert bad == 0
    o = dict(classes=4, features=256, clauses=64, clause_tolerance=16, clause_maximum=16, specificity=8.0,
             target=16, seed=2)
    dev_m, host_m = tm.classifier(o), tm.classifier(o)
    y = torch.from_numpy(np.concatenate([d["train_y"], d["test_y"]])).cuda()
    dev_m.train(samples=3000, problems=rows[:3000].contiguous(), solutions=y[:3000].contiguous(), iterations=3)
    host_m.train(samples=3000, problems=d["train_rows"], solutions=d["train_y"], iterations=3)
    assert np.array_equal(dev_m.checkpoint(), host_m.checkpoint())
    pred = dev_m.predict(rows[3000:].contiguous(), 1000)
    score = E.class_accuracy(pred, y[3000:].contiguous(), 1000, 4)
    want = E.class_accuracy(host_m.predict(d["test_rows"], 1000), d["test_y"], 1000, 4)
    assert score == want and score["f1"] > 0
