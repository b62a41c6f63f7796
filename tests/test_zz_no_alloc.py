"""tm_predict_classifier_dev_noalloc through the C ABI against
tm_predict_classifier_dev: the same predictions, on two calls in a row (the
second with fewer rows, so stale class sums in the shared scratch array would
show), with rows used in place and with rows repacked into whole words, and
a defined error past the scratch size."""
import ctypes

import numpy as np
import pytest

import santoku.tsetlin as tm
from santoku.tsetlin import data as D

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def machine(features, classes, clauses, rng):
    eng = tm.classifier(classes=classes, features=features, clauses=clauses, clause_tolerance=7,
                        clause_maximum=7, specificity=2.0)
    info = eng.info()
    dens = min(0.2, 10.0 / (2 * features))
    acts = np.packbits(rng.random(info["action_chunks"] * 8) < dens, bitorder="little")
    eng.restore(acts)
    return eng


def noalloc(eng, rows, n):
    out = torch.full((n,), -7, dtype=torch.int64, device="cuda")
    rc = eng._lib.tm_predict_classifier_dev_noalloc(eng._handle(), rows.data_ptr(), n, out.data_ptr())
    return rc, out


@pytest.mark.parametrize("features,classes,clauses", [(256, 4, 64), (37, 5, 24), (4096, 20, 128)])
def test_noalloc_predict_equals_predict_dev(features, classes, clauses):
    rng = np.random.default_rng(features)
    eng = machine(features, classes, clauses, rng)
    for n in (3000, 517):
        rows = torch.from_numpy(D.bits_to_cvec(rng.random((n, features)) < 0.3)).cuda()
        rc, got = noalloc(eng, rows, n)
        assert rc == 0, eng._lib.tm_last_error()
        want = eng.predict(rows, n)
        assert torch.equal(got, want), (features, n)
    # rows starting 4 bytes into a buffer (not 16-byte aligned) take the
    # repack path through the scratch array
    buf = torch.from_numpy(D.bits_to_cvec(rng.random((401, features)) < 0.3).reshape(-1)).cuda()
    rows = buf[4:]
    rc, got = noalloc(eng, rows, 400)
    assert rc == 0, eng._lib.tm_last_error()
    assert torch.equal(got, eng.predict(rows.contiguous(), 400))


def test_noalloc_predict_size_limit():
    rng = np.random.default_rng(3)
    eng = machine(16, 3, 8, rng)
    n = (32 << 20) // (3 * 8) + 64
    rows = torch.zeros((n, 4), dtype=torch.uint8, device="cuda")
    rc, out = noalloc(eng, rows, n)
    assert rc == -5 and b"limit" in eng._lib.tm_last_error()
    assert int(out[0]) == -7                      # nothing was written
    rc, out = noalloc(eng, rows, 1000)
    assert rc == 0 and torch.equal(out, eng.predict(rows[:1000].contiguous(), 1000))
