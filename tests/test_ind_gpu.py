"""GPU parity of individualized machines (capi.c:528-626 create, :1284-1482
train, :997-1086 predict, :752-900 restrict, :1509-1531 / :1675-1737 persist
and load) against the oracle's IndMachine, through the C ABI.

Class h reads rows of its own width from byte dim_offsets[h] of one problems
buffer.  The shapes mix widths that fall on different train tiles, rows that
are not whole words, and both padding regimes of the last byte.
"""
import struct

import numpy as np
import pytest

import santoku.tsetlin as tm
from oracle import oracle as O
from santoku.tsetlin import data as D
from state import assert_same_state

pytestmark = pytest.mark.gpu


def rows_for(features, n, rng, density=0.3):
    return D.bits_to_cvec(rng.random((n, features)) < density)


def ragged(widths, n, rng, gap=0):
    """One problems buffer holding each class's rows, plus dim_offsets; `gap`
    bytes of noise between the class blocks."""
    parts, offs, o = [], [], 0
    for k in widths:
        parts.append(rng.integers(0, 256, gap, dtype=np.uint8))
        o += gap
        b = rows_for(k, n, rng).reshape(-1)
        parts.append(b)
        offs.append(o)
        o += b.size
    return np.concatenate(parts), np.array(offs, np.int64)


def offsets_of(widths):
    return np.concatenate([[0], np.cumsum(widths)]).astype(np.int64)


def pair(kind, widths, **kw):
    # reusable by default: the machine keeps its counters for the state checks
    o = dict(clauses=16, clause_tolerance=6, clause_maximum=8, specificity=3.0, target=8, seed=77, reusable=True)
    o.update(kw)
    K = len(widths)
    eng = tm.create(kind, dict(o, classes=K, individualized=True, feat_offsets=offsets_of(widths)))
    po = {k: v for k, v in o.items() if k != "reusable"}
    orc = O.IndMachine(O.params(kind, classes=K, features=1, **po), widths)
    return eng, orc


def train_both(eng, orc, kind, prob, offs, n, rng, epochs, each=None):
    K = len(orc.feat_sizes)
    if kind == "classifier":
        labels = rng.integers(0, K, n).astype(np.int64)
        eng.train(samples=n, problems=prob, solutions=labels, dim_offsets=offs, iterations=epochs, each=each)
        orc.train_replay(prob, offs, n, labels=labels, epochs=epochs)
    else:
        codes = rng.integers(0, 256, (n, (K + 7) // 8), dtype=np.uint8)
        eng.train(samples=n, sentences=prob, codes=codes, dim_offsets=offs, iterations=epochs, each=each)
        orc.train_replay(prob, offs, n, codes=codes, epochs=epochs)


IND_SHAPES = [
    # widths, kind, n, epochs, negative
    ([5, 40, 64], "classifier", 180, 3, 0.5),         # 1..4 words; masked and 16-byte rows
    ([3, 29, 100, 700], "classifier", 150, 2, -1),    # 8-lane rows up to 44 words (64 x 1 tile)
    ([2048, 17, 300], "encoder", 90, 2, -1),          # 64 x 2 tile, 5-byte and 75-byte rows
    ([63, 64, 65, 9, 1], "encoder", 160, 3, -1),      # widths around one 16-byte block
    ([4096, 1000], "classifier", 80, 1, 0.3),         # 64 x 4 tile
]


@pytest.mark.parametrize("widths,kind,n,epochs,negative", IND_SHAPES)
def test_ind_train_bitexact_vs_replay(widths, kind, n, epochs, negative):
    rng = np.random.default_rng(sum(widths) + n)
    spec = max(2.0, max(widths) / 32.0)
    eng, orc = pair(kind, widths, specificity=spec, negative=negative)
    prob, offs = ragged(widths, n, rng, gap=3)
    train_both(eng, orc, kind, prob, offs, n, rng, epochs)
    assert orc.stats[1] > 0 and orc.stats[3] > 0
    assert_same_state(eng, orc, label=str(widths))
    tprob, toffs = ragged(widths, 120, rng)
    assert np.array_equal(eng.predict(tprob, toffs, 120), orc.predict(tprob, toffs, 120))


def test_ind_epoch_callback_matches_grouped_epochs():
    # with a callback every epoch is its own launch; without one epochs run in
    # groups inside one launch: both are the same machine
    widths, n = [21, 130, 8], 200
    rng = np.random.default_rng(3)
    prob, offs = ragged(widths, n, rng)
    labels = rng.integers(0, 3, n).astype(np.int64)
    a, orc = pair("classifier", widths)
    b, _ = pair("classifier", widths)
    seen = []
    a.train(samples=n, problems=prob, solutions=labels, dim_offsets=offs, iterations=10,
            each=lambda e: seen.append(e) or True)
    b.train(samples=n, problems=prob, solutions=labels, dim_offsets=offs, iterations=10)
    orc.train_replay(prob, offs, n, labels=labels, epochs=10)
    assert seen == list(range(1, 11))
    assert_same_state(a, orc, label="per-epoch launches")
    assert_same_state(b, orc, label="grouped epochs")


@pytest.mark.parametrize("kind", ["classifier", "encoder"])
def test_ind_shared_rows_equal_plain_engine(kind):
    # every class reading the same rows makes the plain machine
    F, K, n = 45, 4, 220
    rng = np.random.default_rng(9)
    rows = rows_for(F, n, rng)
    o = dict(classes=K, clauses=24, clause_tolerance=5, clause_maximum=7, specificity=2.5, target=6, seed=5)
    plain = tm.create(kind, dict(o, features=F))
    ind = tm.create(kind, dict(o, individualized=True, feat_offsets=offsets_of([F] * K)))
    offs = np.zeros(K, np.int64)
    if kind == "classifier":
        labels = rng.integers(0, K, n).astype(np.int64)
        plain.train(samples=n, problems=rows, solutions=labels, iterations=3)
        ind.train(samples=n, problems=rows, solutions=labels, dim_offsets=offs, iterations=3)
    else:
        codes = rng.integers(0, 16, (n, 1), dtype=np.uint8)
        plain.train(samples=n, sentences=rows, codes=codes, iterations=3)
        ind.train(samples=n, sentences=rows, codes=codes, dim_offsets=offs, iterations=3)
    assert np.array_equal(ind.checkpoint(), plain.checkpoint())
    assert ind.train_stats() == plain.train_stats()
    test = rows_for(F, 100, rng)
    assert np.array_equal(ind.predict(test, offs, 100), plain.predict(test, 100))


def test_ind_predict_random_actions():
    # predict alone on a random include plane, widths on every tile size
    widths = [1, 7, 33, 300, 5000, 2]
    rng = np.random.default_rng(17)
    eng, orc = pair("encoder", widths, clauses=72, clause_tolerance=9)
    acts = (rng.random(orc.actions.size) < 0.02).astype(np.uint8) * rng.integers(1, 256, orc.actions.size)
    acts = acts.astype(np.uint8)
    eng.restore(acts)
    orc.actions[:] = acts
    prob, offs = ragged(widths, 333, rng, gap=1)
    assert np.array_equal(eng.predict(prob, offs, 333), orc.predict(prob, offs, 333))
    cls, orc_c = pair("classifier", widths, clauses=72, clause_tolerance=9)
    cls.restore(acts)
    orc_c.actions[:] = acts
    assert np.array_equal(cls.predict(prob, offs, 333), orc_c.predict(prob, offs, 333))


def test_ind_persist_layout_and_load(tmp_path):
    widths = [13, 4, 70]
    rng = np.random.default_rng(21)
    eng, orc = pair("classifier", widths, clauses=10, clause_tolerance=5, clause_maximum=7, specificity=3.5,
                    target=9)
    info = eng.info()
    B = [(2 * k + 7) // 8 for k in widths]
    assert info["individualized"] == 1 and info["action_chunks"] == 16 * sum(B)
    assert eng.feature_sizes() == widths
    acts = rng.integers(0, 256, info["action_chunks"], dtype=np.uint8)
    eng.restore(acts)
    blob = eng.persist()
    hdr = struct.pack("<i13Id", 0, 3, 0, 16, 5, 7, 9, 8, 1, 0, 0, 2, 16 * sum(B) * 7, 16 * sum(B), 3.5)
    assert blob[:len(hdr)] == hdr
    assert blob[len(hdr)] == 0
    a0 = len(hdr) + 1
    assert blob[a0:a0 + acts.size] == acts.tobytes()
    assert blob[a0 + acts.size] == 1
    assert np.frombuffer(blob[a0 + acts.size + 1:], np.uint64).tolist() == widths
    loaded = tm.load(blob, True)
    assert loaded.individualized() and loaded.feature_sizes() == widths
    assert np.array_equal(loaded.checkpoint(), acts)
    orc.actions[:] = acts
    prob, offs = ragged(widths, 64, rng)
    assert np.array_equal(loaded.predict(prob, offs, 64), orc.predict(prob, offs, 64))
    path = tmp_path / "ind.bin"
    eng.persist(str(path))
    assert path.read_bytes() == blob
    with pytest.raises(tm.TsetlinError, match="TM_E_STATE"):
        loaded.train(samples=64, problems=prob, solutions=np.zeros(64, np.int64), dim_offsets=offs, iterations=1)
    with pytest.raises(tm.TsetlinError, match="TM_E_FORMAT"):
        tm.load(blob[:-5], True)


def test_ind_restrict():
    widths, n = [11, 50, 3, 90], 160
    rng = np.random.default_rng(31)
    eng, orc = pair("classifier", widths, reusable=True)
    prob, offs = ragged(widths, n, rng)
    train_both(eng, orc, "classifier", prob, offs, n, rng, 2)
    assert_same_state(eng, orc)
    full, full_counts = eng.checkpoint(), eng.counts()
    sl = orc.action_slices()
    sel = [3, 0, 2]
    eng.restrict(sel)
    info = eng.info()
    assert info["classes"] == 3 and eng.feature_sizes() == [widths[i] for i in sel]
    assert np.array_equal(eng.checkpoint(), np.concatenate([full[sl[i]] for i in sel]))
    # the selected classes' counters move with them (7 planes per action byte)
    csl = [slice(7 * x.start, 7 * x.stop) for x in sl]
    assert np.array_equal(eng.counts(), np.concatenate([full_counts[csl[i]] for i in sel]))
    # restrict keeps the machine's negative sampling rate, 1/4 from its
    # original 4 classes (capi.c:752-900 leaves tm->negative alone)
    sub = O.IndMachine(O.params("classifier", classes=3, features=1, clauses=16, clause_tolerance=6,
                                clause_maximum=8, specificity=3.0, target=8, seed=77, negative=0.25),
                       [widths[i] for i in sel])
    sub.actions[:] = eng.checkpoint()
    assert np.array_equal(eng.predict(prob, offs[sel], n), sub.predict(prob, offs[sel], n))
    # training after restrict runs on the new per-class layout
    labels = rng.integers(0, 3, n).astype(np.int64)
    eng.train(samples=n, problems=prob, solutions=labels, dim_offsets=offs[sel], iterations=2)
    sub.train_replay(prob, offs[sel], n, labels=labels, epochs=2)
    assert_same_state(eng, sub, label="after restrict")
    for bad, msg in (([], "selected_dims cannot be empty"), ([0, 3], r"selected_dims\[1\] = 3 out of range \[0, 3\)"),
                     ([1, 2, 1], "duplicate dimension 1 in selected_dims"), ([-1], "out of range")):
        with pytest.raises(tm.TsetlinError, match=msg):
            eng.restrict(bad)
    plain = tm.classifier(classes=2, features=8, clauses=8, clause_tolerance=3, clause_maximum=3, specificity=2)
    with pytest.raises(tm.TsetlinError, match="restrict only supported for individualized TMs"):
        plain.restrict([0])


def test_ind_reconfigure_then_train():
    widths, n = [30, 7], 150
    rng = np.random.default_rng(41)
    eng, _ = pair("encoder", widths, reusable=True)
    prob, offs = ragged(widths, n, rng)
    codes = rng.integers(0, 4, (n, 1), dtype=np.uint8)
    eng.train(samples=n, sentences=prob, codes=codes, dim_offsets=offs, iterations=1)
    eng.reconfigure(clauses=40, clause_tolerance=4, clause_maximum=5, specificity=2.0, target=5, seed=8)
    info = eng.info()
    assert info["clauses"] == 40 and info["action_chunks"] == 40 * (8 + 2)
    eng.train(samples=n, sentences=prob, codes=codes, dim_offsets=offs, iterations=2)
    orc = O.IndMachine(O.params("encoder", classes=2, features=1, clauses=40, clause_tolerance=4, clause_maximum=5,
                                specificity=2.0, target=5, seed=8), widths)
    orc.train_replay(prob, offs, n, codes=codes, epochs=2)
    assert_same_state(eng, orc, label="after reconfigure")


def test_ind_argument_errors():
    widths, n = [12, 5], 40
    rng = np.random.default_rng(51)
    eng, _ = pair("classifier", widths)
    prob, offs = ragged(widths, n, rng)
    labels = np.zeros(n, np.int64)
    with pytest.raises(tm.TsetlinError, match="TM_E_SIZE"):
        eng.train(samples=n, problems=prob[:-1], solutions=labels, dim_offsets=offs, iterations=1)
    with pytest.raises(tm.TsetlinError, match="TM_E_SIZE"):
        eng.predict(prob, offs + 1, n)
    with pytest.raises(tm.TsetlinError, match="negative"):
        eng.predict(prob, np.array([-1, 0], np.int64), n)
    out = np.zeros(n, np.int64)
    assert eng._lib.tm_predict_classifier(eng._handle(), prob.ctypes.data, n, out.ctypes.data) == -1
    assert b"dim_offsets" in eng._lib.tm_last_error()
    assert eng._lib.tm_train_classifier(eng._handle(), n, prob.ctypes.data, labels.ctypes.data, 1,
                                        tm._c.EPOCH_FN(), None) == -1
    assert b"dim_offsets" in eng._lib.tm_last_error()
    with pytest.raises(tm.TsetlinError, match="classes \\+ 1 elements"):
        tm.create("classifier", classes=2, individualized=True, feat_offsets=[0, 4], clauses=8, clause_tolerance=3,
                  clause_maximum=3, specificity=2)
    with pytest.raises(tm.TsetlinError, match="empty feature set"):
        tm.create("classifier", classes=2, individualized=True, feat_offsets=[0, 4, 4], clauses=8,
                  clause_tolerance=3, clause_maximum=3, specificity=2)
    plain = tm.classifier(classes=2, features=8, clauses=8, clause_tolerance=3, clause_maximum=3, specificity=2)
    with pytest.raises(tm.TsetlinError, match="not individualized"):
        plain._predict_ind(prob, offs, n)


def test_ind_zero_samples_and_iterations():
    # setup still runs with no sample or no epoch (capi.c:1315-1327, :1413-1425)
    widths = [6, 19]
    rng = np.random.default_rng(61)
    eng, orc = pair("classifier", widths, reusable=True)
    acts = rng.integers(0, 256, eng.info()["action_chunks"], dtype=np.uint8)
    eng.restore(acts)
    prob, offs = ragged(widths, 10, rng)
    eng.train(samples=0, problems=prob[:0], solutions=np.zeros(0, np.int64), dim_offsets=offs, iterations=3)
    assert not eng.checkpoint().any()
    eng.restore(acts)
    eng.train(samples=10, problems=prob, solutions=np.zeros(10, np.int64), dim_offsets=offs, iterations=0)
    assert not eng.checkpoint().any()
    assert eng.predict(prob, offs, 0).shape == (0,)
    enc, _ = pair("encoder", widths)
    assert enc.predict(prob, offs, 0).shape == (0, 1)


# ---------------------------------------------------------------------------
# Device-resident individualized data: the packer and the _ind_dev calls.

def test_bits_to_cvec_ind_dev_equals_host_packer():
    """tm_bits_to_cvec_ind_dev against the host restatement: ragged widths
    (rows of 1 to 75 bytes, so most class blocks start off a word boundary),
    an odd sample count, duplicates, out-of-range entries (counted, skipped),
    a class without entries, and an output tensor whose start is not word
    aligned."""
    import torch
    rng = np.random.default_rng(21)
    n, union = 37, 500
    x = rng.random((n, union)) < 0.15
    widths = [1, 300, 5, 64, 9, 17]
    ids = [rng.choice(union, k, replace=False) for k in widths]
    fo = np.concatenate([[0], np.cumsum(widths)]).astype(np.int64)
    s, f = np.nonzero(x)
    sb = np.concatenate([s * union + f, (s * union + f)[:40]])
    ind, io = D.bits_individualize(sb, fo, np.concatenate(ids), union)
    io = io.copy()
    # class 3 gets no entries; class 1 gets two out-of-range ones
    ind = np.concatenate([ind[:io[1]], ind[io[1]:io[2]], [-3, n * widths[1]], ind[io[2]:io[3]], ind[io[4]:]])
    sizes = np.diff(io)
    sizes[1] += 2
    sizes[3] = 0
    io = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    want, wdim = D.bits_to_cvec_ind(ind, io, fo, n)
    total = int(wdim[-1])
    big = torch.full((total + 5,), 0xA5, dtype=torch.uint8, device="cuda")
    out = big[1:1 + total]
    got, dim, bad = D.bits_to_cvec_ind_dev(torch.from_numpy(ind).cuda(), io, fo, n, out=out)
    assert bad == 2
    assert np.array_equal(dim, wdim)
    assert np.array_equal(got.cpu().numpy(), want)
    b = big.cpu().numpy()
    assert b[0] == 0xA5 and (b[1 + total:] == 0xA5).all()     # nothing written outside the output


@pytest.mark.parametrize("kind", ["classifier", "encoder"])
def test_ind_dev_calls_equal_host_calls(kind):
    """Pack on the device, then train and predict with the problems in device
    memory (the _ind_dev calls): the state and the predictions equal the host
    calls' on the host-packed bytes, bit for bit."""
    import torch
    rng = np.random.default_rng(33 if kind == "classifier" else 34)
    n, union = 140, 300
    x = rng.random((n, union)) < 0.25
    widths = [40, 5, 120, 64] if kind == "classifier" else [40, 5, 120, 64, 9, 200, 31, 2, 77]
    K = len(widths)
    ids = [rng.choice(union, k, replace=False) for k in widths]
    fo = np.concatenate([[0], np.cumsum(widths)]).astype(np.int64)
    s, f = np.nonzero(x)
    ind, io = D.bits_individualize(s * union + f, fo, np.concatenate(ids), union)
    hb, hdim = D.bits_to_cvec_ind(ind, io, fo, n)
    db, ddim, bad = D.bits_to_cvec_ind_dev(torch.from_numpy(ind).cuda(), io, fo, n)
    assert bad == 0 and np.array_equal(ddim, hdim) and np.array_equal(db.cpu().numpy(), hb)
    o = dict(classes=K, individualized=True, feat_offsets=fo, clauses=16, clause_tolerance=6, clause_maximum=8,
             specificity=3.0, target=8, seed=5, reusable=True)
    host, dev = tm.create(kind, dict(o)), tm.create(kind, dict(o))
    if kind == "classifier":
        y = rng.integers(0, K, n).astype(np.int64)
        host.train(samples=n, problems=hb, solutions=y, dim_offsets=hdim[:K], iterations=3)
        dev.train(samples=n, problems=db, solutions=torch.from_numpy(y).cuda(), dim_offsets=ddim[:K], iterations=3)
    else:
        codes = rng.integers(0, 256, (n, (K + 7) // 8), dtype=np.uint8)
        host.train(samples=n, sentences=hb, codes=codes, dim_offsets=hdim[:K], iterations=3)
        dev.train(samples=n, sentences=db, codes=torch.from_numpy(codes).cuda(), dim_offsets=ddim[:K], iterations=3)
    assert np.array_equal(host.checkpoint(), dev.checkpoint())
    assert np.array_equal(host.counts(), dev.counts())
    want = host.predict(hb, hdim[:K], n)
    got = dev.predict(db, ddim[:K], n)
    assert np.array_equal(got.cpu().numpy(), want)
    # the byte count the library checks is the tensor's: a short one is refused
    with pytest.raises(tm.TsetlinError):
        dev.predict(db[:int(ddim[-1]) - 1], ddim[:K], n)
