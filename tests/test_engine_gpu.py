"""GPU parity tests: the HIP engine (through the C ABI) against the CPU oracle.

Predict must match bit for bit on the same include plane; training must match
the oracle's replay of the same RNG stream bit for bit.  Shapes cover every
train tile (lanes per clause 8/16/32/64, words per lane 1..16), tail bytes in
both of the reference's padding regimes (masked byte path and unmasked 128-bit
path), rows whose length is not a multiple of 4 bytes, and the encoder.
"""
import ctypes
import struct

import numpy as np
import pytest

import santoku.tsetlin as tm
from oracle import oracle as O
from santoku.tsetlin import data as D
from state import assert_same_state, count_mask

pytestmark = pytest.mark.gpu


def opts(**kw):
    o = dict(classes=3, features=8, clauses=16, clause_tolerance=4, clause_maximum=4, specificity=2.0)
    o.update(kw)
    return o


def pair(kind="classifier", **kw):
    o = opts(**kw)
    eng = tm.create(kind, dict(o, reusable=True))
    po = {k: v for k, v in o.items() if k not in ("state", "reusable")}
    if kind == "encoder":
        po = dict(po)
    orc = O.Machine(O.params(kind, **po))
    return eng, orc


def rows_for(features, n, rng, density=0.3):
    return D.bits_to_cvec(rng.random((n, features)) < density)


# ---------------------------------------------------------------------------
# predict

PREDICT_SHAPES = [
    # features, classes, clauses, kind
    (5, 3, 8, "classifier"),
    (37, 9, 24, "encoder"),
    (64, 2, 40, "classifier"),
    (100, 5, 16, "classifier"),
    (256, 4, 256, "classifier"),
    (4096, 20, 64, "classifier"),
    (8192, 11, 16, "encoder"),
]


@pytest.mark.parametrize("features,classes,clauses,kind", PREDICT_SHAPES)
def test_predict_bitexact_random_actions(features, classes, clauses, kind):
    rng = np.random.default_rng(features + classes)
    eng, orc = pair(kind, features=features, classes=classes, clauses=clauses, clause_tolerance=5)
    dens = min(0.2, 12.0 / (2 * features))
    acts = np.packbits((rng.random(orc.actions.size * 8) < dens).astype(np.uint8), bitorder="little")
    # a few empty clauses (vote +tolerance)
    B = orc.p.input_chunks
    acts = acts.reshape(-1, B)
    acts[rng.random(acts.shape[0]) < 0.1] = 0
    acts = acts.reshape(-1)
    eng.restore(acts)
    orc.actions[:] = acts
    n = 300
    rows = rows_for(features, n, rng, 0.5)
    got = eng.predict(rows, n)
    want = orc.predict(rows)
    assert np.array_equal(got, want)
    assert np.array_equal(eng.checkpoint(), acts)


def test_predict_device_pointers():
    import torch
    rng = np.random.default_rng(1)
    eng, orc = pair(features=64, classes=4, clauses=32)
    acts = np.packbits((rng.random(orc.actions.size * 8) < 0.05).astype(np.uint8), bitorder="little")
    eng.restore(acts)
    orc.actions[:] = acts
    rows = rows_for(64, 777, rng)
    got = eng.predict(torch.from_numpy(rows).cuda(), 777).cpu().numpy()
    assert np.array_equal(got, orc.predict(rows))


# ---------------------------------------------------------------------------
# training

TRAIN_SHAPES = [
    # features, classes, clauses, kind, include_bits, n, epochs
    (5, 3, 16, "classifier", 1, 200, 3),        # 1 word, masked tail byte
    (63, 2, 16, "classifier", 1, 150, 3),       # 16-byte row: unmasked tail bits
    (100, 3, 8, "classifier", 2, 150, 2),       # 25-byte rows (repacked), include_bits 2
    (256, 4, 32, "classifier", 1, 400, 2),      # 16 lanes
    (512, 2, 16, "encoder", 1, 120, 2),         # 32 lanes
    (1024, 2, 8, "classifier", 1, 100, 2),      # 64 lanes x 1 word
    (2048, 3, 8, "encoder", 1, 80, 2),          # 64 x 2
    (4096, 5, 8, "classifier", 1, 120, 2),      # 64 x 4
    (8192, 2, 8, "encoder", 1, 60, 1),          # 64 x 8
    (16384, 2, 8, "classifier", 1, 40, 1),      # 64 x 16
]


@pytest.mark.parametrize("features,classes,clauses,kind,include_bits,n,epochs", TRAIN_SHAPES)
def test_train_bitexact_vs_replay(features, classes, clauses, kind, include_bits, n, epochs):
    rng = np.random.default_rng(features * 7 + classes)
    spec = max(2.0, features / 16.0)
    eng, orc = pair(kind, features=features, classes=classes, clauses=clauses, clause_tolerance=6,
                    clause_maximum=8, specificity=spec, target=8, include_bits=include_bits, seed=1234)
    rows = rows_for(features, n, rng, 0.3)
    if kind == "classifier":
        labels = rng.integers(0, classes, n).astype(np.int64)
        eng.train(samples=n, problems=rows, solutions=labels, iterations=epochs)
        orc.train_replay(rows, labels=labels, epochs=epochs)
    else:
        codes = rng.integers(0, 256, (n, (classes + 7) // 8), dtype=np.uint8)
        eng.train(samples=n, sentences=rows, codes=codes, iterations=epochs)
        orc.train_replay(rows, codes=codes, epochs=epochs)
    assert orc.stats[1] > 0 and orc.stats[3] > 0
    assert_same_state(eng, orc)
    test_rows = rows_for(features, 200, rng, 0.3)
    assert np.array_equal(eng.predict(test_rows, 200), orc.predict(test_rows))


def test_train_workload1_bitexact():
    d = D.xor_and_task(2)
    o = dict(classes=4, features=256, clauses=256, clause_tolerance=64, clause_maximum=64, specificity=8.0,
             target=64, seed=99)
    eng = tm.classifier(dict(o, reusable=True))
    orc = O.Machine(O.params("classifier", **o))
    eng.train(samples=8192, problems=d["train_rows"], solutions=d["train_y"], iterations=3)
    orc.train_replay(d["train_rows"], labels=d["train_y"], epochs=3)
    assert_same_state(eng, orc)
    assert np.array_equal(eng.predict(d["test_rows"], 2048), orc.predict(d["test_rows"]))


@pytest.mark.parametrize("features,kind,state", [(6, "classifier", 8), (40, "encoder", 8), (300, "classifier", 8),
                                                 (24, "classifier", 3), (1024, "classifier", 8),
                                                 (16, "classifier", 8), (1024, "encoder", 8),
                                                 (8192, "encoder", 8)])
def test_repeat_hits_bitexact(features, kind, state):
    """Specificity 1: an event draws as many random literals as there are
    literals, so literals are hit two, three and four or more times by one
    event (Poisson(1)), often while their counters are at 1 or 2.  Covers
    the clamped multi-level subtract of the random-literal draws, its
    saturation at zero, the rounds left for a fourth hit, and the
    round-per-hit path of the four-workgroup tile (1024 features).  32 and
    2048 literals (16 and 1024 features outside the four-workgroup tile) take
    the power-of-two draw shift, with and without tail masks.  8192 features
    is the full-wave 8-word encoder tile with its three hit levels, at 16,384
    draws per event (256 per lane)."""
    rng = np.random.default_rng(features + state)
    n = 160
    o = dict(classes=4, features=features, clauses=16, clause_tolerance=6, clause_maximum=8, specificity=1.0,
             target=6, seed=4321, state=state)
    eng = tm.create(kind, dict(o, reusable=True))
    orc = O.Machine(O.params(kind, **o))
    rows = rows_for(features, n, rng, 0.3)
    if kind == "classifier":
        labels = rng.integers(0, 4, n).astype(np.int64)
        eng.train(samples=n, problems=rows, solutions=labels, iterations=3)
        orc.train_replay(rows, labels=labels, epochs=3)
    else:
        codes = rng.integers(0, 256, (n, 1), dtype=np.uint8)
        eng.train(samples=n, sentences=rows, codes=codes, iterations=3)
        orc.train_replay(rows, codes=codes, epochs=3)
    # oracle counters: steps, Type I true, Type II, Type I false, draws, skips
    assert orc.stats[3] > 0 and orc.stats[4] == orc.stats[3] * 2 * features
    assert_same_state(eng, orc)
    st = eng.train_stats()
    assert st["type1_false"] == int(orc.stats[3]) and st["draws"] == int(orc.stats[4])


def test_negative_sampling_variants():
    rng = np.random.default_rng(5)
    rows = rows_for(40, 300, rng)
    labels = rng.integers(0, 4, 300).astype(np.int64)
    labels[::17] = -1          # unsigned cast: never the target class (capi.c:1163)
    for neg in (0.0, 0.3, 1.0):
        eng, orc = pair(features=40, classes=4, clauses=16, negative=neg, seed=3, target=6)
        eng.train(samples=300, problems=rows, solutions=labels, iterations=2)
        orc.train_replay(rows, labels=labels, epochs=2)
        assert_same_state(eng, orc, label="negative %s" % neg)


QUEUE_SHAPES = [
    # features, classes, clauses, negative, n -- tiles 8x1, 16x2, 64x1, 64x4, 64x16
    (5, 3, 16, 0.5, 300),
    (200, 4, 24, 0.0, 500),
    (1024, 3, 16, 1.0, 400),
    (4096, 5, 8, 0.1, 700),
    (16384, 2, 8, 0.3, 90),
]


@pytest.mark.parametrize("features,classes,clauses,neg,n", QUEUE_SHAPES)
def test_step_queue_matches_scan(features, classes, clauses, neg, n, monkeypatch):
    """A classifier trains from per-epoch step queues; the in-kernel scan of
    the epoch order (TSETLIN_AMD_QUEUE=0) must give the same machine and the
    same work counters, and both the oracle's replay."""
    rng = np.random.default_rng(features + n)
    rows = rows_for(features, n, rng, 0.3)
    labels = rng.integers(0, classes, n).astype(np.int64)
    labels[::13] = -1
    o = dict(features=features, classes=classes, clauses=clauses, negative=neg, seed=21, target=6,
             clause_tolerance=5, clause_maximum=7, specificity=max(2.0, features / 16.0))
    got = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("TSETLIN_AMD_QUEUE", mode)
        eng, orc = pair(**o)
        eng.train(samples=n, problems=rows, solutions=labels, iterations=3)
        got[mode] = (eng.checkpoint(), eng.train_stats(), eng.queue_timing(), eng.counts())
    assert np.array_equal(got["1"][0], got["0"][0])
    assert np.array_equal(got["1"][3], got["0"][3])
    assert got["1"][1] == got["0"][1]
    assert got["1"][2] > 0 and got["0"][2] == 0
    orc.train_replay(rows, labels=labels, epochs=3)
    assert np.array_equal(got["1"][0], orc.actions)
    mask = count_mask(orc)
    assert np.array_equal(got["1"][3] & mask, orc.counts & mask)
    # the event counters (kept in LDS on the four-workgroup tile) against the
    # oracle's: steps, Type I true, Type II, Type I false, draws
    st = got["1"][1]
    assert (st["steps"], st["type1_true"], st["type2_true"], st["type1_false"], st["draws"]) == \
        tuple(int(v) for v in orc.stats[:5])


@pytest.mark.parametrize("tolerance", [8191, 8192, 9000])
def test_vote_width_at_large_tolerance(tolerance):
    """Tolerances (and clause maxima) past the 16-bit range of four summed
    votes, on the four-workgroup tile (2048 literals, a full 64-lane tile):
    the votes, the chunk clamp and the fire test must keep the replay's
    machine."""
    rng = np.random.default_rng(tolerance)
    n = 300
    rows = rows_for(1024, n, rng, 0.3)
    labels = rng.integers(0, 3, n).astype(np.int64)
    o = dict(features=1024, classes=3, clauses=16, clause_tolerance=tolerance, clause_maximum=tolerance,
             specificity=8.0, target=tolerance // 2, seed=17)
    eng, orc = pair(**o)
    eng.train(samples=n, problems=rows, solutions=labels, iterations=2)
    orc.train_replay(rows, labels=labels, epochs=2)
    assert_same_state(eng, orc)
    assert eng.train_stats()["steps"] == int(orc.stats[0])


@pytest.mark.parametrize("kind", ["classifier", "encoder"])
def test_every_tile_gives_the_same_machine(kind, monkeypatch):
    """32 words per row: tiles 16x2, 32x1, 32x2, 64x1 and 64x2 (lanes x words)
    all cover it; each must train the replay's machine bit for bit."""
    rng = np.random.default_rng(77)
    n = 300
    rows = rows_for(512, n, rng, 0.3)
    o = dict(features=512, classes=3, clauses=16, clause_tolerance=6, clause_maximum=8, specificity=16.0,
             target=8, seed=55)
    orc = None
    for tile in ("16x2", "32x1", "32x2", "64x1", "64x2"):
        monkeypatch.setenv("TSETLIN_AMD_TILE", tile)
        eng, orc = pair(kind, **o)
        if kind == "classifier":
            labels = np.random.default_rng(5).integers(0, 3, n).astype(np.int64)
            eng.train(samples=n, problems=rows, solutions=labels, iterations=2)
        else:
            codes = np.random.default_rng(5).integers(0, 8, (n, 1), dtype=np.uint8)
            eng.train(samples=n, sentences=rows, codes=codes, iterations=2)
        if tile == "16x2":
            if kind == "classifier":
                orc.train_replay(rows, labels=labels, epochs=2)
            else:
                orc.train_replay(rows, codes=codes, epochs=2)
            ref = orc
        assert_same_state(eng, ref, label=tile)


def test_each_callback_epochs_and_stop():
    rng = np.random.default_rng(6)
    rows = rows_for(64, 200, rng)
    labels = rng.integers(0, 3, 200).astype(np.int64)
    eng, orc = pair(features=64, classes=3, clauses=16, seed=8)
    seen = []
    preds = []

    def each(epoch):
        seen.append(epoch)
        preds.append(eng.predict(rows, 200))
        return epoch < 3          # stop after epoch 3

    eng.train(samples=200, problems=rows, solutions=labels, iterations=10, each=each)
    assert seen == [1, 2, 3]
    orc.train_replay(rows, labels=labels, epochs=3)
    assert_same_state(eng, orc)
    assert np.array_equal(preds[-1], orc.predict(rows))
    # the callback path (one launch per epoch) equals one launch over all epochs
    eng2, _ = pair(features=64, classes=3, clauses=16, seed=8)
    eng2.train(samples=200, problems=rows, solutions=labels, iterations=3)
    assert_same_state(eng2, orc)


def test_each_callback_error_stops_training(capsys):
    rng = np.random.default_rng(7)
    rows = rows_for(16, 50, rng)
    labels = rng.integers(0, 2, 50).astype(np.int64)
    eng, _ = pair(features=16, classes=2, clauses=8)
    seen = []

    def each(epoch):
        seen.append(epoch)
        raise ValueError("boom")

    eng.train(samples=50, problems=rows, solutions=labels, iterations=5, each=each)
    assert seen == [1]
    assert "boom" in capsys.readouterr().err


def test_train_restarts_from_setup_each_call():
    # capi.c:1136-1146: every train() call re-runs setup
    rng = np.random.default_rng(9)
    rows = rows_for(32, 100, rng)
    labels = rng.integers(0, 2, 100).astype(np.int64)
    eng, orc = pair(features=32, classes=2, clauses=8, seed=4)
    eng.train(samples=100, problems=rows, solutions=labels, iterations=2)
    eng.train(samples=100, problems=rows, solutions=labels, iterations=2)
    orc.train_replay(rows, labels=labels, epochs=2)
    assert_same_state(eng, orc)


def test_device_resident_training():
    import torch
    rng = np.random.default_rng(10)
    rows = rows_for(256, 500, rng)
    labels = rng.integers(0, 4, 500).astype(np.int64)
    eng, orc = pair(features=256, classes=4, clauses=16, seed=2)
    eng.train(samples=500, problems=torch.from_numpy(rows).cuda(), solutions=torch.from_numpy(labels).cuda(),
              iterations=2)
    orc.train_replay(rows, labels=labels, epochs=2)
    assert_same_state(eng, orc)


# ---------------------------------------------------------------------------
# persistence, checkpoint, reconfigure, errors (capi.c:1509-1772)

def test_persist_layout_and_roundtrip(tmp_path):
    rng = np.random.default_rng(11)
    eng = tm.classifier(classes=3, features=13, clauses=10, clause_tolerance=5, clause_maximum=7,
                        specificity=3.5, target=9)
    info = eng.info()
    acts = rng.integers(0, 256, info["action_chunks"], dtype=np.uint8)
    eng.restore(acts)
    blob = eng.persist()
    hdr = struct.pack("<i13Id", 0, 3, 13, 16, 5, 7, 9, 8, 1, 26, 4, 2, 3 * 16 * 7 * 4, 3 * 16 * 4, 3.5)
    assert blob[:len(hdr)] == hdr
    assert blob[len(hdr)] == 0x03                     # tail mask of 26 bits
    assert blob[len(hdr) + 1:len(hdr) + 1 + len(acts)] == acts.tobytes()
    assert blob[-1] == 0 and len(blob) == 65 + len(acts) + 1
    loaded = tm.load(blob, True)
    assert loaded.type() == "classifier"
    rows = rows_for(13, 64, rng)
    assert np.array_equal(loaded.predict(rows, 64), eng.predict(rows, 64))
    path = tmp_path / "m.bin"
    eng.persist(str(path))
    assert path.read_bytes() == blob
    assert np.array_equal(tm.load(str(path)).checkpoint(), acts)
    with pytest.raises(tm.TsetlinError, match="TM_E_STATE"):
        loaded.train(samples=64, problems=rows, solutions=np.zeros(64, np.int64), iterations=1)
    with pytest.raises(tm.TsetlinError, match="TM_E_FORMAT"):
        tm.load(blob[:40], True)
    with pytest.raises(tm.TsetlinError, match="TM_E_FORMAT"):
        tm.load(blob[:-3], True)
    bad = bytearray(blob)
    bad[0] = 7
    with pytest.raises(tm.TsetlinError, match="unexpected tsetlin machine type"):
        tm.load(bytes(bad), True)


def test_encoder_persist_type():
    eng = tm.encoder(hidden=10, visible=20, clauses=8, clause_tolerance=3, clause_maximum=3, specificity=2)
    blob = eng.persist()
    assert struct.unpack("<i", blob[:4])[0] == 1
    assert tm.load(blob, True).type() == "encoder"


def test_checkpoint_restore_and_errors():
    eng = tm.classifier(classes=2, features=8, clauses=8, clause_tolerance=3, clause_maximum=3, specificity=2)
    ck = eng.checkpoint()
    assert ck.size == 2 * 8 * 2
    with pytest.raises(tm.TsetlinError, match="checkpoint size mismatch"):
        eng.restore(ck[:-1])
    with pytest.raises(tm.TsetlinError, match="reusable"):
        eng.reconfigure(clauses=16, clause_tolerance=2, clause_maximum=2, specificity=2)
    with pytest.raises(tm.TsetlinError, match="classes must be greater"):
        tm.classifier(classes=0, features=8, clauses=8, clause_tolerance=3, clause_maximum=3, specificity=2)
    with pytest.raises(tm.TsetlinError, match="clause_tolerance"):
        tm.classifier(classes=2, features=8, clauses=8, clause_tolerance=0, clause_maximum=3, specificity=2)
    enc = tm.encoder(hidden=2, visible=8, clauses=8, clause_tolerance=3, clause_maximum=3, specificity=2)
    out = np.zeros(1, np.int64)
    rc = enc._lib.tm_predict_classifier(enc._handle(), np.zeros(2, np.uint8).ctypes.data, 1, out.ctypes.data)
    assert rc == -1 and b"not a" in enc._lib.tm_last_error()
    # a non-reusable machine releases its counters after training (capi.c:1187-1188)
    rows = np.zeros((4, 2), np.uint8)
    eng.train(samples=4, problems=rows, solutions=np.zeros(4, np.int64), iterations=1)
    with pytest.raises(tm.TsetlinError, match="TM_E_STATE"):
        eng.train(samples=4, problems=rows, solutions=np.zeros(4, np.int64), iterations=1)


def test_reconfigure_matches_fresh_machine():
    rng = np.random.default_rng(12)
    rows = rows_for(48, 160, rng)
    labels = rng.integers(0, 3, 160).astype(np.int64)
    eng = tm.classifier(classes=3, features=48, clauses=8, clause_tolerance=3, clause_maximum=3,
                        specificity=2, reusable=True, seed=5)
    eng.train(samples=160, problems=rows, solutions=labels, iterations=1)
    eng.reconfigure(clauses=20, clause_tolerance=4, clause_maximum=6, specificity=3.0, target=0.25, seed=6)
    info = eng.info()
    assert info["clauses"] == 24 and info["target"] == 6 and info["trained"] == 0
    assert (eng.checkpoint() == 0).all()
    eng.train(samples=160, problems=rows, solutions=labels, iterations=2)
    orc = O.Machine(O.params("classifier", classes=3, features=48, clauses=20, clause_tolerance=4,
                             clause_maximum=6, specificity=3.0, target=0.25, seed=6))
    orc.train_replay(rows, labels=labels, epochs=2)
    assert_same_state(eng, orc)


def test_zero_iterations_runs_setup():
    eng = tm.classifier(classes=2, features=8, clauses=8, clause_tolerance=3, clause_maximum=3, specificity=2,
                        reusable=True)
    eng.restore(np.full(32, 0xFF, np.uint8))
    eng.train(samples=0, problems=np.zeros(0, np.uint8), solutions=np.zeros(0, np.int64), iterations=0)
    assert (eng.checkpoint() == 0).all()


@pytest.mark.parametrize("kind,target", [("classifier", 40000), ("classifier", 3_000_000),
                                         ("classifier", (1 << 30) - 1), ("encoder", 123_457)])
def test_fire_test_at_large_targets(kind, target):
    """Targets past 2^15, up to the 2^30 - 1 limit: every clause's feedback
    draw is 32 bits and the fire test mulhi(u, 2T) < T -+ v stays exact, on the
    four-workgroup queue tile (classifier) and the scan tile (encoder)."""
    rng = np.random.default_rng(target % 1000)
    n = 200
    rows = rows_for(1024, n, rng, 0.3)
    o = dict(features=1024, classes=3, clauses=16, clause_tolerance=64, clause_maximum=64, specificity=16.0,
             target=target, seed=29)
    eng, orc = pair(kind, **o)
    if kind == "classifier":
        labels = rng.integers(0, 3, n).astype(np.int64)
        eng.train(samples=n, problems=rows, solutions=labels, iterations=2)
        orc.train_replay(rows, labels=labels, epochs=2)
    else:
        codes = rng.integers(0, 8, (n, 1), dtype=np.uint8)
        eng.train(samples=n, sentences=rows, codes=codes, iterations=2)
        orc.train_replay(rows, codes=codes, epochs=2)
    assert eng.info()["target"] == target
    assert orc.stats[1] > 0 and orc.stats[3] > 0
    assert_same_state(eng, orc, label="target %d" % target)
