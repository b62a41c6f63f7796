"""Workload 3 (plain encoder) at its full machine shape: 256 hidden bits x
1024 clauses per bit over 8192 visible bits (16384 literals, 512 words per
row, 32,768 chunks, 4 GiB of automata state).

Predict is bit-exact against the oracle on the same include plane; training
is bit-exact against the oracle's replay of the same stream; predict over
all 500k samples is batch-independent and bit-exact on a sample of rows.
"""
import numpy as np
import pytest

import santoku.tsetlin as tm
from oracle import oracle as O
from santoku.tsetlin import data as D
from state import assert_same_state

pytestmark = pytest.mark.gpu

C3 = dict(classes=256, features=8192, clauses=1024, clause_tolerance=64, clause_maximum=64, specificity=20.0,
          target=128)


def random_plane(info, rng, density):
    """A random include plane with some empty clauses."""
    B = info["input_chunks"]
    nclauses = info["classes"] * info["clauses"]
    acts = np.zeros(nclauses * B, np.uint8)
    idx = rng.integers(0, acts.size * 8, int(acts.size * 8 * density))
    np.bitwise_or.at(acts, idx >> 3, (1 << (idx & 7)).astype(np.uint8))
    acts = acts.reshape(nclauses, B)
    acts[rng.random(nclauses) < 0.05] = 0
    return acts.reshape(-1)


@pytest.fixture(scope="module")
def c3_machine():
    eng = tm.encoder(dict(C3, reusable=True, seed=31))
    orc = O.Machine(O.params("encoder", **dict(C3, seed=31)))
    yield eng, orc
    eng.destroy()


def test_c3_shape_predict_bitexact(c3_machine):
    eng, orc = c3_machine
    rng = np.random.default_rng(3)
    acts = random_plane(eng.info(), rng, 0.0015)
    eng.restore(acts)
    orc.actions[:] = acts
    d = D.projection_codes_task(5, n=48)
    got = eng.predict(d["rows"], 48)
    want = orc.predict(d["rows"])
    assert got.shape == (48, 32)
    assert np.array_equal(got, want)
    # neither all-zero nor all-one codes: the sums really vary
    assert 0.05 < np.unpackbits(got).mean() < 0.95


def test_c3_shape_train_bitexact_vs_replay(c3_machine):
    eng, orc = c3_machine
    d = D.projection_codes_task(6, n=12)
    eng.train(samples=12, sentences=d["rows"], codes=d["codes"], iterations=1)
    orc.counts[:] = 0
    orc.actions[:] = 0
    orc.train_replay(d["rows"], codes=d["codes"], epochs=1)
    st = eng.train_stats()
    assert st["steps"] == 12 * 256 * 128 == int(orc.stats[0])
    assert st["draws"] > 0 and st["type1_false"] > 0
    assert_same_state(eng, orc, label="workload-3 shape, 12 samples")


def test_c3_full_size_predict_batch_independent(c3_machine):
    """All 500k samples in one call: equal to the same rows predicted in
    slices, and bit-exact with the oracle on 32 rows spread over the set."""
    import torch
    eng, orc = c3_machine
    rng = np.random.default_rng(9)
    acts = random_plane(eng.info(), rng, 0.0015)
    eng.restore(acts)
    orc.actions[:] = acts
    n = 500_000
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    rows = torch.empty((n, 2048), dtype=torch.uint8, device=dev)
    weights = (2 ** torch.arange(8, device=dev, dtype=torch.int32)).view(1, 1, 8)
    for s in range(0, n, 50_000):
        x = torch.rand((50_000, 8192), generator=g, device=dev) < 0.1
        lits = torch.cat([x, ~x], dim=1).view(50_000, 2048, 8).to(torch.int32)
        rows[s:s + 50_000] = (lits * weights).sum(2).to(torch.uint8)
    full = eng.predict(rows, n)
    assert full.shape == (n, 32)
    for s, e in ((0, 1000), (123_456, 124_001), (n - 777, n)):
        assert torch.equal(eng.predict(rows[s:e].contiguous(), e - s), full[s:e])
    pick = np.sort(rng.choice(n, 32, replace=False))
    host = rows[torch.from_numpy(pick).to(dev)].cpu().numpy()
    assert np.array_equal(full[torch.from_numpy(pick).to(dev)].cpu().numpy(), orc.predict(host))


def test_c3_shape_train_longer_vs_chunk_replay(c3_machine):
    # 2048 samples x 2 epochs at the full machine shape; the oracle replays a
    # few chunks (chunks are independent), first and last hidden bits included
    eng, orc = c3_machine
    d = D.projection_codes_task(8, n=2048)
    eng.train(samples=2048, sentences=d["rows"], codes=d["codes"], iterations=2)
    chunks = [0, 127, 128 * 100 + 5, 128 * 200 + 64, 128 * 255 + 127]
    orc.actions[:] = 0
    orc.train_replay_chunks(d["rows"], chunks, codes=d["codes"], epochs=2)
    assert orc.stats[1] > 0 and orc.stats[3] > 0
    assert_same_state(eng, orc, chunks=chunks, label="workload-3 shape, 2048 samples x 2 epochs")
