"""Pins of the oracle's individualized machine (capi.c:258-402, :528-626,
:997-1086, :1284-1482), CPU only.

An individualized machine is per-class one-class machines over per-class
feature sets.  The pins:
- with every class reading the same rows (equal widths, all dim_offsets 0) it
  must be the plain machine, byte for byte;
- a class's state depends on its own rows and width only, not on the other
  classes' (a wrong per-class offset breaks this);
- a class's sums equal a plain one-class machine's on its own rows.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from santoku.tsetlin import data as D


def rows_for(features, n, rng, density=0.3):
    return D.bits_to_cvec(rng.random((n, features)) < density)


def base(kind, classes, features, **kw):
    o = dict(classes=classes, features=features, clauses=16, clause_tolerance=5, clause_maximum=7,
             specificity=3.0, target=6, seed=13)
    o.update(kw)
    return O.params(kind, **o)


@pytest.mark.parametrize("kind", ["classifier", "encoder"])
def test_shared_rows_equal_plain_machine(kind):
    rng = np.random.default_rng(1)
    K, F, n = 3, 37, 160
    rows = rows_for(F, n, rng)
    p = base(kind, K, F, negative=0.4)
    plain = O.Machine(p)
    ind = O.IndMachine(p, [F] * K)
    offs = np.zeros(K, np.int64)
    if kind == "classifier":
        labels = rng.integers(0, K, n).astype(np.int64)
        plain.train_replay(rows, labels=labels, epochs=3)
        ind.train_replay(rows, offs, n, labels=labels, epochs=3)
    else:
        codes = rng.integers(0, 8, (n, 1), dtype=np.uint8)
        plain.train_replay(rows, codes=codes, epochs=3)
        ind.train_replay(rows, offs, n, codes=codes, epochs=3)
    assert plain.stats[1] > 0 and plain.stats[3] > 0
    assert np.array_equal(ind.actions, plain.actions)
    assert np.array_equal(ind.counts, plain.counts)
    assert np.array_equal(ind.stats, plain.stats)
    test = rows_for(F, 90, rng)
    assert np.array_equal(ind.predict(test, offs, 90), plain.predict(test))


def ragged_problems(widths, n, rng, blocks=None):
    """Concatenated per-class row blocks; `blocks` overrides some classes."""
    parts, offs, o = [], [], 0
    for h, k in enumerate(widths):
        b = blocks[h] if blocks and h in blocks else rows_for(k, n, rng)
        parts.append(b.reshape(-1))
        offs.append(o)
        o += b.size
    return np.concatenate(parts), np.array(offs, np.int64)


@pytest.mark.parametrize("kind", ["classifier", "encoder"])
def test_class_state_depends_on_its_own_rows_only(kind):
    rng = np.random.default_rng(2)
    n = 140
    keep = {0: rows_for(5, n, rng), 2: rows_for(64, n, rng)}
    w1, w2 = [5, 40, 64], [5, 17, 64]
    prob1, off1 = ragged_problems(w1, n, rng, keep)
    prob2, off2 = ragged_problems(w2, n, rng, keep)
    p = base(kind, 3, 1, negative=0.5)
    m1, m2 = O.IndMachine(p, w1), O.IndMachine(p, w2)
    if kind == "classifier":
        labels = rng.integers(0, 3, n).astype(np.int64)
        m1.train_replay(prob1, off1, n, labels=labels, epochs=2)
        m2.train_replay(prob2, off2, n, labels=labels, epochs=2)
    else:
        codes = rng.integers(0, 8, (n, 1), dtype=np.uint8)
        m1.train_replay(prob1, off1, n, codes=codes, epochs=2)
        m2.train_replay(prob2, off2, n, codes=codes, epochs=2)
    s1, s2 = m1.action_slices(), m2.action_slices()
    assert np.array_equal(m1.actions[s1[0]], m2.actions[s2[0]])
    assert np.array_equal(m1.actions[s1[2]], m2.actions[s2[2]])
    assert m1.actions[s1[0]].any() and m1.actions[s1[2]].any()
    assert m1.actions[s1[1]].size != m2.actions[s2[1]].size


def test_class_sums_equal_one_class_machines():
    rng = np.random.default_rng(3)
    widths, n = [3, 29, 64, 100], 70
    prob, offs = ragged_problems(widths, n, rng)
    p = base("classifier", 4, 1)
    ind = O.IndMachine(p, widths)
    ind.actions[:] = (rng.random(ind.actions.size) < 0.05).astype(np.uint8) * rng.integers(1, 256, ind.actions.size)
    sums = ind.class_sums(prob, offs, n)
    for h, (k, sl) in enumerate(zip(widths, ind.action_slices())):
        one = O.Machine(base("classifier", 1, k))
        one.actions[:] = ind.actions[sl]
        B = (2 * k + 7) // 8
        rows_h = prob[offs[h]:offs[h] + n * B].reshape(n, B)
        assert np.array_equal(one.class_sums(rows_h)[:, 0], sums[:, h]), h
    assert np.array_equal(ind.predict(prob, offs, n), np.argmax(sums, axis=1))


def test_class_params():
    p = base("classifier", 5, 1, negative=-1, target=0.25, clauses=30)
    ph = O.Params()
    O.lib().orc_ind_class(ctypes.byref(p), 13, ctypes.byref(ph))
    assert ph.features == 13 and ph.input_bits == 26 and ph.input_chunks == 4 and ph.tail_mask == 0x03
    assert ph.target == p.target and ph.negative == pytest.approx(0.2) and ph.clauses == 32


# ---------------------------------------------------------------------------
# The individualized packers (santoku-matrix's bits_individualize and
# bits_to_cvec_ind, restated from their call sites and doc/sth.md:30-91).

def test_ind_packers_hand_example():
    """Union of 3 features; class 0 lists union features [2, 0], class 1 [1].
    Sample 0 has features {0, 2}, sample 1 has {1}.  Class 0 (k = 2, one byte
    per row): sample 0 has x = (1, 1), bits 0-1 set -> 0x03; sample 1 has
    x = (0, 0), the not-x bits 2-3 set -> 0x0C.  Class 1 (k = 1): sample 0
    absent -> not-x bit 1 -> 0x02; sample 1 present -> 0x01."""
    from santoku.tsetlin import data as D
    union = 3
    set_bits = np.array([0 * union + 0, 0 * union + 2, 1 * union + 1])
    ind, ind_off = D.bits_individualize(set_bits, [0, 2, 3], [2, 0, 1], union)
    assert ind.tolist() == [0, 1, 1] and ind_off.tolist() == [0, 2, 3]
    bitmap, dim = D.bits_to_cvec_ind(ind, ind_off, [0, 2, 3], 2)
    assert bitmap.tolist() == [0x03, 0x0C, 0x02, 0x01] and dim.tolist() == [0, 2, 4]


def test_ind_packers_equal_the_plain_packer_on_selected_columns():
    """Class h's block is the plain packer over the union columns it selects,
    in its order; blocks follow each other in class order.  Ragged widths
    (1 to 300 features, rows that are not whole words), duplicates in the
    input and samples without any feature."""
    from santoku.tsetlin import data as D
    rng = np.random.default_rng(4)
    n, union = 23, 400
    x = rng.random((n, union)) < 0.2
    x[5] = False
    widths = [1, 7, 300, 64, 13]
    ids = [rng.choice(union, k, replace=False) for k in widths]
    feat_offsets = np.concatenate([[0], np.cumsum(widths)])
    s, f = np.nonzero(x)
    set_bits = np.concatenate([s * union + f, (s * union + f)[:50]])      # with duplicates
    rng.shuffle(set_bits)
    ind, ind_off = D.bits_individualize(set_bits, feat_offsets, np.concatenate(ids), union)
    bitmap, dim = D.bits_to_cvec_ind(ind, ind_off, feat_offsets, n)
    assert dim[0] == 0 and dim[-1] == bitmap.size
    for h, k in enumerate(widths):
        want = D.bits_to_cvec(x[:, ids[h]]).reshape(-1)
        assert dim[h + 1] - dim[h] == n * ((2 * k + 7) // 8)
        assert np.array_equal(bitmap[dim[h]:dim[h + 1]], want), h
    # without flip: the x bits only
    raw, rdim = D.bits_to_cvec_ind(ind, ind_off, feat_offsets, n, flip=False)
    for h, k in enumerate(widths):
        want = np.packbits(x[:, ids[h]], axis=1, bitorder="little").reshape(-1)
        assert np.array_equal(raw[rdim[h]:rdim[h + 1]], want), h


def test_ind_packers_argument_errors():
    from santoku.tsetlin import data as D
    with pytest.raises(ValueError):
        D.bits_individualize([0, 1], [0, 2], [1, 1], 4)          # a feature listed twice
    with pytest.raises(ValueError):
        D.bits_individualize([0, 1], [0, 1], [9], 4)             # outside the union
    with pytest.raises(ValueError):
        D.bits_to_cvec_ind([0], [0, 1], [0, 0], 1)               # an empty class
    # entries past the rows are skipped
    bitmap, dim = D.bits_to_cvec_ind([0, 5, -1], [0, 3], [0, 2], 1)
    assert bitmap.tolist() == [0x01 | 0x08] and dim.tolist() == [0, 1]
