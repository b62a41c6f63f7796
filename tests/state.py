"""Whole-state comparison of an engine machine with the oracle: the include
plane (checkpoint) and every count plane (tm_counts), both in the reference's
layout (automata.h:7-31).  Imported by the GPU parity tests.

One set of count bits is left out: the padding bits of a row's last byte when
that byte goes through the reference's masked byte path (input_chunks not a
multiple of 16).  There the reference masks only the carry into the action
plane (automata.h:97,106), so those counters count modulo 2^(state_bits-1)
and never set their action bit; the engine leaves them at their setup value.
They reach nothing: not the action plane, not evaluation (tail-masked,
capi.c:106-109), not "included" (its action bit is 0), not the persisted
state (capi.c:1509-1531).  Every other count bit is compared."""
import numpy as np


def _diff(got, want):
    return int((got != want).sum())


def _row_mask(input_bits):
    """Compared bits of each byte of a row of input_bits literals."""
    B = (input_bits + 7) // 8
    m = np.full(B, 0xFF, np.uint8)
    tail = input_bits & 7
    if tail and B % 16 != 0:
        m[-1] = (1 << tail) - 1
    return m


def count_mask(orc):
    """Byte mask over orc.counts of the compared count bits."""
    M = orc.p.state_bits - 1
    if hasattr(orc, "feat_sizes"):            # individualized: class after class
        parts = [np.tile(_row_mask(2 * int(k)), orc.p.clauses * M) for k in orc.feat_sizes]
        return np.concatenate(parts)
    return np.tile(_row_mask(orc.p.input_bits), orc.p.classes * orc.p.clauses * M)


def assert_same_state(eng, orc, chunks=None, label=""):
    """Actions and the count planes of `eng` equal the oracle machine's, for
    every clause or only the listed 8-clause chunks.  The engine must hold
    its counters (created reusable)."""
    acts, cnt = eng.checkpoint(), eng.counts()
    want_a, want_c = orc.actions, orc.counts
    assert acts.size == want_a.size and cnt.size == want_c.size, (label, acts.size, want_a.size, cnt.size,
                                                                    want_c.size)
    mask = count_mask(orc)
    cnt, want_c = cnt & mask, want_c & mask
    M = orc.p.state_bits - 1
    if chunks is not None:
        B = orc.p.input_chunks
        cl = np.concatenate([np.arange(8 * c, 8 * c + 8) for c in chunks])
        acts, want_a = acts.reshape(-1, B)[cl], want_a.reshape(-1, B)[cl]
        cnt, want_c = cnt.reshape(-1, M, B)[cl], want_c.reshape(-1, M, B)[cl]
    assert np.array_equal(acts, want_a), "%s: actions differ in %d bytes" % (label, _diff(acts, want_a))
    if not np.array_equal(cnt, want_c):
        if hasattr(orc, "feat_sizes"):
            per = _diff(cnt, want_c)
        else:
            B = orc.p.input_chunks
            per = [_diff(cnt.reshape(-1, M, B)[:, p], want_c.reshape(-1, M, B)[:, p]) for p in range(M)]
        raise AssertionError("%s: count planes differ (bytes per plane: %s)" % (label, per))
