"""Literal-row packing and the synthetic workloads the engine is measured on.

``bits_to_cvec`` produces the engine's input contract: one flip-interleaved row
per sample, ``ceil(2F/8)`` bytes, bits ``0..F-1`` = x and bits ``F..2F-1`` = not
x, least-significant bit first, padding bits zero.  It follows the layout of
santoku-matrix's ``ivec:bits_to_cvec(n, F, true)`` as the reference documents it
(doc/sth.md:85-91) and uses it (test/spec/santoku/tsetlin/classify/mnist.lua:60).
"""
import numpy as np


def bits_to_cvec(x):
    """x: (n, F) bool/0-1 array -> (n, ceil(2F/8)) uint8 flip-interleaved rows."""
    x = np.asarray(x).astype(bool, copy=False)
    lits = np.concatenate([x, ~x], axis=1)
    return np.packbits(lits, axis=1, bitorder="little")


def bits_individualize(set_bits, feat_offsets, feat_ids, union_size):
    """toks:bits_individualize(offsets, ids, union_size) of santoku-matrix, as
    the reference's callers use it (doc/sth.md:38; classify/newsgroups.lua:96,
    encoder/mnist.lua:396): set_bits holds sample * union_size + f for each
    present feature f of the union vocabulary; class h's features are
    feat_ids[feat_offsets[h] : feat_offsets[h + 1]] (indices into the union,
    distinct within a class).  Returns (ind, ind_offsets): class h's entries
    ind[ind_offsets[h] : ind_offsets[h + 1]] = sample * k_h + i for every
    present feature feat_ids[feat_offsets[h] + i], sorted, and the classes + 1
    offsets.  Entries outside [0, n * union_size) are the caller's error:
    negative ones raise, others map to samples past n and are skipped by the
    packers' range check."""
    set_bits = np.asarray(set_bits, np.int64).ravel()
    feat_offsets = np.asarray(feat_offsets, np.int64).ravel()
    feat_ids = np.asarray(feat_ids, np.int64).ravel()
    if union_size <= 0:
        raise ValueError("bits_individualize: union_size must be positive")
    if (set_bits < 0).any():
        raise ValueError("bits_individualize: negative set bit")
    s, f = set_bits // union_size, set_bits % union_size
    out, offs = [], [0]
    for h in range(feat_offsets.size - 1):
        ids = feat_ids[feat_offsets[h]:feat_offsets[h + 1]]
        if np.unique(ids).size != ids.size:
            raise ValueError("bits_individualize: class %d lists a feature twice" % h)
        if ids.size and (ids.min() < 0 or ids.max() >= union_size):
            raise ValueError("bits_individualize: class %d lists a feature outside the union" % h)
        lut = np.full(union_size, -1, np.int64)
        lut[ids] = np.arange(ids.size)
        loc = lut[f]
        keep = loc >= 0
        e = np.unique(s[keep] * ids.size + loc[keep])
        out.append(e)
        offs.append(offs[-1] + e.size)
    return (np.concatenate(out) if out else np.zeros(0, np.int64)), np.asarray(offs, np.int64)


def bits_to_cvec_ind(ind, ind_offsets, feat_offsets, n, flip=True):
    """ivec:bits_to_cvec_ind(ind_offsets, feat_offsets, n, flip) of
    santoku-matrix (doc/sth.md:39-55): class h's entries (sample * k_h +
    feature, k_h = feat_offsets[h + 1] - feat_offsets[h]) packed into n rows
    of ceil(2 k_h / 8) bytes, flip-interleaved as bits_to_cvec (ceil(k_h / 8)
    bytes of x bits without flip), the classes' blocks back to back.  Returns
    (bitmap, dim_offsets): a uint8 array and the classes + 1 byte offsets of
    the blocks (the last one the total), which the individualized train and
    predict calls take.  Entries outside [0, n * k_h) are skipped."""
    ind = np.asarray(ind, np.int64).ravel()
    ind_offsets = np.asarray(ind_offsets, np.int64).ravel()
    feat_offsets = np.asarray(feat_offsets, np.int64).ravel()
    classes = feat_offsets.size - 1
    if classes < 1 or ind_offsets.size != classes + 1:
        raise ValueError("bits_to_cvec_ind: offsets must have classes + 1 entries")
    blocks, dim = [], [0]
    for h in range(classes):
        k = int(feat_offsets[h + 1] - feat_offsets[h])
        if k <= 0:
            raise ValueError("bits_to_cvec_ind: class %d has no features" % h)
        e = ind[ind_offsets[h]:ind_offsets[h + 1]]
        e = e[(e >= 0) & (e < n * k)]
        x = np.zeros((n, k), bool)
        x[e // k, e % k] = True
        rows = bits_to_cvec(x) if flip else np.packbits(x, axis=1, bitorder="little")
        blocks.append(rows.reshape(-1))
        dim.append(dim[-1] + rows.size)
    return np.concatenate(blocks), np.asarray(dim, np.int64)


def codes_from_bits(h):
    """h: (n, H) bool -> (n, ceil(H/8)) uint8 codes, LSB first (capi.c:1254)."""
    return np.packbits(np.asarray(h).astype(bool, copy=False), axis=1, bitorder="little")


# ---------------------------------------------------------------------------
# Workload 1: plain classifier, CPU-runnable (8192 train + 2048 test,
# 256 Bernoulli(0.5) features, 4 classes, 5% label flips).
# ---------------------------------------------------------------------------

def xor_and_task(seed, n_train=8192, n_test=2048, features=256, flip=0.05):
    """class = 2*(x3 AND x17) + (x40 XOR x91); a flipped label moves to one of
    the 3 other classes uniformly (Bayes ceiling 95%)."""
    rng = np.random.default_rng(seed)
    n = n_train + n_test
    x = rng.random((n, features)) < 0.5
    y = 2 * (x[:, 3] & x[:, 17]).astype(np.int64) + (x[:, 40] ^ x[:, 91]).astype(np.int64)
    flip_mask = rng.random(n) < flip
    y[flip_mask] = (y[flip_mask] + rng.integers(1, 4, flip_mask.sum())) % 4
    rows = bits_to_cvec(x)
    return dict(train_rows=rows[:n_train], train_y=y[:n_train], test_rows=rows[n_train:], test_y=y[n_train:],
                features=features, classes=4, x=x)


# ---------------------------------------------------------------------------
# Workload 2: plain classifier at 2048 clauses/class (200k train + 20k test,
# 4096 sparse Bernoulli(0.05) features, 20 classes from planted conjunctions).
# ---------------------------------------------------------------------------

def planted_conjunctions(seed, classes=20, features=4096, per_class=3):
    """Three conjunctions of 2..4 literals per class 1..classes-1.  Each has one
    positive literal and the rest negated, so it holds with probability about
    0.05 * 0.95^(k-1) at density 0.05 (all-positive conjunctions would almost
    never fire)."""
    rng = np.random.default_rng(seed ^ 0x5EED)
    conj = []
    for c in range(1, classes):
        for _ in range(per_class):
            k = int(rng.integers(2, 5))
            feats = rng.choice(features, size=k, replace=False)
            pol = np.zeros(k, bool)
            pol[0] = True
            conj.append((c, feats, pol))
    return conj


def label_planted(x, conj, classes):
    n = x.shape[0]
    y = np.zeros(n, np.int64)
    done = np.zeros(n, bool)
    for c, feats, pol in conj:
        m = np.ones(n, bool)
        for f, p in zip(feats, pol):
            m &= x[:, f] if p else ~x[:, f]
        hit = m & ~done
        y[hit] = c
        done |= hit
    return y


def sparse_planted_task(seed, n_train=200_000, n_test=20_000, features=4096, classes=20,
                        density=0.05, noise=0.02, block=8192):
    """First matching conjunction wins (class order 1..classes-1), else class 0;
    2% of labels are replaced by a uniformly random different class.  Rows are
    generated in blocks so peak memory stays at a few hundred MB."""
    rng = np.random.default_rng(seed)
    conj = planted_conjunctions(seed, classes, features)
    n = n_train + n_test
    rows = np.empty((n, (2 * features + 7) // 8), np.uint8)
    y = np.empty(n, np.int64)
    for s in range(0, n, block):
        e = min(n, s + block)
        xb = rng.random((e - s, features), dtype=np.float32) < density
        y[s:e] = label_planted(xb, conj, classes)
        rows[s:e] = bits_to_cvec(xb)
    flip = rng.random(n) < noise
    y[flip] = (y[flip] + rng.integers(1, classes, flip.sum())) % classes
    return dict(train_rows=rows[:n_train], train_y=y[:n_train], test_rows=rows[n_train:], test_y=y[n_train:],
                features=features, classes=classes, conj=conj)


# ---------------------------------------------------------------------------
# Workload 3: plain encoder (8192 visible Bernoulli(0.1) bits -> 256 hidden
# bits = sign of a fixed Gaussian projection of x - 0.1, 3% bit flips).
# ---------------------------------------------------------------------------

def projection_codes_task(seed, n=500_000, visible=8192, hidden=256, density=0.1, flip=0.03, block=4096):
    rng = np.random.default_rng(seed)
    w = np.random.default_rng(seed ^ 0xC0DE).standard_normal((visible, hidden)).astype(np.float32)
    rows = np.empty((n, (2 * visible + 7) // 8), np.uint8)
    codes = np.empty((n, (hidden + 7) // 8), np.uint8)
    for s in range(0, n, block):
        e = min(n, s + block)
        xb = rng.random((e - s, visible), dtype=np.float32) < density
        h = (xb.astype(np.float32) - density) @ w > 0
        h ^= rng.random(h.shape) < flip
        rows[s:e] = bits_to_cvec(xb)
        codes[s:e] = codes_from_bits(h)
    return dict(rows=rows, codes=codes, visible=visible, hidden=hidden)


# ---------------------------------------------------------------------------
# Device packers (the C ABI's tm_bits_to_cvec_dev / tm_flip_interleave_dev):
# torch tensors on the GPU in, a uint8 tensor of rows out.
# ---------------------------------------------------------------------------

def _dev_lib():
    from . import capi
    return capi.load_library()


def _dev_check(rc):
    if rc != 0:
        from . import TsetlinError, _c
        raise TsetlinError("%s: %s" % (_c.ERRORS.get(rc, rc), _dev_lib().tm_last_error().decode()))


def bits_to_cvec_dev(set_bits, n, features, out=None):
    """ivec:bits_to_cvec(n, features, true) on the device: set_bits is an int64
    tensor of sample * features + feature indices.  Returns (rows, invalid):
    an (n, ceil(2F/8)) uint8 tensor and the count of indices outside
    [0, n * features), which are skipped."""
    import ctypes
    import torch
    from . import _device_buffer
    set_bits = set_bits.contiguous()
    sp = _device_buffer(set_bits, np.int64, set_bits.numel() * 8, "bits_to_cvec: set_bits")
    if out is None:
        out = torch.empty((n, (2 * features + 7) // 8), dtype=torch.uint8, device=set_bits.device)
    op = _device_buffer(out, np.uint8, n * ((2 * features + 7) // 8), "bits_to_cvec: out")
    bad = ctypes.c_uint64(0)
    _dev_check(_dev_lib().tm_bits_to_cvec_dev(sp, set_bits.numel(), n, features, op, ctypes.byref(bad)))
    return out, bad.value


def bits_to_cvec_ind_dev(ind, ind_offsets, feat_offsets, n, out=None):
    """bits_to_cvec_ind(ind_offsets, feat_offsets, n, true) on the device: ind
    is an int64 GPU tensor of class-major entries (sample * k_h + feature),
    ind_offsets and feat_offsets host arrays of classes + 1 entries.  Returns
    (bitmap, dim_offsets, invalid): a uint8 GPU tensor, the classes + 1 byte
    offsets (numpy int64) and the count of skipped out-of-range entries."""
    import ctypes
    import torch
    from . import _device_buffer
    ind = ind.contiguous()
    io = np.ascontiguousarray(ind_offsets, np.int64)
    fo = np.ascontiguousarray(feat_offsets, np.int64)
    classes = fo.size - 1
    if classes < 1 or io.size != classes + 1:
        raise ValueError("bits_to_cvec_ind: offsets must have classes + 1 entries")
    if io[-1] > ind.numel():
        raise ValueError("bits_to_cvec_ind: ind_offsets run past the %d entries of ind" % ind.numel())
    k = np.diff(fo)
    total = int((n * ((2 * k + 7) // 8)).sum()) if (k > 0).all() else 0
    ip = _device_buffer(ind, np.int64, int(io[-1]) * 8, "bits_to_cvec_ind: ind")
    if out is None:
        out = torch.empty(max(total, 1), dtype=torch.uint8, device=ind.device)
    op = _device_buffer(out, np.uint8, total, "bits_to_cvec_ind: out")
    dim = np.zeros(classes + 1, np.int64)
    bad = ctypes.c_uint64(0)
    _dev_check(_dev_lib().tm_bits_to_cvec_ind_dev(ip, io.ctypes.data, fo.ctypes.data, classes, n, op,
                                                  out.numel(), dim.ctypes.data, ctypes.byref(bad)))
    return out, dim, bad.value


def flip_interleave_dev(raw, n, features, out=None):
    """Flip-interleaved rows from bit-packed rows (ceil(F/8) bytes, LSB first)."""
    import torch
    from . import _device_buffer
    raw = raw.contiguous()
    rp = _device_buffer(raw, np.uint8, n * ((features + 7) // 8), "flip_interleave: raw")
    if out is None:
        out = torch.empty((n, (2 * features + 7) // 8), dtype=torch.uint8, device=raw.device)
    op = _device_buffer(out, np.uint8, n * ((2 * features + 7) // 8), "flip_interleave: out")
    _dev_check(_dev_lib().tm_flip_interleave_dev(rp, n, features, op))
    return out
