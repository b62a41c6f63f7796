"""ctypes front-end of the CPU oracle (tm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the engine package.  See tm_oracle.c for
what it restates and where each part is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_U8P = ctypes.POINTER(ctypes.c_uint8)
_I64P = ctypes.POINTER(ctypes.c_int64)
_U32P = ctypes.POINTER(ctypes.c_uint32)
_U64P = ctypes.POINTER(ctypes.c_uint64)


class Params(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_uint32),
        ("classes", ctypes.c_uint32),
        ("features", ctypes.c_uint32),
        ("clauses", ctypes.c_uint32),
        ("clause_tolerance", ctypes.c_uint32),
        ("clause_maximum", ctypes.c_uint32),
        ("state_bits", ctypes.c_uint32),
        ("include_bits", ctypes.c_uint32),
        ("target_opt", ctypes.c_double),
        ("negative_opt", ctypes.c_double),
        ("specificity", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("target", ctypes.c_uint32),
        ("input_bits", ctypes.c_uint32),
        ("input_chunks", ctypes.c_uint32),
        ("clause_chunks", ctypes.c_uint32),
        ("class_chunks", ctypes.c_uint32),
        ("state_chunks", ctypes.c_uint32),
        ("action_chunks", ctypes.c_uint32),
        ("tail_mask", ctypes.c_uint8),
        ("pad_", ctypes.c_uint8 * 7),
        ("negative", ctypes.c_double),
    ]


def build(native=False):
    """Compile the oracle with its Makefile (portable, or -march=native)."""
    target = "libtm_oracle_native.so" if native else "libtm_oracle.so"
    subprocess.run(["make", "-s", "-C", _HERE, target], check=True)
    return os.path.join(_HERE, target)


def lib(path=None):
    global _LIB
    if path is not None:
        return _bind(ctypes.CDLL(path))
    if _LIB is None:
        so = os.path.join(_HERE, "libtm_oracle.so")
        if not os.path.exists(so):
            build()
        _LIB = _bind(ctypes.CDLL(so))
    return _LIB


def _bind(L):
    P = ctypes.POINTER(Params)
    sig = {
        "orc_derive": (None, [P]),
        "orc_philox": (None, [_U32P, _U32P, _U32P]),
        "orc_lowbias32": (ctypes.c_uint32, [ctypes.c_uint32]),
        "orc_epoch_keys": (None, [ctypes.c_uint64, ctypes.c_uint32, _U32P]),
        "orc_perm": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _U32P]),
        "orc_skip_draw": (ctypes.c_uint32, [_U32P, ctypes.c_uint32, ctypes.c_uint32]),
        "orc_step_draws": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _U32P]),
        "orc_literal_key": (ctypes.c_uint32, [_U32P, ctypes.c_uint32, ctypes.c_uint32]),
        "orc_literal_draw": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
        "orc_threshold32": (ctypes.c_uint64, [ctypes.c_double]),
        "orc_setup": (None, [P, _U8P, _U8P, ctypes.c_uint64, ctypes.c_uint64]),
        "orc_inc_byte": (None, [P, _U8P, _U8P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint8]),
        "orc_dec_byte": (None, [P, _U8P, _U8P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint8]),
        "orc_inc": (None, [P, _U8P, _U8P, ctypes.c_uint64, _U8P]),
        "orc_dec_not_excluded": (None, [P, _U8P, _U8P, ctypes.c_uint64, _U8P]),
        "orc_inc_not_excluded": (None, [P, _U8P, _U8P, ctypes.c_uint64, _U8P]),
        "orc_calculate": (ctypes.c_uint8, [P, _U8P, _U8P, ctypes.c_uint32, _U32P, _U32P]),
        "orc_sums": (ctypes.c_long, [ctypes.c_uint8, _U32P]),
        "orc_class_sums": (None, [P, _U8P, _U8P, ctypes.c_uint32, _I64P]),
        "orc_predict_classifier": (None, [P, _U8P, _U8P, ctypes.c_uint32, _I64P]),
        "orc_predict_classifier_threads": (None, [P, _U8P, _U8P, ctypes.c_uint32, _I64P, ctypes.c_int]),
        "orc_predict_encoder_threads": (None, [P, _U8P, _U8P, ctypes.c_uint32, _U8P, ctypes.c_int]),
        "orc_predict_encoder": (None, [P, _U8P, _U8P, ctypes.c_uint32, _U8P]),
        "orc_s_draws": (ctypes.c_uint32, [P]),
        "orc_train_replay": (None, [P, _U8P, _U8P, ctypes.c_uint32, _U8P, _I64P, _U8P,
                                    ctypes.c_uint32, ctypes.c_uint32, _U64P]),
        "orc_train_replay_chunks": (None, [P, _U8P, _U8P, ctypes.c_uint32, _U8P, _I64P, _U8P,
                                    ctypes.c_uint32, ctypes.c_uint32, _U32P, ctypes.c_uint32, _U64P]),
        "orc_train_refshape": (None, [P, _U8P, _U8P, ctypes.c_uint32, _U8P, _I64P, _U8P,
                                      ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32]),
        "orc_max_threads": (ctypes.c_int, []),
        "orc_ind_class": (None, [P, ctypes.c_uint64, P]),
        "orc_train_replay_ind": (None, [P, _U64P, _U8P, _U8P, ctypes.c_uint32, _U8P, _I64P, _I64P, _U8P,
                                        ctypes.c_uint32, ctypes.c_uint32, _U64P]),
        "orc_class_sums_ind": (None, [P, _U64P, _U8P, _U8P, _I64P, ctypes.c_uint32, _I64P]),
        "orc_predict_ind": (None, [P, _U64P, _U8P, _U8P, _I64P, ctypes.c_uint32, _I64P, _U8P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def ptr(a, t=_U8P):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def params(type_="classifier", *, classes, features, clauses, clause_tolerance, clause_maximum,
           specificity, state=8, include_bits=1, target=-1.0, negative=-1.0, seed=0):
    p = Params()
    p.type = 0 if type_ == "classifier" else 1
    p.classes = classes
    p.features = features
    p.clauses = clauses
    p.clause_tolerance = clause_tolerance
    p.clause_maximum = clause_maximum
    p.state_bits = state
    p.include_bits = include_bits
    p.target_opt = target
    p.negative_opt = negative
    p.specificity = specificity
    p.seed = seed
    lib().orc_derive(ctypes.byref(p))
    return p


class Machine:
    """Oracle machine: counts [clause][plane][B] and actions [clause][B] bytes,
    the reference layout (automata.h:7-31)."""

    def __init__(self, p):
        self.p = p
        nclauses = p.classes * p.clauses
        self.counts = np.zeros(nclauses * (p.state_bits - 1) * p.input_chunks, np.uint8)
        self.actions = np.zeros(nclauses * p.input_chunks, np.uint8)
        self.stats = np.zeros(6, np.uint64)

    @property
    def nclauses(self):
        return self.p.classes * self.p.clauses

    def setup(self):
        lib().orc_setup(ctypes.byref(self.p), ptr(self.counts), ptr(self.actions), 0, self.nclauses - 1)

    def train_replay(self, rows, labels=None, codes=None, epochs=1, epoch_first=0):
        rows = np.ascontiguousarray(rows, np.uint8)
        n = rows.shape[0]
        lab = None if labels is None else np.ascontiguousarray(labels, np.int64)
        cod = None if codes is None else np.ascontiguousarray(codes, np.uint8)
        lib().orc_train_replay(ctypes.byref(self.p), ptr(self.counts), ptr(self.actions), n, ptr(rows),
                               ptr(lab, _I64P), ptr(cod), epoch_first, epochs, ptr(self.stats, _U64P))

    def train_replay_chunks(self, rows, chunks, labels=None, codes=None, epochs=1, epoch_first=0):
        """Replay of the listed chunks only (the others keep their state)."""
        rows = np.ascontiguousarray(rows, np.uint8)
        n = rows.shape[0]
        lab = None if labels is None else np.ascontiguousarray(labels, np.int64)
        cod = None if codes is None else np.ascontiguousarray(codes, np.uint8)
        lst = np.ascontiguousarray(chunks, np.uint32)
        lib().orc_train_replay_chunks(ctypes.byref(self.p), ptr(self.counts), ptr(self.actions), n, ptr(rows),
                                      ptr(lab, _I64P), ptr(cod), epoch_first, epochs, ptr(lst, _U32P), lst.size,
                                      ptr(self.stats, _U64P))

    def train_refshape(self, rows, labels=None, codes=None, epochs=1, threads=0, sample_limit=0, L=None):
        rows = np.ascontiguousarray(rows, np.uint8)
        n = rows.shape[0]
        lab = None if labels is None else np.ascontiguousarray(labels, np.int64)
        cod = None if codes is None else np.ascontiguousarray(codes, np.uint8)
        (L or lib()).orc_train_refshape(ctypes.byref(self.p), ptr(self.counts), ptr(self.actions), n, ptr(rows),
                                        ptr(lab, _I64P), ptr(cod), epochs, threads, sample_limit)

    def predict(self, rows):
        rows = np.ascontiguousarray(rows, np.uint8)
        n = rows.shape[0]
        if self.p.type == 0:
            out = np.zeros(n, np.int64)
            lib().orc_predict_classifier(ctypes.byref(self.p), ptr(self.actions), ptr(rows), n, ptr(out, _I64P))
        else:
            out = np.zeros((n, self.p.class_chunks), np.uint8)
            lib().orc_predict_encoder(ctypes.byref(self.p), ptr(self.actions), ptr(rows), n, ptr(out))
        return out

    def class_sums(self, rows):
        rows = np.ascontiguousarray(rows, np.uint8)
        n = rows.shape[0]
        out = np.zeros((n, self.p.classes), np.int64)
        lib().orc_class_sums(ctypes.byref(self.p), ptr(self.actions), ptr(rows), n, ptr(out, _I64P))
        return out


class IndMachine:
    """Oracle individualized machine (capi.c:528-606): class h is a one-class
    machine over feat_sizes[h] features; counts and actions of the classes
    sit one after another in class order, the reference's layout."""

    def __init__(self, p, feat_sizes):
        self.p = p
        self.feat_sizes = np.ascontiguousarray(feat_sizes, np.uint64)
        self.chunks = [(2 * int(k) + 7) // 8 for k in self.feat_sizes]
        self.counts = np.zeros(sum(p.clauses * (p.state_bits - 1) * b for b in self.chunks), np.uint8)
        self.actions = np.zeros(sum(p.clauses * b for b in self.chunks), np.uint8)
        self.stats = np.zeros(6, np.uint64)

    def action_slices(self):
        out, o = [], 0
        for b in self.chunks:
            out.append(slice(o, o + self.p.clauses * b))
            o += self.p.clauses * b
        return out

    def train_replay(self, problems, dim_offsets, n, labels=None, codes=None, epochs=1, epoch_first=0):
        prob = np.ascontiguousarray(problems, np.uint8)
        offs = np.ascontiguousarray(dim_offsets, np.int64)
        lab = None if labels is None else np.ascontiguousarray(labels, np.int64)
        cod = None if codes is None else np.ascontiguousarray(codes, np.uint8)
        lib().orc_train_replay_ind(ctypes.byref(self.p), ptr(self.feat_sizes, _U64P), ptr(self.counts),
                                   ptr(self.actions), n, ptr(prob), ptr(offs, _I64P), ptr(lab, _I64P), ptr(cod),
                                   epoch_first, epochs, ptr(self.stats, _U64P))

    def class_sums(self, problems, dim_offsets, n):
        prob = np.ascontiguousarray(problems, np.uint8)
        offs = np.ascontiguousarray(dim_offsets, np.int64)
        out = np.zeros((n, self.p.classes), np.int64)
        lib().orc_class_sums_ind(ctypes.byref(self.p), ptr(self.feat_sizes, _U64P), ptr(self.actions), ptr(prob),
                                 ptr(offs, _I64P), n, ptr(out, _I64P))
        return out

    def predict(self, problems, dim_offsets, n):
        prob = np.ascontiguousarray(problems, np.uint8)
        offs = np.ascontiguousarray(dim_offsets, np.int64)
        if self.p.type == 0:
            out = np.zeros(n, np.int64)
            lib().orc_predict_ind(ctypes.byref(self.p), ptr(self.feat_sizes, _U64P), ptr(self.actions), ptr(prob),
                                  ptr(offs, _I64P), n, ptr(out, _I64P), None)
        else:
            out = np.zeros((n, self.p.class_chunks), np.uint8)
            lib().orc_predict_ind(ctypes.byref(self.p), ptr(self.feat_sizes, _U64P), ptr(self.actions), ptr(prob),
                                  ptr(offs, _I64P), n, None, ptr(out))
        return out


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox(ptr(c, _U32P), ptr(k, _U32P), ptr(o, _U32P))
    return o


def perm(seed, epoch, n):
    o = np.zeros(n, np.uint32)
    lib().orc_perm(seed, epoch, n, ptr(o, _U32P))
    return o


def epoch_keys(seed, epoch):
    o = np.zeros(4, np.uint32)
    lib().orc_epoch_keys(seed, epoch, ptr(o, _U32P))
    return o
