"""santoku.tsetlin.trials -- batched search trials on the GPU engine.

A Batch trains N classifiers that share the data (classes, features, negative)
but differ in clauses, tolerance, maximum, target, specificity, include_bits
and seed, one launch per epoch for all of them (tm_batch_* in
include/tsetlin_amd.h).  Each trial's automata equal a solo run with the same
options.  optimize.lua runs the same trials one after another on a reusable
machine (optimize.lua:730-768); `batch_trial_fn` plugs the batch into
santoku.tsetlin.optimize.search_batched.
"""
import ctypes
import sys

import numpy as np

from . import Machine, TsetlinError, _check, _device_buffer, _host_buffer, _is_device_tensor, _options
from . import capi as _c


class Batch:
    def __init__(self, base, trials, encoder=False, feat_offsets=None):
        """base: options shared by every trial (classes, features, negative, ...);
        trials: list of per-trial option dicts merged over base.  encoder:
        the trials are encoders (classes = hidden bits), trained on sentences
        and codes (optimize.lua:842's search).  feat_offsets: individualized
        trials, one feature set per class shared by all of them (classes + 1
        offsets, as create(..., individualized=True) takes them); train then
        takes dim_offsets."""
        self._lib = _c.load_library()
        self.encoder = bool(encoder)
        self.individualized = feat_offsets is not None
        opts = (_c.Options * len(trials))()
        for i, t in enumerate(trials):
            opts[i] = _options(dict(base, **t))
        h = ctypes.c_void_p()
        if self.individualized:
            fa, fp = _host_buffer(feat_offsets, np.int64)
            create = self._lib.tm_batch_create_encoder_ind if self.encoder else self._lib.tm_batch_create_classifier_ind
            _check(create(opts, len(trials), fp, fa.size, ctypes.byref(h)))
        else:
            create = self._lib.tm_batch_create_encoder if self.encoder else self._lib.tm_batch_create
            _check(create(opts, len(trials), ctypes.byref(h)))
        self._h = h
        self.n = len(trials)
        self._machines = []
        for t in range(self.n):
            mh = ctypes.c_void_p()
            _check(self._lib.tm_batch_machine(self._h, t, ctypes.byref(mh)))
            m = Machine(mh.value)
            m.destroy = lambda: None          # owned by the batch
            m.__del__ = lambda: None
            self._machines.append(m)

    def machine(self, t):
        """Borrowed machine of trial t (predict / checkpoint / persist)."""
        return self._machines[t]

    def train(self, samples, problems, solutions, iterations, each=None, dim_offsets=None):
        """each(epoch, active) -> False stops all; clearing active[t] freezes
        trial t.  For an encoder batch, problems are the sentences and
        solutions the packed codes ([n][(classes + 7) // 8] bytes).  An
        individualized batch takes dim_offsets (class h's rows start at byte
        dim_offsets[h] of problems, as Machine.train)."""
        keep = []

        def on_epoch(_u, epoch, active_ptr):
            act = np.ctypeslib.as_array(active_ptr, shape=(self.n,))
            try:
                r = each(epoch, act)
            except Exception as ex:
                print("Error in epoch callback: %s" % ex, file=sys.stderr)
                return -1
            return 0 if r is False else 1

        cb = _c.BATCH_EPOCH_FN(on_epoch) if each is not None else _c.BATCH_EPOCH_FN()
        if self.individualized:
            if dim_offsets is None:
                raise TsetlinError("batch train: dim_offsets is required for individualized trials")
            da, dp = _host_buffer(dim_offsets, np.int64)
            keep += [da]
            n = int(samples)
            if _is_device_tensor(problems):
                # the library checks dim_offsets and n against the byte size
                problems = problems.contiguous()
                pp = _device_buffer(problems, np.uint8, 0, "batch train: problems")
                if self.encoder:
                    tp = _device_buffer(solutions, np.uint8, n * self._machines[0].info()["class_chunks"],
                                        "batch train: codes")
                    fn = self._lib.tm_batch_train_encoder_ind_dev
                else:
                    tp = _device_buffer(solutions, np.int64, n * 8, "batch train: solutions")
                    fn = self._lib.tm_batch_train_classifier_ind_dev
                _check(fn(self._h, n, pp, problems.numel(), tp, dp, int(iterations), cb, None))
            else:
                ra, rp = _host_buffer(problems)
                ta, tp = _host_buffer(solutions, np.uint8 if self.encoder else np.int64)
                keep += [ra, ta]
                fn = self._lib.tm_batch_train_encoder_ind if self.encoder else self._lib.tm_batch_train_classifier_ind
                _check(fn(self._h, n, rp, ra.nbytes, tp, dp, int(iterations), cb, None))
        elif self.encoder:
            if _is_device_tensor(problems):
                info = self._machines[0].info()
                pp = _device_buffer(problems, np.uint8, int(samples) * info["input_chunks"], "batch train: sentences")
                cp = _device_buffer(solutions, np.uint8, int(samples) * info["class_chunks"], "batch train: codes")
                _check(self._lib.tm_batch_train_encoder_dev(self._h, int(samples), pp, cp, int(iterations), cb, None))
            else:
                ra, rp = _host_buffer(problems)
                ca, cp = _host_buffer(solutions)
                keep += [ra, ca]
                _check(self._lib.tm_batch_train_encoder(self._h, int(samples), rp, cp, int(iterations), cb, None))
        elif _is_device_tensor(problems):
            row = self._machines[0].info()["input_chunks"]
            pp = _device_buffer(problems, np.uint8, int(samples) * row, "batch train: problems")
            lp = _device_buffer(solutions, np.int64, int(samples) * 8, "batch train: solutions")
            _check(self._lib.tm_batch_train_classifier_dev(self._h, int(samples), pp, lp, int(iterations), cb, None))
        else:
            ra, rp = _host_buffer(problems)
            la, lp = _host_buffer(solutions, np.int64)
            keep += [ra, la]
            _check(self._lib.tm_batch_train_classifier(self._h, int(samples), rp, lp, int(iterations), cb, None))

    def accuracy(self, problems_dev, solutions_dev, n):
        """Held-out accuracy of every trial (device tensors)."""
        out = (ctypes.c_double * self.n)()
        row = self._machines[0].info()["input_chunks"]
        pp = _device_buffer(problems_dev, np.uint8, int(n) * row, "batch accuracy: problems")
        lp = _device_buffer(solutions_dev, np.int64, int(n) * 8, "batch accuracy: solutions")
        _check(self._lib.tm_batch_accuracy_dev(self._h, pp, lp, int(n), out))
        return np.array(out[:], dtype=np.float64)

    def class_accuracy(self, problems_dev, solutions_dev, n, dim_offsets=None):
        """evaluator.class_accuracy of every trial (tm_batch_class_accuracy_dev):
        an (ntrials, 3) array of macro precision, recall, F1.  Individualized
        trials take the dev problems buffer and its dim_offsets
        (tm_batch_class_accuracy_ind_dev)."""
        out = (ctypes.c_double * (3 * self.n))()
        if self.individualized:
            if dim_offsets is None:
                raise TsetlinError("batch class accuracy: dim_offsets is required for individualized trials")
            da, dp = _host_buffer(dim_offsets, np.int64)
            problems_dev = problems_dev.contiguous()
            pp = _device_buffer(problems_dev, np.uint8, 0, "batch class accuracy: problems")
            lp = _device_buffer(solutions_dev, np.int64, int(n) * 8, "batch class accuracy: solutions")
            _check(self._lib.tm_batch_class_accuracy_ind_dev(self._h, pp, problems_dev.numel(), lp, dp, int(n), out))
            return np.array(out[:], dtype=np.float64).reshape(self.n, 3)
        row = self._machines[0].info()["input_chunks"]
        pp = _device_buffer(problems_dev, np.uint8, int(n) * row, "batch class accuracy: problems")
        lp = _device_buffer(solutions_dev, np.int64, int(n) * 8, "batch class accuracy: solutions")
        _check(self._lib.tm_batch_class_accuracy_dev(self._h, pp, lp, int(n), out))
        return np.array(out[:], dtype=np.float64).reshape(self.n, 3)

    def tuning(self):
        lanes, wpl, order, ms = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
        _check(self._lib.tm_batch_tuning(self._h, ctypes.byref(lanes), ctypes.byref(wpl), ctypes.byref(order),
                                         ctypes.byref(ms)))
        return {"lanes": lanes.value, "words_per_lane": wpl.value, "order": order.value, "epoch_ms": ms.value}

    def timing(self):
        t, e = ctypes.c_double(), ctypes.c_double()
        _check(self._lib.tm_batch_timing(self._h, ctypes.byref(t), ctypes.byref(e)))
        return t.value, e.value

    def destroy(self):
        if self._h:
            self._lib.tm_batch_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def batch_trial_fn(base, train_rows, train_y, dev_rows, dev_y, iterations, patience=None, tolerance=1e-8,
                   seed=0, metric="f1", encoder=False, feat_offsets=None, dim_offsets=None, dev_dim_offsets=None):
    """A batch_fn for optimize.search_batched: every trial of a round trains
    in one Batch, scored after each epoch with the reference's search metric
    (macro F1 of evaluator.class_accuracy on the dev rows, as the classifier
    specs' search_metric, classify/mnist.lua:93-97; metric="accuracy" scores
    plain accuracy).  As train_tm (optimize.lua:586-653): a trial that improves
    by more than `tolerance` checkpoints its actions (kept in device memory),
    `patience` epochs without improvement freeze it, and after training every
    checkpointed trial is restored to its best epoch and scored again, which is
    its result.  Without patience there is no checkpoint and a trial's result
    is its last epoch's score.  Trial t takes seed + t.

    encoder=True: the trials are encoders (base's classes = hidden bits) on
    sentences train_rows and packed codes train_y, scored by the encoder
    specs' search metric, evaluator.encoding_accuracy's mean_hamming (one
    minus the mean bit error) of each trial's codes for dev_rows against
    dev_y (encoder/mnist.lua:368-372).

    feat_offsets: individualized trials sharing those per-class feature
    sets; train_rows and dev_rows are then individualized problems buffers
    read at dim_offsets and dev_dim_offsets, and each trial is scored through
    its own machine's individualized predict."""
    import torch

    from . import evaluator
    from .optimize import _swap_tolerance

    ind = feat_offsets is not None
    if ind and (dim_offsets is None or dev_dim_offsets is None):
        raise TsetlinError("batch_trial_fn: individualized trials need dim_offsets and dev_dim_offsets")
    n_dev = dev_y.shape[0]

    def predict(m):
        return m.predict(dev_rows, dev_dim_offsets, n_dev) if ind else m.predict(dev_rows, n_dev)

    def score(b):
        if encoder:
            return np.array([evaluator.encoding_accuracy(predict(b.machine(t)), dev_y, n_dev,
                                                         base["classes"])["mean_hamming"] for t in range(b.n)])
        if ind:
            if metric != "accuracy":
                return b.class_accuracy(dev_rows, dev_y, n_dev, dim_offsets=dev_dim_offsets)[:, 2]
            return np.array([float((predict(b.machine(t)) == dev_y).sum()) / n_dev for t in range(b.n)])
        if metric == "accuracy":
            return b.accuracy(dev_rows, dev_y, dev_y.shape[0])
        return b.class_accuracy(dev_rows, dev_y, dev_y.shape[0])[:, 2]

    def fn(batch_params, info):
        trials = [dict(_swap_tolerance(dict(p)), seed=seed + i) for i, p in enumerate(batch_params)]
        b = Batch(base, trials, encoder=encoder, feat_offsets=feat_offsets)
        nt = len(trials)
        best = np.full(nt, -np.inf)
        last = np.full(nt, -np.inf)
        since = np.zeros(nt, np.int64)
        best_epoch = np.zeros(nt, np.int64)
        ckpt = [None] * nt

        def each(epoch, active):
            sc = score(b)
            for t in range(nt):
                if not active[t]:
                    continue
                last[t] = sc[t]
                if sc[t] > best[t] + tolerance:
                    best[t], since[t], best_epoch[t] = sc[t], 0, epoch
                    if patience:
                        m = b.machine(t)
                        if ckpt[t] is None:
                            ckpt[t] = torch.empty(m.info()["action_chunks"], dtype=torch.uint8,
                                                  device=dev_rows.device)
                        m.checkpoint(ckpt[t])
                else:
                    since[t] += 1
                if patience and since[t] >= patience:
                    active[t] = 0
            return True

        b.train(train_y.shape[0], train_rows, train_y, iterations, each, dim_offsets=dim_offsets)
        if any(c is not None for c in ckpt):
            for t in range(nt):
                if ckpt[t] is not None:
                    b.machine(t).restore(ckpt[t])
            sc = score(b)
            for t in range(nt):
                if ckpt[t] is not None:
                    last[t] = sc[t]
        fn.last_batch = b
        fn.last_best = best.copy()
        fn.last_best_epoch = best_epoch.copy()
        return [float(x) for x in last]
    return fn
