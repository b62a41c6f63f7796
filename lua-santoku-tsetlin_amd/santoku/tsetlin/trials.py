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

from . import Machine, TsetlinError, _check, _host_buffer, _is_device_tensor, _options
from . import capi as _c


class Batch:
    def __init__(self, base, trials):
        """base: options shared by every trial (classes, features, negative, ...);
        trials: list of per-trial option dicts merged over base."""
        self._lib = _c.load_library()
        opts = (_c.Options * len(trials))()
        for i, t in enumerate(trials):
            opts[i] = _options(dict(base, **t))
        h = ctypes.c_void_p()
        _check(self._lib.tm_batch_create(opts, len(trials), ctypes.byref(h)))
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

    def train(self, samples, problems, solutions, iterations, each=None):
        """each(epoch, active) -> False stops all; clearing active[t] freezes trial t."""
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
        if _is_device_tensor(problems):
            _check(self._lib.tm_batch_train_classifier_dev(self._h, int(samples), problems.data_ptr(),
                                                           solutions.data_ptr(), int(iterations), cb, None))
        else:
            ra, rp = _host_buffer(problems)
            la, lp = _host_buffer(solutions, np.int64)
            keep += [ra, la]
            _check(self._lib.tm_batch_train_classifier(self._h, int(samples), rp, lp, int(iterations), cb, None))

    def accuracy(self, problems_dev, solutions_dev, n):
        """Held-out accuracy of every trial (device tensors)."""
        out = (ctypes.c_double * self.n)()
        _check(self._lib.tm_batch_accuracy_dev(self._h, problems_dev.data_ptr(), solutions_dev.data_ptr(), int(n), out))
        return np.array(out[:], dtype=np.float64)

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
                   seed=0):
    """A batch_fn for optimize.search_batched: trains every trial of a round in
    one Batch with per-epoch dev accuracy, patience early stop (a frozen trial
    keeps the state of its last epoch; its score is its best epoch's accuracy,
    optimize.lua:600-626) and returns the scores."""
    from .optimize import _swap_tolerance

    def fn(batch_params, info):
        trials = [dict(_swap_tolerance(dict(p)), seed=seed + i) for i, p in enumerate(batch_params)]
        b = Batch(base, trials)
        best = np.full(len(trials), -np.inf)
        since = np.zeros(len(trials), np.int64)

        def each(epoch, active):
            acc = b.accuracy(dev_rows, dev_y, dev_y.shape[0])
            for t in range(len(trials)):
                if not active[t]:
                    continue
                if acc[t] > best[t] + tolerance:
                    best[t], since[t] = acc[t], 0
                else:
                    since[t] += 1
                if patience and since[t] >= patience:
                    active[t] = 0
            return True

        b.train(train_rows.shape[0], train_rows, train_y, iterations, each)
        fn.last_batch = b
        return [float(x) for x in best]
    return fn
