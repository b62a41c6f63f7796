"""santoku.tsetlin -- host mirror of the reference's Lua module over the C ABI.

Mirrors module santoku.tsetlin.capi (lib/santoku/tsetlin/capi.c:1790-1804,
re-exported by lib/santoku/tsetlin.lua:1): ``create(type, opts)``,
``load(data, from_string)``, ``align``; and the machine methods of the
``santoku_tsetlin`` metatable (capi.c:419-431): train, predict, persist,
checkpoint, restore, reconfigure, type, destroy.  The README spellings
``classifier(opts)`` / ``encoder(opts)`` are exported too (README.md:17,32).

Errors the reference raises with luaL_error / tk_lua_verror raise TsetlinError
here, carrying the engine's message.  Byte buffers (problems, sentences, codes,
checkpoints) may be bytes, bytearray, numpy uint8 arrays or, for train and
predict, torch tensors on the GPU (taken as device pointers, no copy).
"""
import ctypes
import sys

import numpy as np

from . import capi as _c

align = 8


class TsetlinError(RuntimeError):
    """Raised where the reference raises a Lua error."""


def _check(rc):
    if rc != _c.TM_OK:
        msg = _c.load_library().tm_last_error()
        raise TsetlinError("%s: %s" % (_c.ERRORS.get(rc, rc), msg.decode() if msg else ""))


def _is_device_tensor(x):
    t = type(x)
    return t.__module__.startswith("torch") and t.__name__ == "Tensor" and x.is_cuda


_TORCH_DTYPES = {np.uint8: "uint8", np.int64: "int64"}


def _device_buffer(x, dtype, nbytes, what):
    """Pointer of a device tensor after the checks the kernels rely on: a
    contiguous GPU tensor of the right dtype holding at least nbytes.  A bad
    buffer raises here, before any kernel could read past it."""
    if not _is_device_tensor(x):
        raise TsetlinError("%s: expected a GPU tensor (the other inputs of this call are on the GPU)" % what)
    want = _TORCH_DTYPES[dtype]
    if str(x.dtype) != "torch." + want:
        raise TsetlinError("%s: expected dtype %s, got %s" % (what, want, x.dtype))
    if not x.is_contiguous():
        raise TsetlinError("%s: tensor must be contiguous" % what)
    have = x.numel() * x.element_size()
    if have < nbytes:
        raise TsetlinError("%s: tensor holds %d bytes, the call reads %d" % (what, have, nbytes))
    return x.data_ptr()


def _host_buffer(x, dtype=np.uint8):
    """(array, pointer) for a host buffer; keeps the array alive."""
    if isinstance(x, (bytes, bytearray, memoryview)):
        a = np.frombuffer(x, dtype=np.uint8)
        if dtype != np.uint8:
            a = a.view(dtype)
    else:
        a = np.ascontiguousarray(np.asarray(x), dtype=dtype)
    return a, a.ctypes.data


def _options(o):
    lib = _c.load_library()
    opts = _c.Options()
    lib.tm_options_default(ctypes.byref(opts))
    for key, field in (("classes", "classes"), ("hidden", "classes"), ("features", "features"),
                       ("visible", "features"), ("clauses", "clauses"), ("clause_tolerance", "clause_tolerance"),
                       ("clause_maximum", "clause_maximum"), ("state", "state"), ("include_bits", "include_bits"),
                       ("target", "target"), ("negative", "negative"), ("specificity", "specificity"),
                       ("seed", "seed")):
        if o.get(key) is not None:
            setattr(opts, field, type(getattr(opts, field))(o[key]))
    opts.reusable = 1 if o.get("reusable") else 0
    opts.individualized = 1 if o.get("individualized") else 0
    return opts


class Machine:
    """A classifier or encoder (the reference's tk_tsetlin_t userdata)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        self._lib = _c.load_library()

    def _handle(self):
        if not self._h:
            raise TsetlinError("machine destroyed")
        return self._h

    # capi.c:1774-1788
    def type(self):
        t = self._lib.tm_type(self._handle())
        if t < 0:
            _check(t)
        return "classifier" if t == _c.TM_CLASSIFIER else "encoder"

    def info(self):
        i = _c.Info()
        _check(self._lib.tm_get_info(self._handle(), ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in _c.Info._fields_}

    def individualized(self):
        return bool(self.info()["individualized"])

    def feature_sizes(self):
        """Per-class feature counts k_h of an individualized machine."""
        k = self.info()["classes"]
        out = (ctypes.c_uint64 * k)()
        _check(self._lib.tm_get_feature_sizes(self._handle(), out, k))
        return [int(v) for v in out]

    # capi.c:1484-1507 (dispatch), :1112-1191 (classifier), :1193-1282 (encoder),
    # :1284-1482 (individualized, with dim_offsets)
    def train(self, opts=None, **kw):
        o = dict(opts or {}, **kw)
        n = int(o["samples"])
        iters = int(o["iterations"])
        each = o.get("each")
        enc = self.type() == "encoder"
        rows = o["sentences"] if enc else o["problems"]
        tgt = o["codes"] if enc else o["solutions"]
        keep = []

        def on_epoch(_user, epoch):
            try:
                r = each(epoch)
            except Exception as ex:  # the reference prints and stops (capi.c:1171-1175)
                print("Error in Lua callback: %s" % ex, file=sys.stderr)
                return -1
            return 0 if r is False else 1

        cb = _c.EPOCH_FN(on_epoch) if each is not None else _c.EPOCH_FN()
        if self.individualized():
            if o.get("dim_offsets") is None:
                raise TsetlinError("train: dim_offsets is required for an individualized machine")
            da, dp = _host_buffer(o["dim_offsets"], np.int64)
            keep += [da]
            if _is_device_tensor(rows):
                # problems and targets in device memory (the _ind_dev calls);
                # the library checks dim_offsets and n against the byte size
                rows = rows.contiguous()
                rp = _device_buffer(rows, np.uint8, 0, "train: problems")
                nbytes = rows.numel()
                if enc:
                    tp = _device_buffer(tgt, np.uint8, n * ((self.info()["classes"] + 7) // 8), "train: codes")
                else:
                    tp = _device_buffer(tgt, np.int64, n * 8, "train: solutions")
                fn = self._lib.tm_train_encoder_ind_dev if enc else self._lib.tm_train_classifier_ind_dev
                _check(fn(self._handle(), n, rp, nbytes, tp, dp, iters, cb, None))
                return
            ra, rp = _host_buffer(rows)
            ta, tp = _host_buffer(tgt, np.uint8 if enc else np.int64)
            keep += [ra, ta]
            fn = self._lib.tm_train_encoder_ind if enc else self._lib.tm_train_classifier_ind
            _check(fn(self._handle(), n, rp, ra.nbytes, tp, dp, iters, cb, None))
            return
        if _is_device_tensor(rows):
            info = self.info()
            rp = _device_buffer(rows, np.uint8, n * info["input_chunks"], "train: problems")
            if enc:
                tp = _device_buffer(tgt, np.uint8, n * info["class_chunks"], "train: codes")
            else:
                tp = _device_buffer(tgt, np.int64, n * 8, "train: solutions")
            fn = self._lib.tm_train_encoder_dev if enc else self._lib.tm_train_classifier_dev
        else:
            ra, rp = _host_buffer(rows)
            ta, tp = _host_buffer(tgt, np.uint8 if enc else np.int64)
            keep += [ra, ta]
            fn = self._lib.tm_train_encoder if enc else self._lib.tm_train_classifier
        _check(fn(self._handle(), n, rp, tp, iters, cb, None))

    # capi.c:1088-1110, :903-948, :950-995; an individualized machine takes
    # predict(problems, dim_offsets, n[, threads]) (capi.c:997-1086)
    def predict(self, problems, *args):
        if self.individualized():
            if len(args) < 2:
                raise TsetlinError("predict: an individualized machine takes (problems, dim_offsets, n)")
            return self._predict_ind(problems, args[0], int(args[1]))
        n = int(args[0])
        enc = self.type() == "encoder"
        if _is_device_tensor(problems):
            import torch
            pp = _device_buffer(problems, np.uint8, n * self.info()["input_chunks"], "predict: problems")
            if enc:
                cc = (self.info()["classes"] + 7) // 8
                out = torch.empty((n, cc), dtype=torch.uint8, device=problems.device)
                _check(self._lib.tm_predict_encoder_dev(self._handle(), pp, n, out.data_ptr()))
            else:
                out = torch.empty(n, dtype=torch.int64, device=problems.device)
                _check(self._lib.tm_predict_classifier_dev(self._handle(), pp, n, out.data_ptr()))
            return out
        pa, pp = _host_buffer(problems)
        if enc:
            cc = (self.info()["classes"] + 7) // 8
            out = np.zeros((n, cc), np.uint8)
            _check(self._lib.tm_predict_encoder(self._handle(), pp, n, out.ctypes.data))
        else:
            out = np.zeros(n, np.int64)
            _check(self._lib.tm_predict_classifier(self._handle(), pp, n, out.ctypes.data))
        return out

    def _predict_ind(self, problems, dim_offsets, n):
        da, dp = _host_buffer(dim_offsets, np.int64)
        if _is_device_tensor(problems):
            import torch
            problems = problems.contiguous()
            pp = _device_buffer(problems, np.uint8, 0, "predict: problems")
            if self.type() == "encoder":
                out = torch.empty((n, (self.info()["classes"] + 7) // 8), dtype=torch.uint8, device=problems.device)
                _check(self._lib.tm_predict_encoder_ind_dev(self._handle(), pp, problems.numel(), dp, n,
                                                            out.data_ptr()))
            else:
                out = torch.empty(n, dtype=torch.int64, device=problems.device)
                _check(self._lib.tm_predict_classifier_ind_dev(self._handle(), pp, problems.numel(), dp, n,
                                                               out.data_ptr()))
            return out
        pa, pp = _host_buffer(problems)
        if self.type() == "encoder":
            out = np.zeros((n, (self.info()["classes"] + 7) // 8), np.uint8)
            _check(self._lib.tm_predict_encoder_ind(self._handle(), pp, pa.nbytes, dp, n, out.ctypes.data))
        else:
            out = np.zeros(n, np.int64)
            _check(self._lib.tm_predict_classifier_ind(self._handle(), pp, pa.nbytes, dp, n, out.ctypes.data))
        return out

    # capi.c:752-900
    def restrict(self, selected_dims):
        a, p = _host_buffer(selected_dims, np.int64)
        _check(self._lib.tm_restrict(self._handle(), p, a.size))

    # capi.c:1538-1562: returns the bytes when path is None
    def persist(self, path=None, persist_state=None):
        if path is not None:
            _check(self._lib.tm_persist_file(self._handle(), str(path).encode()))
            return None
        ln = ctypes.c_size_t(0)
        rc = self._lib.tm_persist(self._handle(), None, 0, ctypes.byref(ln))
        if rc not in (_c.TM_OK, -5):
            _check(rc)
        buf = np.zeros(ln.value, np.uint8)
        _check(self._lib.tm_persist(self._handle(), buf.ctypes.data, ln.value, ctypes.byref(ln)))
        return buf.tobytes()

    # capi.c:1564-1576: the buffer is resized by the callee; a GPU tensor of
    # uint8 is filled in place (tm_checkpoint_dev)
    def checkpoint(self, buf=None):
        ln = ctypes.c_size_t(0)
        if buf is not None and _is_device_tensor(buf):
            size = self.info()["action_chunks"]
            p = _device_buffer(buf, np.uint8, size, "checkpoint")
            _check(self._lib.tm_checkpoint_dev(self._handle(), p, size, ctypes.byref(ln)))
            return buf
        rc = self._lib.tm_checkpoint(self._handle(), None, 0, ctypes.byref(ln))
        if rc not in (_c.TM_OK, -5):
            _check(rc)
        out = np.zeros(ln.value, np.uint8)
        _check(self._lib.tm_checkpoint(self._handle(), out.ctypes.data, ln.value, ctypes.byref(ln)))
        if isinstance(buf, bytearray):
            buf[:] = out.tobytes()
            return buf
        return out

    # capi.c:1578-1588
    def restore(self, buf):
        if _is_device_tensor(buf):
            p = _device_buffer(buf, np.uint8, 0, "restore")
            _check(self._lib.tm_restore_dev(self._handle(), p, buf.numel()))
            return
        a, p = _host_buffer(buf)
        _check(self._lib.tm_restore(self._handle(), p, a.nbytes))

    def counts(self):
        """The counters in the reference layout counts[clause][plane][input_chunks]
        (tm_counts; a diagnostic read, no reference counterpart)."""
        ln = ctypes.c_size_t(0)
        rc = self._lib.tm_counts(self._handle(), None, 0, ctypes.byref(ln))
        if rc not in (_c.TM_OK, -5):
            _check(rc)
        out = np.zeros(ln.value, np.uint8)
        _check(self._lib.tm_counts(self._handle(), out.ctypes.data, ln.value, ctypes.byref(ln)))
        return out

    # capi.c:1590-1673
    def reconfigure(self, opts=None, **kw):
        o = dict(opts or {}, **kw)
        if "include_bits" not in o:
            o["include_bits"] = self.info()["include_bits"]
        _check(self._lib.tm_reconfigure(self._handle(), ctypes.byref(_options(o))))

    # capi.c:743-750
    def destroy(self):
        if self._h:
            self._lib.tm_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def timing(self):
        t, p = ctypes.c_double(0), ctypes.c_double(0)
        _check(self._lib.tm_last_timing(self._handle(), ctypes.byref(t), ctypes.byref(p)))
        return t.value, p.value

    def queue_timing(self):
        """Milliseconds the last train call spent building step queues."""
        q = ctypes.c_double(0)
        _check(self._lib.tm_last_queue_timing(self._handle(), ctypes.byref(q)))
        return q.value

    def train_stats(self):
        """Work counters of the last train call (see tm_last_train_stats)."""
        out = (ctypes.c_uint64 * 6)()
        _check(self._lib.tm_last_train_stats(self._handle(), out))
        keys = ("steps", "type1_true", "type1_grow", "type2_true", "type1_false", "draws")
        return dict(zip(keys, [int(v) for v in out]))

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# capi.c:1702-1724 (create dispatch)
def create(type_, opts=None, **kw):
    o = dict(opts or {}, **kw)
    lib = _c.load_library()
    h = ctypes.c_void_p()
    if type_ in ("classifier", "encoder") and o.get("individualized"):
        # capi.c:631-645, :668-682: the feature sets come from feat_offsets
        if o.get("feat_offsets") is None:
            raise TsetlinError("create: feat_offsets is required for an individualized machine")
        fa, fp = _host_buffer(o["feat_offsets"], np.int64)
        fn = lib.tm_create_classifier_ind if type_ == "classifier" else lib.tm_create_encoder_ind
        _check(fn(ctypes.byref(_options(o)), fp, fa.size, ctypes.byref(h)))
        return Machine(h.value)
    if type_ == "classifier":
        _check(lib.tm_create_classifier(ctypes.byref(_options(o)), ctypes.byref(h)))
    elif type_ == "encoder":
        _check(lib.tm_create_encoder(ctypes.byref(_options(o)), ctypes.byref(h)))
    else:
        raise TsetlinError("unexpected tsetlin machine type in create")
    return Machine(h.value)


def classifier(opts=None, **kw):
    return create("classifier", opts, **kw)


def encoder(opts=None, **kw):
    return create("encoder", opts, **kw)


# capi.c:1751-1772: a path, or the persisted bytes when from_string is true
def load(data, from_string=False, read_state=None):
    lib = _c.load_library()
    h = ctypes.c_void_p()
    if from_string:
        a, p = _host_buffer(data)
        _check(lib.tm_load(p, a.nbytes, ctypes.byref(h)))
    else:
        _check(lib.tm_load_file(str(data).encode(), ctypes.byref(h)))
    return Machine(h.value)
