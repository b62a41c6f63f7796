"""santoku.tsetlin.evaluator -- the TM metrics used by the search driver.

Restates tm_class_accuracy (lib/santoku/tsetlin/evaluator/capi.c:119-203) and
tm_encoding_accuracy (evaluator/capi.c:497-574).  The other evaluator
functions (retrieval, clustering, dendrograms, bit selection) belong to the
spectral-hashing pipeline and are not part of this engine.
"""
import numpy as np


def _as_array(x, dtype):
    if isinstance(x, (bytes, bytearray, memoryview)):
        return np.frombuffer(x, dtype=np.uint8).view(dtype)
    t = type(x)
    if t.__module__.startswith("torch"):
        x = x.detach().cpu().numpy()
    return np.asarray(x).reshape(-1).astype(dtype, copy=False)


def _on_device(*xs):
    for x in xs:
        t = type(x)
        if not (t.__module__.startswith("torch") and t.__name__ == "Tensor" and x.is_cuda):
            return False
    return True


def _dev_call(fn, *args):
    from . import TsetlinError, capi
    rc = fn(*args)
    if rc != 0:
        raise TsetlinError("%s: %s" % (capi.ERRORS.get(rc, rc), capi.load_library().tm_last_error().decode()))


def class_accuracy(predicted, expected, n_samples, n_classes, threads=None):
    """Per-class and macro precision/recall/F1 (evaluator/capi.c:119-203).
    Samples whose predicted or expected class is >= n_classes (after the
    reference's unsigned conversion) are ignored.  GPU tensors are scored on
    the device (tm_class_accuracy_dev), with the same doubles."""
    if n_classes == 0:
        raise ValueError("class_accuracy: n_classes must be > 0")
    if _on_device(predicted, expected):
        import ctypes
        from . import capi
        per = (ctypes.c_double * (3 * n_classes))()
        mac = (ctypes.c_double * 3)()
        from . import _device_buffer
        predicted, expected = predicted.contiguous(), expected.contiguous()
        pp = _device_buffer(predicted, np.int64, n_samples * 8, "class_accuracy: predicted")
        ep = _device_buffer(expected, np.int64, n_samples * 8, "class_accuracy: expected")
        _dev_call(capi.load_library().tm_class_accuracy_dev, pp, ep, n_samples, n_classes, per, mac)
        classes = [{"precision": per[3 * c], "recall": per[3 * c + 1], "f1": per[3 * c + 2]} for c in range(n_classes)]
        return {"classes": classes, "precision": mac[0], "recall": mac[1], "f1": mac[2]}
    p = _as_array(predicted, np.int64)[:n_samples].astype(np.uint64).astype(np.uint32)
    e = _as_array(expected, np.int64)[:n_samples].astype(np.uint64).astype(np.uint32)
    ok = (p < n_classes) & (e < n_classes)
    p, e = p[ok].astype(np.int64), e[ok].astype(np.int64)
    hit = p == e
    tp = np.bincount(e[hit], minlength=n_classes).astype(np.float64)
    fp = np.bincount(p[~hit], minlength=n_classes).astype(np.float64)
    fn = np.bincount(e[~hit], minlength=n_classes).astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        prec = np.where(tp + fp > 0, tp / (tp + fp), 0.0)
        rec = np.where(tp + fn > 0, tp / (tp + fn), 0.0)
        f1 = np.where(prec + rec > 0, 2 * prec * rec / (prec + rec), 0.0)
    classes = [{"precision": float(prec[c]), "recall": float(rec[c]), "f1": float(f1[c])} for c in range(n_classes)]
    # sums in class order, as the reference's loop (evaluator/capi.c:156-166)
    pa = ra = fa = 0.0
    for c in range(n_classes):
        pa += float(prec[c])
        ra += float(rec[c])
        fa += float(f1[c])
    return {"classes": classes, "precision": pa / n_classes, "recall": ra / n_classes, "f1": fa / n_classes}


def encoding_accuracy(predicted, expected, n_samples, n_hidden, threads=None):
    """Per-bit error rates and their summary (evaluator/capi.c:497-574):
    bits[j] = fraction of samples whose bit j differs; mean_hamming =
    1 - mean error; ber_min/ber_max = 1 - min/max per-bit error; ber_std =
    sample standard deviation of the per-bit errors."""
    if _on_device(predicted, expected):
        import ctypes
        from . import capi
        bits = (ctypes.c_double * n_hidden)()
        summ = (ctypes.c_double * 4)()
        from . import _device_buffer
        predicted, expected = predicted.contiguous(), expected.contiguous()
        nb = n_samples * ((n_hidden + 7) // 8)
        pp = _device_buffer(predicted, np.uint8, nb, "encoding_accuracy: predicted")
        ep = _device_buffer(expected, np.uint8, nb, "encoding_accuracy: expected")
        _dev_call(capi.load_library().tm_encoding_accuracy_dev, pp, ep, n_samples, n_hidden, bits, summ)
        return {"bits": list(bits), "mean_hamming": summ[0], "ber_min": summ[1], "ber_max": summ[2],
                "ber_std": summ[3]}
    chunks = (n_hidden + 7) // 8
    pr = _as_array(predicted, np.uint8)[: n_samples * chunks].reshape(n_samples, chunks)
    ex = _as_array(expected, np.uint8)[: n_samples * chunks].reshape(n_samples, chunks)
    diff = np.unpackbits(pr ^ ex, axis=1, bitorder="little")[:, :n_hidden]
    bdiff = diff.sum(0).astype(np.float64)
    ber = bdiff / n_samples
    # the reference divides by n_samples * n_dims as an unsigned int product
    mean = float(bdiff.sum() / ((n_samples * n_hidden) & 0xFFFFFFFF))
    var = 0.0
    for j in range(n_hidden):                 # in bit order (evaluator/capi.c:560-564)
        var += (float(ber[j]) - mean) * (float(ber[j]) - mean)
    std = float(np.sqrt(var / (n_hidden - 1))) if n_hidden > 1 else 0.0
    return {"bits": [float(b) for b in ber], "mean_hamming": 1.0 - mean,
            "ber_min": 1.0 - float(min(1.0, ber.min())), "ber_max": 1.0 - float(max(0.0, ber.max())), "ber_std": std}
