"""The C ABI on the CPU (no GPU call): libtsetlin_amd.so loads and exports
every function include/tsetlin_amd.h declares, the ctypes table binds each of
them, and the device-pointer paths refuse malformed buffers before any kernel
could read past them."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tsetlin_amd.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(tm_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_the_entry_points():
    names = declared()
    for n in ("tm_create_classifier", "tm_train_classifier", "tm_predict_classifier", "tm_persist", "tm_load",
              "tm_checkpoint", "tm_restore", "tm_reconfigure", "tm_type", "tm_destroy", "tm_counts",
              "tm_batch_class_accuracy_dev", "tm_mlp_train"):
        assert n in names, n
    assert len(names) >= 50


def test_library_exports_every_declared_symbol():
    from santoku.tsetlin import capi
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_binds_every_declared_symbol():
    from santoku.tsetlin import capi
    from santoku.tsetlin import surrogate
    lib = capi.load_library()
    surrogate._lib()
    mlp = [n for n in declared() if n.startswith("tm_mlp_")]
    unbound = [n for n in declared() if n not in capi.EXPORTS and n not in mlp]
    assert not unbound, unbound
    for n in capi.EXPORTS:
        assert getattr(lib, n).argtypes is not None or n == "tm_last_error", n


class FakeTensor:
    """Stands in for a GPU tensor in the argument checks (no device needed)."""

    def __init__(self, dtype, numel, contiguous=True, itemsize=1):
        self.dtype, self._n, self._c, self._s = dtype, numel, contiguous, itemsize
        self.is_cuda = True

    def numel(self):
        return self._n

    def element_size(self):
        return self._s

    def is_contiguous(self):
        return self._c

    def data_ptr(self):
        return 0x1000


FakeTensor.__module__ = "torch"
FakeTensor.__name__ = "Tensor"


def test_device_buffer_checks():
    import santoku.tsetlin as tm
    torch = pytest.importorskip("torch")
    ok = FakeTensor("torch.uint8", 100)
    assert tm._device_buffer(ok, np.uint8, 100, "x") == 0x1000
    with pytest.raises(tm.TsetlinError, match="expected a GPU tensor"):
        tm._device_buffer(torch.zeros(10, dtype=torch.uint8), np.uint8, 10, "x")
    with pytest.raises(tm.TsetlinError, match="dtype int64"):
        tm._device_buffer(FakeTensor("torch.int32", 100, itemsize=4), np.int64, 400, "labels")
    with pytest.raises(tm.TsetlinError, match="contiguous"):
        tm._device_buffer(FakeTensor("torch.uint8", 100, contiguous=False), np.uint8, 100, "rows")
    with pytest.raises(tm.TsetlinError, match="holds 99 bytes, the call reads 100"):
        tm._device_buffer(FakeTensor("torch.uint8", 99), np.uint8, 100, "rows")
    with pytest.raises(tm.TsetlinError, match="holds 72 bytes"):
        tm._device_buffer(FakeTensor("torch.int64", 9, itemsize=8), np.int64, 80, "labels")


def test_width_limits_refused_before_any_device_call():
    """Tolerances from 2^25 and targets from 2^30 exceed the kernels' 32-bit
    vote and fire arithmetic: refused with TM_E_UNSUPPORTED while the options
    are checked, before the device is touched (without a GPU, a machine that
    passes these checks fails later with TM_E_DEVICE)."""
    import santoku.tsetlin as tm
    from conftest import gpu_available
    base = dict(classes=2, features=8, clauses=8, clause_maximum=3, specificity=2.0)
    with pytest.raises(tm.TsetlinError, match="TM_E_UNSUPPORTED.*clause_tolerance 33554432"):
        tm.classifier(dict(base, clause_tolerance=1 << 25))
    with pytest.raises(tm.TsetlinError, match="TM_E_UNSUPPORTED.*target"):
        tm.classifier(dict(base, clause_tolerance=4, target=float(1 << 30)))
    # the default target sqrt(clauses / 2) * tolerance can reach the limit too
    with pytest.raises(tm.TsetlinError, match="TM_E_UNSUPPORTED.*target"):
        tm.classifier(dict(base, clauses=1 << 20, clause_tolerance=(1 << 25) - 1))
    if not gpu_available():
        with pytest.raises(tm.TsetlinError, match="TM_E_DEVICE"):
            tm.classifier(dict(base, clause_tolerance=(1 << 25) - 1, target=float((1 << 30) - 1)))


def test_sizes_that_would_wrap_are_refused():
    """Clause and class counts whose rounding to whole bytes would wrap 32 bits
    are refused as arguments, before the device is touched."""
    import santoku.tsetlin as tm
    base = dict(classes=2, features=8, clause_tolerance=3, clause_maximum=3, specificity=2.0)
    for o in (dict(base, clauses=0xFFFFFFFF), dict(base, clauses=8, classes=0xFFFFFFFE)):
        with pytest.raises(tm.TsetlinError, match="TM_E_ARG.*too large"):
            tm.classifier(o)
