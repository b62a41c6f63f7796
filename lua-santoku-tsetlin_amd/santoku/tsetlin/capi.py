"""ctypes binding of libtsetlin_amd.so (the C ABI in include/tsetlin_amd.h).

The library is loaded from the package directory (built in place by
lua-santoku-tsetlin_amd/build.py).  There is no fallback: if the library is
missing or no GPU is present, calls raise.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.path.join(_PKG, "libtsetlin_amd.so")

TM_CLASSIFIER = 0
TM_ENCODER = 1
TM_OK = 0
ERRORS = {
    -1: "TM_E_ARG", -2: "TM_E_ALLOC", -3: "TM_E_STATE", -4: "TM_E_FORMAT",
    -5: "TM_E_SIZE", -6: "TM_E_DEVICE", -7: "TM_E_UNSUPPORTED",
}

EXPORTS = [
    "tm_options_default", "tm_create_classifier", "tm_create_encoder", "tm_train_classifier",
    "tm_train_encoder", "tm_train_classifier_dev", "tm_train_encoder_dev", "tm_predict_classifier",
    "tm_predict_encoder", "tm_predict_classifier_dev", "tm_predict_encoder_dev",
    "tm_predict_classifier_dev_noalloc", "tm_persist",
    "tm_persist_file", "tm_load", "tm_load_file", "tm_checkpoint", "tm_restore", "tm_checkpoint_dev",
    "tm_restore_dev", "tm_counts", "tm_reconfigure",
    "tm_type", "tm_destroy", "tm_get_info", "tm_last_error", "tm_last_timing", "tm_last_train_stats",
    "tm_last_queue_timing",
    "tm_create_classifier_ind", "tm_create_encoder_ind", "tm_train_classifier_ind", "tm_train_encoder_ind",
    "tm_predict_classifier_ind", "tm_predict_encoder_ind", "tm_train_classifier_ind_dev", "tm_train_encoder_ind_dev",
    "tm_predict_classifier_ind_dev", "tm_predict_encoder_ind_dev", "tm_restrict", "tm_get_feature_sizes",
    "tm_bits_to_cvec_dev", "tm_bits_to_cvec_ind_dev", "tm_flip_interleave_dev", "tm_class_accuracy_dev", "tm_encoding_accuracy_dev",
    "tm_batch_create", "tm_batch_machine", "tm_batch_train_classifier", "tm_batch_train_classifier_dev",
    "tm_batch_create_encoder", "tm_batch_train_encoder", "tm_batch_train_encoder_dev",
    "tm_batch_create_classifier_ind", "tm_batch_create_encoder_ind", "tm_batch_train_classifier_ind",
    "tm_batch_train_classifier_ind_dev", "tm_batch_train_encoder_ind", "tm_batch_train_encoder_ind_dev",
    "tm_batch_class_accuracy_ind_dev",
    "tm_batch_accuracy_dev", "tm_batch_class_accuracy_dev", "tm_batch_tuning", "tm_batch_timing", "tm_batch_destroy",
]


class Options(ctypes.Structure):
    _fields_ = [
        ("classes", ctypes.c_uint32),
        ("features", ctypes.c_uint32),
        ("clauses", ctypes.c_uint32),
        ("clause_tolerance", ctypes.c_uint32),
        ("clause_maximum", ctypes.c_uint32),
        ("state", ctypes.c_uint32),
        ("include_bits", ctypes.c_uint32),
        ("target", ctypes.c_double),
        ("negative", ctypes.c_double),
        ("specificity", ctypes.c_double),
        ("reusable", ctypes.c_int32),
        ("individualized", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
    ]


class Info(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "type", "classes", "features", "clauses", "clause_tolerance", "clause_maximum", "target",
        "state_bits", "include_bits", "input_bits", "input_chunks", "clause_chunks", "state_chunks",
        "action_chunks", "class_chunks")] + [
        ("specificity", ctypes.c_double), ("negative", ctypes.c_double),
        ("tail_mask", ctypes.c_uint8), ("reusable", ctypes.c_uint8), ("has_state", ctypes.c_uint8),
        ("trained", ctypes.c_uint8), ("individualized", ctypes.c_uint8)]


EPOCH_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32)
BATCH_EPOCH_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint8))

_lib = None


def load_library(path=LIB_PATH):
    """Load and bind the engine library (cached)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libtsetlin_amd.so is not built (%s); run lua-santoku-tsetlin_amd/build.py" % path)
    L = ctypes.CDLL(path)
    vp, u8p, i64p = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p
    opt_p = ctypes.POINTER(Options)
    m_pp = ctypes.POINTER(ctypes.c_void_p)
    sz_p = ctypes.POINTER(ctypes.c_size_t)
    u32 = ctypes.c_uint32
    sigs = {
        "tm_options_default": (None, [opt_p]),
        "tm_create_classifier": (ctypes.c_int, [opt_p, m_pp]),
        "tm_create_encoder": (ctypes.c_int, [opt_p, m_pp]),
        "tm_train_classifier": (ctypes.c_int, [vp, u32, u8p, i64p, u32, EPOCH_FN, vp]),
        "tm_train_encoder": (ctypes.c_int, [vp, u32, u8p, u8p, u32, EPOCH_FN, vp]),
        "tm_train_classifier_dev": (ctypes.c_int, [vp, u32, u8p, i64p, u32, EPOCH_FN, vp]),
        "tm_train_encoder_dev": (ctypes.c_int, [vp, u32, u8p, u8p, u32, EPOCH_FN, vp]),
        "tm_predict_classifier": (ctypes.c_int, [vp, u8p, u32, i64p]),
        "tm_predict_encoder": (ctypes.c_int, [vp, u8p, u32, u8p]),
        "tm_predict_classifier_dev": (ctypes.c_int, [vp, u8p, u32, i64p]),
        "tm_predict_encoder_dev": (ctypes.c_int, [vp, u8p, u32, u8p]),
        "tm_predict_classifier_dev_noalloc": (ctypes.c_int, [vp, u8p, u32, i64p]),
        "tm_persist": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, sz_p]),
        "tm_persist_file": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "tm_load": (ctypes.c_int, [u8p, ctypes.c_size_t, m_pp]),
        "tm_load_file": (ctypes.c_int, [ctypes.c_char_p, m_pp]),
        "tm_checkpoint": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, sz_p]),
        "tm_restore": (ctypes.c_int, [vp, u8p, ctypes.c_size_t]),
        "tm_checkpoint_dev": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, sz_p]),
        "tm_restore_dev": (ctypes.c_int, [vp, u8p, ctypes.c_size_t]),
        "tm_counts": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, sz_p]),
        "tm_reconfigure": (ctypes.c_int, [vp, opt_p]),
        "tm_type": (ctypes.c_int, [vp]),
        "tm_destroy": (None, [vp]),
        "tm_get_info": (ctypes.c_int, [vp, ctypes.POINTER(Info)]),
        "tm_last_error": (ctypes.c_char_p, []),
        "tm_last_timing": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "tm_last_train_stats": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_uint64)]),
        "tm_last_queue_timing": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double)]),
        "tm_create_classifier_ind": (ctypes.c_int, [opt_p, i64p, u32, m_pp]),
        "tm_create_encoder_ind": (ctypes.c_int, [opt_p, i64p, u32, m_pp]),
        "tm_train_classifier_ind": (ctypes.c_int, [vp, u32, u8p, ctypes.c_size_t, i64p, i64p, u32, EPOCH_FN, vp]),
        "tm_train_encoder_ind": (ctypes.c_int, [vp, u32, u8p, ctypes.c_size_t, u8p, i64p, u32, EPOCH_FN, vp]),
        "tm_predict_classifier_ind": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, i64p, u32, i64p]),
        "tm_predict_encoder_ind": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, i64p, u32, u8p]),
        "tm_train_classifier_ind_dev": (ctypes.c_int, [vp, u32, vp, ctypes.c_size_t, vp, i64p, u32, EPOCH_FN, vp]),
        "tm_train_encoder_ind_dev": (ctypes.c_int, [vp, u32, vp, ctypes.c_size_t, vp, i64p, u32, EPOCH_FN, vp]),
        "tm_predict_classifier_ind_dev": (ctypes.c_int, [vp, vp, ctypes.c_size_t, i64p, u32, vp]),
        "tm_predict_encoder_ind_dev": (ctypes.c_int, [vp, vp, ctypes.c_size_t, i64p, u32, vp]),
        "tm_restrict": (ctypes.c_int, [vp, i64p, u32]),
        "tm_get_feature_sizes": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_uint64), u32]),
        "tm_bits_to_cvec_dev": (ctypes.c_int, [vp, ctypes.c_uint64, u32, u32, vp, ctypes.POINTER(ctypes.c_uint64)]),
        "tm_bits_to_cvec_ind_dev": (ctypes.c_int, [vp, i64p, i64p, u32, u32, vp, ctypes.c_uint64, i64p,
                                                   ctypes.POINTER(ctypes.c_uint64)]),
        "tm_flip_interleave_dev": (ctypes.c_int, [vp, u32, u32, vp]),
        "tm_class_accuracy_dev": (ctypes.c_int, [vp, vp, u32, u32, ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double)]),
        "tm_encoding_accuracy_dev": (ctypes.c_int, [vp, vp, u32, u32, ctypes.POINTER(ctypes.c_double),
                                                    ctypes.POINTER(ctypes.c_double)]),
        "tm_batch_create": (ctypes.c_int, [opt_p, u32, m_pp]),
        "tm_batch_machine": (ctypes.c_int, [vp, u32, m_pp]),
        "tm_batch_train_classifier": (ctypes.c_int, [vp, u32, u8p, i64p, u32, BATCH_EPOCH_FN, vp]),
        "tm_batch_train_classifier_dev": (ctypes.c_int, [vp, u32, u8p, i64p, u32, BATCH_EPOCH_FN, vp]),
        "tm_batch_create_encoder": (ctypes.c_int, [opt_p, u32, m_pp]),
        "tm_batch_train_encoder": (ctypes.c_int, [vp, u32, u8p, u8p, u32, BATCH_EPOCH_FN, vp]),
        "tm_batch_train_encoder_dev": (ctypes.c_int, [vp, u32, u8p, u8p, u32, BATCH_EPOCH_FN, vp]),
        "tm_batch_create_classifier_ind": (ctypes.c_int, [opt_p, u32, i64p, u32, m_pp]),
        "tm_batch_create_encoder_ind": (ctypes.c_int, [opt_p, u32, i64p, u32, m_pp]),
        "tm_batch_train_classifier_ind": (ctypes.c_int, [vp, u32, u8p, ctypes.c_size_t, i64p, i64p, u32,
                                                         BATCH_EPOCH_FN, vp]),
        "tm_batch_train_classifier_ind_dev": (ctypes.c_int, [vp, u32, u8p, ctypes.c_size_t, i64p, i64p, u32,
                                                             BATCH_EPOCH_FN, vp]),
        "tm_batch_train_encoder_ind": (ctypes.c_int, [vp, u32, u8p, ctypes.c_size_t, u8p, i64p, u32,
                                                      BATCH_EPOCH_FN, vp]),
        "tm_batch_train_encoder_ind_dev": (ctypes.c_int, [vp, u32, u8p, ctypes.c_size_t, u8p, i64p, u32,
                                                          BATCH_EPOCH_FN, vp]),
        "tm_batch_class_accuracy_ind_dev": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, i64p, i64p, u32,
                                                           ctypes.POINTER(ctypes.c_double)]),
        "tm_batch_accuracy_dev": (ctypes.c_int, [vp, u8p, i64p, u32, ctypes.POINTER(ctypes.c_double)]),
        "tm_batch_class_accuracy_dev": (ctypes.c_int, [vp, u8p, i64p, u32, ctypes.POINTER(ctypes.c_double)]),
        "tm_batch_tuning": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                           ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double)]),
        "tm_batch_timing": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "tm_batch_destroy": (None, [vp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L
