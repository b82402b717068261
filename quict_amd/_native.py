"""ctypes binding of libqsim_hip.so (the C ABI declared in include/qsim.h).

The library is built in-tree (``quict_amd/libqsim_hip.so``) by
``__graft_entry__.build()`` / ``make -C quict_amd/csrc``.  There is no CPU
fallback: if the library or a HIP device is missing, every simulator call
raises ``NativeRuntimeError``.

torch is imported before the library is opened so that the process uses one
HIP runtime (torch's libamdhip64.so.7 satisfies the library's dependency by
SONAME); torch also owns the device buffers and streams passed in here.
"""
import ctypes
import os
import threading

import numpy as np

from .core import exception as qerr

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libqsim_hip.so")

QS_OK = 0
QS_ERR_ARG = -1
QS_ERR_DEVICE = -2
QS_ERR_UNSUPPORTED = -3
QS_C64 = 0
QS_C128 = 1
QS_MAX_GATE_BITS = 16

GATE_DENSE = 0
GATE_DIAG = 1
GATE_PHASE = 2
GATE_XPERM = 3
GATE_SWAP = 4


class QsGate(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("ncontrols", ctypes.c_int32),
        ("ntargets", ctypes.c_int32),
        ("bits", ctypes.c_int32 * QS_MAX_GATE_BITS),
        ("m", ctypes.c_double * 32),
        ("big", ctypes.c_void_p),
    ]


# numpy mirror of QsGate, used to build gate arrays without per-field ctypes calls
QS_GATE_DTYPE = np.dtype({
    "names": ["kind", "ncontrols", "ntargets", "bits", "m", "big"],
    "formats": ["<i4", "<i4", "<i4", ("<i4", (QS_MAX_GATE_BITS,)), ("<f8", (32,)), "<u8"],
    "offsets": [QsGate.kind.offset, QsGate.ncontrols.offset, QsGate.ntargets.offset, QsGate.bits.offset,
                QsGate.m.offset, QsGate.big.offset],
    "itemsize": ctypes.sizeof(QsGate),
})


class QsPauliTerm(ctypes.Structure):
    _fields_ = [("coef", ctypes.c_double), ("xmask", ctypes.c_uint64), ("zmask", ctypes.c_uint64)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_U64 = ctypes.c_uint64

# name -> (restype, argtypes)
SIGNATURES = {
    "qs_ctx_create": (_P, [_I]),
    "qs_ctx_destroy": (None, [_P]),
    "qs_last_error": (ctypes.c_char_p, []),
    "qs_ctx_scratch_bytes": (_U64, [_P]),
    "qs_set_fusion": (_I, [_P, _I, _I]),
    "qs_set_fusion_launch": (_I, [_P, _I, _I]),
    "qs_profile_enable": (_I, [_P, _I]),
    "qs_profile_read": (_I, [_P, _P, _P, _I]),
    "qs_stream_copy": (_I, [_P, _P, _P, _U64, _I, _P]),
    "qs_sv_init": (_I, [_P, _P, _I, _I, _U64, _P]),
    "qs_sv_apply_gate": (_I, [_P, _P, _I, _I, _P, _P]),
    "qs_sv_apply_gates": (_I, [_P, _P, _I, _I, _P, _I, _I, ctypes.POINTER(_I), _P]),
    "qs_sv_prob_zero": (_I, [_P, _P, _I, _I, _I, ctypes.POINTER(_D), _P]),
    "qs_sv_collapse": (_I, [_P, _P, _I, _I, _I, _I, _D, _P]),
    "qs_sv_reset": (_I, [_P, _P, _I, _I, _I, _D, _P]),
    "qs_sv_sample": (_I, [_P, _P, _I, _I, _P, _I, _P, _P]),
    "qs_sv_sample_noalloc": (_I, [_P, _P, _I, _I, _P, _I, _P, _P]),
    "qs_sample_probs": (_I, [_P, _P, _I, _P, _I, _P, _P]),
    "qs_sv_probabilities": (_I, [_P, _P, _I, _I, _P, _P]),
    "qs_sv_norm2": (_I, [_P, _P, _I, _I, ctypes.POINTER(_D), _P]),
    "qs_sv_permute": (_I, [_P, _P, _P, _I, _I, _P, _P]),
    "qs_dm_diagonal": (_I, [_P, _P, _I, _I, _P, _P]),
    "qs_dm_trace": (_I, [_P, _P, _I, _I, ctypes.POINTER(_D), ctypes.POINTER(_D), _P]),
    "qs_dm_prob_zero": (_I, [_P, _P, _I, _I, _I, ctypes.POINTER(_D), _P]),
    "qs_dm_collapse": (_I, [_P, _P, _I, _I, _I, _I, _D, _P]),
    "qs_graph_capture": (_I, [_P, _P, _I, _I, _P, _I, _I, ctypes.POINTER(_P)]),
    "qs_graph_launch": (_I, [_P, _P]),
    "qs_graph_passes": (_I, [_P]),
    "qs_graph_destroy": (None, [_P]),
    "qs_mat_identity": (_I, [_P, _P, _I, _I, _P]),
    "qs_mat_transpose": (_I, [_P, _P, _I, _I, _P]),
    "qs_mat_vec": (_I, [_P, _P, _I, _I, _I, _P, _P, _P]),
    "qs_sv_apply_pauli_sum": (_I, [_P, _P, _P, _I, _I, _P, _I, _P]),
    "qs_sv_inner": (_I, [_P, _P, _P, _I, _I, ctypes.POINTER(_D), ctypes.POINTER(_D), _P]),
    "qs_adjoint_gradient": (_I, [_P, _P, _P, _I, _I, _P, _I, _P, _P, _P, _I, _P]),
    "qs_adjoint_gradient_pauli": (_I, [_P, _P, _P, _I, _I, _P, _I, _P, _P, _I, _P, _P, _P, _I, _P, _P]),
    "qs_qnn_batch": (_I, [_P, _I, _I, _P, _I, _P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "qs_qnn_batch_programs": (_I, [_P, _I, _I, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
}


class QsQnnConfig(ctypes.Structure):
    _fields_ = [("lds_max_width", ctypes.c_int32), ("block_threads", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32), ("use_graph", ctypes.c_int32), ("fuse_qubits", ctypes.c_int32)]

class QsQnnOp(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("q0", ctypes.c_int32), ("q1", ctypes.c_int32),
                ("param", ctypes.c_int32), ("value", ctypes.c_double)]


QS_QNN_RX, QS_QNN_RY, QS_QNN_RZ, QS_QNN_CX, QS_QNN_CZ, QS_QNN_ENC_RY = range(6)

_lib = None
_lib_lock = threading.Lock()
_contexts = {}


def library_path() -> str:
    return LIB_PATH


def load_library():
    """Open libqsim_hip.so (once).  Raises NativeRuntimeError if it is missing."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        import torch  # noqa: F401  (shares its HIP runtime with the library)
        if not os.path.exists(LIB_PATH):
            raise qerr.NativeRuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C quict_amd/csrc`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def exported_symbols():
    return list(SIGNATURES.keys())


def check(rc: int, what: str = ""):
    if rc == QS_OK:
        return
    msg = _lib.qs_last_error().decode(errors="replace") if _lib is not None else ""
    text = f"{what}: {msg}" if what else msg
    if rc == QS_ERR_ARG:
        raise qerr.ValueError(text)
    if rc == QS_ERR_UNSUPPORTED:
        raise qerr.GateAlgorithmNotImplementError(text)
    raise qerr.NativeRuntimeError(text)


def context(device: int = 0):
    """The (lazily created) library context for a HIP device."""
    lib = load_library()
    ctx = _contexts.get(device)
    if ctx is None:
        import torch
        if not torch.cuda.is_available():
            raise qerr.NativeRuntimeError("no HIP device visible: the simulator has no CPU fallback")
        torch.cuda.set_device(device)
        ctx = lib.qs_ctx_create(device)
        if not ctx:
            raise qerr.NativeRuntimeError(f"qs_ctx_create({device}) failed: {lib.qs_last_error().decode()}")
        _contexts[device] = ctx
    return ctx


def stream_ptr(device: int = 0):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def dtype_code(torch_dtype) -> int:
    import torch
    if torch_dtype == torch.complex128:
        return QS_C128
    if torch_dtype == torch.complex64:
        return QS_C64
    raise qerr.TypeError("state dtype must be complex64 or complex128, got", torch_dtype)


def gate_array_ptr(arr: np.ndarray):
    assert arr.dtype == QS_GATE_DTYPE and arr.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(arr.ctypes.data) if len(arr) else ctypes.c_void_p(0)
