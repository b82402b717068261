"""UnitarySimulator and circuit matrices on MI355X.

Drop-in for QuICT's ``UnitarySimulator`` (QuICT/simulation/unitary/
unitary_simulator.py:12-157): ``run(circuit_or_matrix, state_vector=None,
use_previous=False)`` returns the state ``U x`` (``x`` = |0...0> unless given),
``sample(shots)`` measures it, ``vector`` holds it; and for ``Circuit.matrix``
(QuICT/core/utils/circuit_info.py:177-200 -> CircuitMatrix,
QuICT/core/utils/circuit_matrix.py:69-125).

Mechanism: the reference groups gates by depth, expands every group to a
dense matrix with tensor / MatrixTensorI / MatrixPermutation and multiplies the
groups with cuBLAS GEMMs, O(8^n) flops for the final products.  Here the
unitary is the circuit applied to all 2^n basis states at once: the buffer is
W = U^T (row c = U|c>, i.e. column c of U), initialised to the identity, and
the circuit's gates act on the low n bits of that 2n-bit vector, fused into
tile passes by the same HIP kernels as the state-vector simulator
(O(passes * 4^n) bytes, HBM-bound).  ``U x`` is one matrix-vector pass over W;
the row-major U is one in-place transpose away and is produced only when the
matrix itself is asked for.  Gates of matrix type ``special`` (measure, reset,
barrier, perm) are skipped, as CircuitMatrix does (circuit_matrix.py:82-83).

Differences (documented, not reproduced): the reference skips the product when
``allclose(U, I)`` (unitary_simulator.py:100, 108-110); here ``U x`` is always
computed, which returns ``x`` bit for bit when U is exactly the identity and
differs by at most allclose's tolerance otherwise.
"""
import ctypes
import random
from typing import List, Union

import numpy as np

from ..core import BasicGate, Circuit, CompositeGate, MatrixType
from ..core import exception as qerr
from .. import _native as nat
from .device_array import DeviceArray
from .gate_compiler import compile_gates


def _unitary_gates(gates) -> List[BasicGate]:
    out = []
    for g in gates:
        if not isinstance(g, BasicGate):
            raise qerr.TypeError("a circuit matrix takes BasicGate operations only, got", type(g))
        if g.matrix_type == MatrixType.special:
            continue
        out.append(g)
    return out


def build_columns(gates, n: int, dtype_code: int, device_id: int = 0, fusion: bool = True):
    """W = U^T of the gate list as a flat torch tensor of 4^n amplitudes
    (row c of W = U|c>).  Returns (W, passes)."""
    import torch
    lib = nat.load_library()
    ctx = nat.context(device_id)
    stream = nat.stream_ptr(device_id)
    if n < 1 or n > 20:
        raise qerr.ValueError("circuit matrix width must be in [1, 20], got", n)
    tdt = torch.complex128 if dtype_code == nat.QS_C128 else torch.complex64
    W = torch.empty(1 << (2 * n), dtype=tdt, device=f"cuda:{device_id}")
    nat.check(lib.qs_mat_identity(ctx, ctypes.c_void_p(W.data_ptr()), n, dtype_code, stream), "qs_mat_identity")
    recs, keep = compile_gates(_unitary_gates(gates), n)
    passes = ctypes.c_int(0)
    if len(recs):
        rc = lib.qs_sv_apply_gates(ctx, ctypes.c_void_p(W.data_ptr()), 2 * n, dtype_code, nat.gate_array_ptr(recs),
                                   len(recs), 1 if fusion else 0, ctypes.byref(passes), stream)
        nat.check(rc, "qs_sv_apply_gates(circuit matrix)")
    del keep
    return W, passes.value


def transpose_inplace(W, n: int, dtype_code: int, device_id: int = 0):
    lib = nat.load_library()
    nat.check(lib.qs_mat_transpose(nat.context(device_id), ctypes.c_void_p(W.data_ptr()), n, dtype_code,
                                   nat.stream_ptr(device_id)), "qs_mat_transpose")
    return W


def circuit_matrix(circuit: Union[Circuit, CompositeGate], precision: str = "double", device_id: int = 0,
                   fusion: bool = True) -> DeviceArray:
    """Row-major 2^n x 2^n unitary of a circuit on the GPU (Circuit.matrix(device="GPU"))."""
    n = int(circuit.width())
    code = nat.QS_C128 if precision == "double" else nat.QS_C64
    W, _ = build_columns(circuit.gates, n, code, device_id, fusion)
    transpose_inplace(W, n, code, device_id)
    return DeviceArray(W, (1 << n, 1 << n))


class UnitarySimulator:
    """unitary_simulator.py:12-157 on the GPU."""
    __DEVICE = ["GPU"]
    __PRECISION = ["single", "double"]

    def __init__(self, device: str = "GPU", precision: str = "double", gpu_device_id: int = 0,
                 fusion: bool = True):
        if device not in self.__DEVICE:
            raise qerr.ValueError("UnitarySimulator.device must be 'GPU' (this build runs on HIP only), got", device)
        if precision not in self.__PRECISION:
            raise qerr.ValueError("UnitarySimulator.precision must be single/double, got", precision)
        import torch
        self._torch = torch
        self._device = device
        self._device_id = gpu_device_id
        self._precision = np.complex128 if precision == "double" else np.complex64
        self._tdtype = torch.complex128 if precision == "double" else torch.complex64
        self._dcode = nat.QS_C128 if precision == "double" else nat.QS_C64
        self._fusion = fusion
        self._vector = None
        self._circuit = None
        self._qubits_num = 0
        self._W = None            # U^T (columns layout) after a circuit run
        self._U = None            # row-major U (given matrix, or W transposed on demand)
        self.last_passes = 0
        self._ctx = nat.context(gpu_device_id)
        self._lib = nat.load_library()

    # --- properties --------------------------------------------------------------
    @property
    def vector(self) -> DeviceArray:
        return DeviceArray(self._vector) if self._vector is not None else None

    @vector.setter
    def vector(self, vec):
        self._vector = self._to_device(vec)

    @property
    def unitary_matrix(self) -> DeviceArray:
        """The row-major unitary of the last run (transposed in place from the
        column layout the first time it is asked for)."""
        if self._U is None and self._W is not None:
            transpose_inplace(self._W, self._qubits_num, self._dcode, self._device_id)
            self._U, self._W = self._W, None
        if self._U is None:
            return None
        D = 1 << self._qubits_num
        return DeviceArray(self._U, (D, D))

    def _to_device(self, vec):
        torch = self._torch
        if isinstance(vec, DeviceArray):
            vec = vec.tensor
        if isinstance(vec, torch.Tensor):
            return vec.reshape(-1).to(device=f"cuda:{self._device_id}", dtype=self._tdtype).contiguous()
        return torch.as_tensor(np.asarray(vec, dtype=self._precision).ravel(), device=f"cuda:{self._device_id}")

    def _stream(self):
        return nat.stream_ptr(self._device_id)

    def initial_vector_state(self):
        """unitary_simulator.py:49-57: |0...0>."""
        n = self._qubits_num
        self._vector = self._torch.empty(1 << n, dtype=self._tdtype, device=f"cuda:{self._device_id}")
        nat.check(self._lib.qs_sv_init(self._ctx, ctypes.c_void_p(self._vector.data_ptr()), n, self._dcode, 0,
                                       self._stream()), "qs_sv_init")

    # --- run -----------------------------------------------------------------------
    def run(self, circuit: Union[np.ndarray, Circuit], state_vector=None, use_previous: bool = False) -> DeviceArray:
        """unitary_simulator.py:59-106."""
        if isinstance(circuit, Circuit):
            self._circuit = circuit
            if self._precision != circuit.precision:
                circuit.convert_precision()
            self._qubits_num = int(circuit.width())
            self._W, self.last_passes = build_columns(circuit.gates, self._qubits_num, self._dcode,
                                                      self._device_id, self._fusion)
            self._U = None
            mat, trans = self._W, 1
        else:
            m = circuit.tensor if isinstance(circuit, DeviceArray) else circuit
            shape = tuple(m.shape)
            if len(shape) != 2 or shape[0] != shape[1] or shape[0] < 2 or (shape[0] & (shape[0] - 1)):
                raise qerr.UnitaryMatrixUnmatchedError("The unitary matrix should be square.")
            self._circuit = None
            self._qubits_num = int(shape[0]).bit_length() - 1
            self._U = self._to_device(m)
            self._W = None
            self.last_passes = 0
            mat, trans = self._U, 0

        n = self._qubits_num
        if state_vector is not None:
            sv = state_vector.tensor if isinstance(state_vector, DeviceArray) else state_vector
            size = sv.numel() if isinstance(sv, self._torch.Tensor) else np.asarray(sv).size
            if size != (1 << n):
                raise qerr.StateVectorUnmatchedError("The state vector should has the same qubits with the circuit.")
            self.vector = sv
        elif not use_previous or self._vector is None or self._vector.numel() != (1 << n):
            self.initial_vector_state()

        out = self._torch.empty(1 << n, dtype=self._tdtype, device=f"cuda:{self._device_id}")
        rc = self._lib.qs_mat_vec(self._ctx, ctypes.c_void_p(mat.data_ptr()), n, self._dcode, trans,
                                  ctypes.c_void_p(self._vector.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                  self._stream())
        nat.check(rc, "qs_mat_vec")
        self._vector = out
        self._torch.cuda.synchronize(self._device_id)
        return self.vector

    # --- sample ----------------------------------------------------------------------
    def sample(self, shots: int):
        """unitary_simulator.py:112-157: per shot, measure qubits 0..n-1 in order
        with CPython ``random.random()`` draws (outcome 1 iff u > P0) and restore
        the state; resolved here by one descent of a probability pyramid per shot
        (same draws and thresholds)."""
        if self._vector is None:
            raise qerr.SampleBeforeRunError("UnitarySimulation sample without run any circuit.")
        n = self._qubits_num
        u = np.array([random.random() for _ in range(shots * n)], dtype=np.float64)
        out = np.zeros(shots, dtype=np.uint64)
        rc = self._lib.qs_sv_sample(self._ctx, ctypes.c_void_p(self._vector.data_ptr()), n, self._dcode,
                                    u.ctypes.data_as(ctypes.c_void_p), shots, out.ctypes.data_as(ctypes.c_void_p),
                                    self._stream())
        nat.check(rc, "qs_sv_sample")
        if shots and self._circuit is not None:
            last = int(out[-1])
            for q in range(n):
                self._circuit.qubits[q].measured = (last >> (n - 1 - q)) & 1
        counts = np.bincount(out.astype(np.int64), minlength=1 << n)
        return counts.tolist()
