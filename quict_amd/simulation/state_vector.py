"""StateVectorSimulator on MI355X.

Drop-in for QuICT's ``StateVectorSimulator(device="GPU")``
(QuICT/simulation/state_vector/statevector_simulator.py:26-626):
``run(circuit, state_vector=None, use_previous=False)``, ``sample(shots)``,
``vector``, measure / reset / perm / trigger handling and the precision and
device-id options.  The amplitude vector is a torch-ROCm tensor on the chosen
device; every operation on it is a HIP kernel in libqsim_hip.so.

Differences in mechanism (same results):
  * consecutive unitary gates are compiled to one record array and applied by
    one library call, which fuses them into LDS/register-resident tile passes
    (``fusion=False`` restores one pass per gate);
  * ``sample`` draws the reference's uniforms (CPython ``random.random()``,
    shot-major, qubit 0 first) and resolves all shots by descending a
    probability pyramid on the device, instead of collapsing and restoring the
    state per shot.
"""
import ctypes
import random
from typing import List, Optional

import numpy as np

from ..core import BasicGate, Circuit, CompositeGate, GateType, Trigger
from ..core import exception as qerr
from .. import _native as nat
from .device_array import DeviceArray
from .gate_compiler import compile_gates


class CapturedCircuit:
    """A unitary circuit planned, uploaded and captured once in a HIP graph
    (``qs_graph_capture``).  Every ``replay()`` re-simulates it from |0...0> on
    its own device vector with no host-side compilation or planning; the
    simulator that made it can then ``sample`` the result."""

    def __init__(self, sim, circuit: Circuit, tensor, handle, passes: int):
        self._sim = sim
        self._circuit = circuit
        self._tensor = tensor          # the graph writes this buffer: keep it alive
        self._handle = handle
        self.passes = passes

    def replay(self) -> DeviceArray:
        sim = self._sim
        nat.check(sim._lib.qs_graph_launch(self._handle, sim._stream()), "qs_graph_launch")
        sim._vector = self._tensor
        sim._circuit = self._circuit
        sim._qubits = int(self._circuit.width())
        if sim._sync:
            sim._torch.cuda.synchronize(sim._device_id)
        return DeviceArray(self._tensor)

    def close(self):
        if self._handle:
            self._sim._lib.qs_graph_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StateVectorSimulator:
    __DEVICE = ["GPU"]
    __PRECISION = ["single", "double"]

    def __init__(self, device: str = "GPU", precision: str = "double", gpu_device_id: int = 0,
                 matrix_aggregation: bool = True, sync: bool = True, fusion: bool = True):
        if device not in self.__DEVICE:
            raise qerr.ValueError("StateVectorSimulator.device must be 'GPU' (this build runs on HIP only), got",
                                  device)
        if precision not in self.__PRECISION:
            raise qerr.ValueError("StateVectorSimulator.precision must be single/double, got", precision)
        import torch
        self._torch = torch
        self._device = device
        self._device_id = gpu_device_id
        self._precision = np.complex128 if precision == "double" else np.complex64
        self._tdtype = torch.complex128 if precision == "double" else torch.complex64
        self._dcode = nat.QS_C128 if precision == "double" else nat.QS_C64
        self._sync = sync
        self._fusion = fusion
        self._using_matrix_aggregation = matrix_aggregation  # accepted; matrices travel inline in gate records
        self._circuit: Optional[Circuit] = None
        self._vector = None
        self._qubits = 0
        self.last_passes = 0
        self._ctx = nat.context(gpu_device_id)
        self._lib = nat.load_library()

    # --- properties ----------------------------------------------------------
    @property
    def circuit(self):
        return self._circuit

    @circuit.setter
    def circuit(self, c):
        self._circuit = c

    @property
    def device(self):
        return self._device_id

    @property
    def vector(self) -> DeviceArray:
        return DeviceArray(self._vector) if self._vector is not None else None

    @vector.setter
    def vector(self, vec):
        self._vector = self._to_device(vec)

    def _to_device(self, vec):
        torch = self._torch
        if isinstance(vec, DeviceArray):
            vec = vec.tensor
        if isinstance(vec, torch.Tensor):
            return vec.to(device=f"cuda:{self._device_id}", dtype=self._tdtype).contiguous()
        return torch.as_tensor(np.asarray(vec, dtype=self._precision), device=f"cuda:{self._device_id}")

    # --- helpers ---------------------------------------------------------------
    def _stream(self):
        return nat.stream_ptr(self._device_id)

    def _vptr(self):
        return ctypes.c_void_p(self._vector.data_ptr())

    def initial_circuit(self, circuit: Circuit):
        self._circuit = circuit
        self._qubits = int(circuit.width())
        if self._precision != circuit.precision:
            circuit.convert_precision()
        self._pipeline = list(circuit.gates)

    def initial_state_vector(self, all_zeros: bool = False):
        n = self._qubits
        self._vector = self._torch.empty(1 << n, dtype=self._tdtype, device=f"cuda:{self._device_id}")
        rc = self._lib.qs_sv_init(self._ctx, self._vptr(), n, self._dcode, 0, self._stream())
        nat.check(rc, "qs_sv_init")
        if all_zeros:
            self._vector.zero_()

    # --- run -------------------------------------------------------------------
    def run(self, circuit: Circuit, state_vector=None, use_previous: bool = False) -> DeviceArray:
        self.initial_circuit(circuit)
        if state_vector is not None:
            sv = state_vector.tensor if isinstance(state_vector, DeviceArray) else state_vector
            size = sv.numel() if isinstance(sv, self._torch.Tensor) else np.asarray(sv).size
            if size != (1 << self._qubits):
                raise qerr.StateVectorUnmatchedError("the state vector must have 2^width amplitudes")
            self.vector = sv
        elif not use_previous or self._vector is None or self._vector.numel() != (1 << self._qubits):
            self.initial_state_vector()
        self.last_passes = 0
        pending: List[BasicGate] = []
        idx = 0
        while idx < len(self._pipeline):
            op = self._pipeline[idx]
            idx += 1
            if isinstance(op, BasicGate) and op.type not in (GateType.measure, GateType.reset, GateType.perm):
                pending.append(op)
                continue
            self._flush(pending)
            pending = []
            if isinstance(op, Trigger):
                self.apply_trigger(op, idx)
            elif isinstance(op, BasicGate):
                self.apply_gate(op)
            else:
                raise qerr.TypeError("StateVectorSimulator.run takes BasicGate / Trigger, got", type(op))
        self._flush(pending)
        if self._sync:
            self._torch.cuda.synchronize(self._device_id)
        return self.vector

    def capture(self, circuit: Circuit) -> CapturedCircuit:
        """Plan and capture a circuit of unitary gates once; ``replay()`` then runs
        it from |0...0> as one HIP graph launch (no per-run compilation)."""
        ops = list(circuit.gates)
        for op in ops:
            if not isinstance(op, BasicGate) or op.type in (GateType.measure, GateType.reset, GateType.perm):
                raise qerr.TypeError("capture takes unitary gates only, got", getattr(op, "type", type(op)))
        if self._precision != circuit.precision:
            circuit.convert_precision()
        n = int(circuit.width())
        recs, keep = compile_gates(ops, n)
        tensor = self._torch.empty(1 << n, dtype=self._tdtype, device=f"cuda:{self._device_id}")
        handle = ctypes.c_void_p(0)
        rc = self._lib.qs_graph_capture(self._ctx, ctypes.c_void_p(tensor.data_ptr()), n, self._dcode,
                                        nat.gate_array_ptr(recs), len(recs), 1, ctypes.byref(handle))
        del keep
        nat.check(rc, "qs_graph_capture")
        return CapturedCircuit(self, circuit, tensor, handle, self._lib.qs_graph_passes(handle))

    def _flush(self, gates: List[BasicGate]):
        if not gates:
            return
        recs, keep = compile_gates(gates, self._qubits)
        if len(recs) == 0:
            return
        passes = ctypes.c_int(0)
        rc = self._lib.qs_sv_apply_gates(self._ctx, self._vptr(), self._qubits, self._dcode,
                                         nat.gate_array_ptr(recs), len(recs), 2 if self._fusion else 0,
                                         ctypes.byref(passes), self._stream())
        del keep
        nat.check(rc, "qs_sv_apply_gates")
        self.last_passes += passes.value

    def apply_gate(self, gate: BasicGate):
        """One gate (statevector_simulator.py:202-315)."""
        n = self._qubits
        gt = gate.type
        if gt == GateType.measure:
            idx = n - 1 - gate.targ
            self._circuit.qubits[gate.targ].measured = self._measure(idx)
        elif gt == GateType.reset:
            idx = n - 1 - gate.targ
            p0 = self.get_measured_prob(idx)
            nat.check(self._lib.qs_sv_reset(self._ctx, self._vptr(), n, self._dcode, idx, p0, self._stream()),
                      "qs_sv_reset")
        elif gt == GateType.perm:
            self._apply_perm(gate)
        else:
            self._flush([gate])

    def _apply_perm(self, gate: BasicGate):
        n = self._qubits
        args = gate.cargs + gate.targs
        if not args:
            args = list(range(n))
        mapping = list(range(n))
        for i, p in enumerate(gate.pargs):
            mapping[args[i]] = args[p]
        # VectorPermutation semantics: source qubit i holds the output's qubit mapping[i];
        # the kernel takes, per output bit b, the source bit it comes from.
        inv = [0] * n
        for i, m in enumerate(mapping):
            inv[m] = i
        perm = np.array([n - 1 - inv[n - 1 - b] for b in range(n)], dtype=np.int32)
        out = self._torch.empty_like(self._vector)
        rc = self._lib.qs_sv_permute(self._ctx, self._vptr(), ctypes.c_void_p(out.data_ptr()), n, self._dcode,
                                     perm.ctypes.data_as(ctypes.c_void_p), self._stream())
        nat.check(rc, "qs_sv_permute")
        self._vector = out

    def get_measured_prob(self, index: int, all_measured: bool = False) -> float:
        """P(amplitude bit ``index`` == 0) (measured_prob_calculate, gate_function.py:1751-1776)."""
        p0 = ctypes.c_double(0)
        if all_measured:
            rc = self._lib.qs_sv_norm2(self._ctx, self._vptr(), self._qubits, self._dcode, ctypes.byref(p0),
                                       self._stream())
        else:
            rc = self._lib.qs_sv_prob_zero(self._ctx, self._vptr(), self._qubits, self._dcode, index,
                                           ctypes.byref(p0), self._stream())
        nat.check(rc, "qs_sv_prob_zero")
        return p0.value

    def _measure(self, index: int) -> int:
        """apply_measuregate (gate_function.py:1779-1815): outcome = random() > P0."""
        p0 = self.get_measured_prob(index)
        if self._precision == np.complex64:
            p0 = float(np.float32(p0))
        outcome = int(random.random() > p0)
        prob = (1 - p0) if outcome else p0
        scale = 1 / np.sqrt(prob)
        if self._precision == np.complex64:
            scale = float(np.float32(scale))
        rc = self._lib.qs_sv_collapse(self._ctx, self._vptr(), self._qubits, self._dcode, index, outcome,
                                      float(scale), self._stream())
        nat.check(rc, "qs_sv_collapse")
        return outcome

    def apply_specialgate(self, index: int, type: GateType, prob: float = None):
        if type == GateType.measure:
            r = self._measure(index)
            self._circuit.qubits[self._qubits - 1 - index].measured = r
            return r
        if type == GateType.reset:
            p0 = self.get_measured_prob(index) if prob is None else prob
            nat.check(self._lib.qs_sv_reset(self._ctx, self._vptr(), self._qubits, self._dcode, index, p0,
                                            self._stream()), "qs_sv_reset")
            return None
        raise qerr.ValueError("apply_specialgate takes measure / reset")

    def apply_trigger(self, op: Trigger, current_idx: int):
        """statevector_simulator.py:498-523: measure, then splice the mapped gates."""
        for t in op.targs:
            self._circuit.qubits[t].measured = self._measure(self._qubits - 1 - t)
        cg = op.mapping(int(self._circuit.qubits[op.targs]))
        if cg is None:
            return
        if getattr(cg, "checkpoint", None) is not None:
            raise qerr.GateTypeNotImplementError("Trigger mappings with a checkpoint position")
        # a single BasicGate is applied too (the reference silently drops it)
        gates = cg.gates if isinstance(cg, CompositeGate) else [cg]
        self._pipeline = self._pipeline[:current_idx] + list(gates) + self._pipeline[current_idx:]

    def apply_zeros(self):
        self._vector.zero_()

    # --- sampling --------------------------------------------------------------
    def sample_indices(self, shots: int) -> np.ndarray:
        """Outcome (amplitude index) of every shot, same draws as the reference."""
        if self._circuit is None or self._vector is None:
            raise qerr.SampleBeforeRunError("StateVectorSimulator sample without run any circuit.")
        n = self._qubits
        u = np.array([random.random() for _ in range(shots * n)], dtype=np.float64)
        out = np.zeros(shots, dtype=np.uint64)
        rc = self._lib.qs_sv_sample(self._ctx, self._vptr(), n, self._dcode, u.ctypes.data_as(ctypes.c_void_p),
                                    shots, out.ctypes.data_as(ctypes.c_void_p), self._stream())
        nat.check(rc, "qs_sv_sample")
        if shots:
            last = int(out[-1])
            for q in range(n):
                self._circuit.qubits[q].measured = (last >> (n - 1 - q)) & 1
        return out

    def sample(self, shots: int = 1):
        """Counts of every basis state (list of 2^n ints, statevector_simulator.py:595-626).
        For more than 24 qubits a dict {index: count} of the non-zero entries is returned."""
        out = self.sample_indices(shots)
        n = self._qubits
        if n > 24:
            vals, cnt = np.unique(out, return_counts=True)
            return {int(v): int(c) for v, c in zip(vals, cnt)}
        counts = np.bincount(out.astype(np.int64), minlength=1 << n)
        return counts.tolist()
