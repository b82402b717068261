"""DensityMatrixSimulator on MI355X.

Drop-in for QuICT's ``DensityMatrixSimulator`` (QuICT/simulation/density_matrix/
density_matrix_simulator.py:14-240): ``run(circuit, density_matrix=None,
noise_model=None, use_previous=False)``, ``sample(shots)``, measure, the
``accumulated_mode`` option and readout errors.

Mechanism: the reference builds a dense 2^n x 2^n unitary per gate group and
multiplies rho <- U rho U^dagger with two cuBLAS GEMMs, and expands every Kraus
operator to 2^n x 2^n on the host (O(4^n) per operator).  Here rho (row-major,
entry (r, c) at index r*2^n + c) is treated as a 2n-bit vector:

  * a gate U on qubit q acts as U on row bit 2n-1-q and conj(U) on column bit
    n-1-q — two gate records that the fusion planner usually puts in the same
    tile pass, so each gate costs at most one read+write of rho;
  * a channel {K_i} on qubits A acts as the dense superoperator
    sum_i K_i (x) conj(K_i) on the 2|A| bits [row bits of A, column bits of A];
  * measurement: P0 from the diagonal, then rho <- P rho P / p.
"""
import ctypes
import random
from typing import List

import numpy as np

from ..core import BasicGate, Circuit, GateType, NoiseGate
from ..core import exception as qerr
from ..core.noise import NoiseModel
from .. import _native as nat
from .device_array import DeviceArray
from .gate_compiler import compile_gate, superop_record


def op_records(g, n: int, keep: list) -> list:
    """Records of one unitary gate or noise channel on the 2n-bit image of rho
    (row bits n..2n-1, column bits 0..n-1): a gate U is U on its row bits and
    conj(U) on its column bits; a channel is its superoperator.  A tensor-product
    channel applied after a gate (accumulated mode) is the gate's two records
    followed by one superoperator per single-qubit factor."""
    out = []

    def unitary(gate):
        row = np.zeros(1, dtype=nat.QS_GATE_DTYPE)[0]
        col = np.zeros(1, dtype=nat.QS_GATE_DTYPE)[0]
        if compile_gate(gate, n, row, keep, row_offset=n):
            compile_gate(gate, n, col, keep, row_offset=0, conj=True)
            out.extend([row, col])

    if isinstance(g, NoiseGate):
        factors = g.source[1].factors if g.source is not None else None
        if factors and len(factors) > 1 and len(factors) == len(g.targs) and all(f.qubits == 1 for f in factors):
            unitary(g.source[0])
            for f, q in zip(factors, g.targs):
                rec = np.zeros(1, dtype=nat.QS_GATE_DTYPE)[0]
                superop_record(f.kraus, [q], n, rec, keep)
                out.append(rec)
        else:
            rec = np.zeros(1, dtype=nat.QS_GATE_DTYPE)[0]
            superop_record(g.noise_matrix, list(g.targs), n, rec, keep)
            out.append(rec)
    else:
        unitary(g)
    return out


def program_records(ops, n: int):
    """Record array of a measurement-free list of gates and channels (the program
    the density-matrix simulator hands to qs_sv_apply_gates on 2n bits)."""
    keep: list = []
    recs = []
    for g in ops:
        recs.extend(op_records(g, n, keep))
    arr = np.zeros(len(recs), dtype=nat.QS_GATE_DTYPE)
    for i, r in enumerate(recs):
        arr[i] = r
    return arr, keep


class DensityMatrixSimulator:
    def __init__(self, device: str = "GPU", precision: str = "double", accumulated_mode: bool = False,
                 gpu_device_id: int = 0, fusion: bool = True):
        if device != "GPU":
            raise qerr.ValueError("DensityMatrixSimulator.device must be 'GPU' (HIP only), got", device)
        if precision not in ("single", "double"):
            raise qerr.ValueError("DensityMatrixSimulator.precision must be single/double, got", precision)
        import torch
        self._torch = torch
        self._device = device
        self._device_id = gpu_device_id
        self._precision = np.complex128 if precision == "double" else np.complex64
        self._tdtype = torch.complex128 if precision == "double" else torch.complex64
        self._dcode = nat.QS_C128 if precision == "double" else nat.QS_C64
        self._accumulated_mode = accumulated_mode
        self._fusion = fusion
        self._density_matrix = None
        self._circuit = None
        self._noise_model = None
        self.last_passes = 0
        self._ctx = nat.context(gpu_device_id)
        self._lib = nat.load_library()

    @property
    def density_matrix(self) -> DeviceArray:
        D = 1 << self._qubits
        return DeviceArray(self._density_matrix, (D, D))

    def _stream(self):
        return nat.stream_ptr(self._device_id)

    def _rptr(self):
        return ctypes.c_void_p(self._density_matrix.data_ptr())

    def initial_circuit(self, circuit: Circuit, noise_model: NoiseModel):
        self._origin_circuit = circuit
        self._circuit = circuit if noise_model is None else noise_model.transpile(circuit, self._accumulated_mode)
        self._noise_model = noise_model
        self._qubits = int(circuit.width())

    def initial_density_matrix(self, qubits: int):
        D = 1 << qubits
        self._density_matrix = self._torch.zeros(D * D, dtype=self._tdtype, device=f"cuda:{self._device_id}")
        self._density_matrix[0] = 1

    def check_matrix(self, matrix) -> bool:
        """density_matrix_simulator.py:70-86: Hermitian, PSD, unit trace."""
        m = np.asarray(matrix.get() if isinstance(matrix, DeviceArray) else matrix)
        if not np.allclose(m.T.conj(), m):
            raise qerr.SimulationMatrixError("density matrix must equal its conjugate transpose")
        ev = np.linalg.eigvalsh(m.astype(np.complex128))
        if np.any((ev < 0) & ~np.isclose(ev, 0, rtol=1e-4)):
            raise qerr.SimulationMatrixError("density matrix eigenvalues must be non-negative")
        if not np.isclose(np.trace(m), 1, rtol=1e-4):
            raise qerr.SimulationMatrixError("density matrix trace must be 1")
        return True

    def run(self, circuit: Circuit, density_matrix=None, noise_model: NoiseModel = None,
            use_previous: bool = False) -> DeviceArray:
        self.initial_circuit(circuit, noise_model)
        if density_matrix is not None:
            self.check_matrix(density_matrix)
            m = np.asarray(density_matrix.get() if isinstance(density_matrix, DeviceArray) else density_matrix)
            self._density_matrix = self._torch.as_tensor(m.astype(self._precision).ravel(),
                                                         device=f"cuda:{self._device_id}")
        elif self._density_matrix is None or not use_previous or \
                self._density_matrix.numel() != (1 << (2 * self._qubits)):
            self.initial_density_matrix(self._qubits)
        self.last_passes = 0
        self._run(self._circuit)
        if noise_model is not None:
            noise_model.apply_readout_error(circuit.qubits)
        self._torch.cuda.synchronize(self._device_id)
        return self.density_matrix

    def _run(self, circuit: Circuit):
        n = self._qubits
        recs: List = []
        keep: list = []

        def flush():
            if not recs:
                return
            arr = np.zeros(len(recs), dtype=nat.QS_GATE_DTYPE)
            for i, r in enumerate(recs):
                arr[i] = r
            passes = ctypes.c_int(0)
            rc = self._lib.qs_sv_apply_gates(self._ctx, self._rptr(), 2 * n, self._dcode, nat.gate_array_ptr(arr),
                                             len(arr), 1 if self._fusion else 0, ctypes.byref(passes),
                                             self._stream())
            nat.check(rc, "qs_sv_apply_gates(density matrix)")
            self.last_passes += passes.value
            recs.clear()

        for g in circuit.gates:
            if isinstance(g, BasicGate) and g.type == GateType.measure:
                flush()
                self.apply_measure(g.targ)
            elif isinstance(g, BasicGate) and g.type in (GateType.reset, GateType.perm):
                raise qerr.GateTypeNotImplementError(f"{g.type.name} in the density-matrix simulator")
            elif isinstance(g, (BasicGate, NoiseGate)):
                recs.extend(op_records(g, n, keep))
            else:
                raise qerr.TypeError("DensityMatrixSimulator takes BasicGate / NoiseGate, got", type(g))
        flush()
        del keep

    def apply_measure(self, index: int):
        """density_matrix_simulator.py:182-208 (``index`` is the qubit)."""
        n = self._qubits
        bit = n - 1 - index
        p0 = ctypes.c_double(0)
        nat.check(self._lib.qs_dm_prob_zero(self._ctx, self._rptr(), n, self._dcode, bit, ctypes.byref(p0),
                                            self._stream()), "qs_dm_prob_zero")
        prob0 = p0.value
        outcome = int(random.random() > prob0)
        p = (1 - prob0) if outcome else prob0
        nat.check(self._lib.qs_dm_collapse(self._ctx, self._rptr(), n, self._dcode, bit, outcome, 1.0 / p,
                                           self._stream()), "qs_dm_collapse")
        self._circuit.qubits[index].measured = outcome
        return outcome

    def diagonal(self) -> np.ndarray:
        n = self._qubits
        d = self._torch.empty(1 << n, dtype=self._torch.float64, device=f"cuda:{self._device_id}")
        nat.check(self._lib.qs_dm_diagonal(self._ctx, self._rptr(), n, self._dcode, ctypes.c_void_p(d.data_ptr()),
                                           self._stream()), "qs_dm_diagonal")
        return d.cpu().numpy()

    def trace(self) -> complex:
        re, im = ctypes.c_double(0), ctypes.c_double(0)
        nat.check(self._lib.qs_dm_trace(self._ctx, self._rptr(), self._qubits, self._dcode, ctypes.byref(re),
                                        ctypes.byref(im), self._stream()), "qs_dm_trace")
        return complex(re.value, im.value)

    def sample(self, shots: int) -> list:
        """density_matrix_simulator.py:210-240 for the accumulated / noiseless case:
        measure every qubit per shot (CPython ``random``), count, restore rho.
        Trajectory mode (noise without accumulated_mode) re-runs the noisy
        circuit per shot, as the reference does."""
        if self._density_matrix is None:
            raise qerr.SampleBeforeRunError("DensityMatrixSimulator sample without run any circuit.")
        n = self._qubits
        last = self._circuit.get_lastcall_for_each_qubits()
        measured = [i for i in range(n) if last[i] != GateType.measure]
        counts = [0] * (1 << n)
        if self._accumulated_mode or self._noise_model is None:
            if len(measured) == n and self._noise_model is None or (len(measured) == n and
                                                                     not self._noise_model._readout_errors):
                u = np.array([random.random() for _ in range(shots * n)], dtype=np.float64)
                d = self._torch.empty(1 << n, dtype=self._torch.float64, device=f"cuda:{self._device_id}")
                nat.check(self._lib.qs_dm_diagonal(self._ctx, self._rptr(), n, self._dcode,
                                                   ctypes.c_void_p(d.data_ptr()), self._stream()), "qs_dm_diagonal")
                out = np.zeros(shots, dtype=np.uint64)
                nat.check(self._lib.qs_sample_probs(self._ctx, ctypes.c_void_p(d.data_ptr()), n,
                                                    u.ctypes.data_as(ctypes.c_void_p), shots,
                                                    out.ctypes.data_as(ctypes.c_void_p), self._stream()),
                          "qs_sample_probs")
                for v in out:
                    counts[int(v)] += 1
                if shots:
                    lastv = int(out[-1])
                    for q in range(n):
                        self._circuit.qubits[q].measured = (lastv >> (n - 1 - q)) & 1
                return counts
            original = self._density_matrix.clone()
            for _ in range(shots):
                for q in measured:
                    self.apply_measure(q)
                if self._noise_model is not None:
                    self._noise_model.apply_readout_error(self._circuit.qubits)
                counts[int(self._circuit.qubits)] += 1
                self._density_matrix.copy_(original)
            return counts
        for _ in range(shots):
            for q in measured:
                self.apply_measure(q)
            self._noise_model.apply_readout_error(self._circuit.qubits)
            counts[int(self._circuit.qubits)] += 1
            self.initial_density_matrix(n)
            noised = self._noise_model.transpile(self._origin_circuit, self._accumulated_mode)
            self._circuit = noised
            self._run(noised)
        return counts
