"""Circuit: the gate container consumed by the simulators.

API of QuICT/core/circuit/circuit.py:30-668 that the simulation path and its
users touch: ``Circuit(wires)``, ``gate | circuit(q)``, ``circuit.append``,
``extend``, ``random_append`` (same RNG draw order), ``qubits``, ``width()``,
``size()``, ``depth()``, ``qasm()``, ``convert_precision()``,
``get_lastcall_for_each_qubits()``, ``matrix()``.
"""
from typing import List, Union

import numpy as np

from .gate import BasicGate, build_random_gate, GATE_TYPE_TO_CLASS
from .gate_type import GateType, MatrixType
from .operator import NoiseGate, Operator, Trigger
from .qubit import Qureg
from .composite_gate import CompositeGate
from .utils import unique_id_generator, matrix_product_to_circuit
from . import exception as qerr


_DEFAULT_RANDOM_TYPES = [
    GateType.rx, GateType.ry, GateType.rz, GateType.cx, GateType.cy, GateType.crz,
    GateType.ch, GateType.cz, GateType.rxx, GateType.ryy, GateType.rzz, GateType.fsim,
]


class Circuit:
    def __init__(self, wires: Union[int, Qureg], name: str = None, topology=None, fidelity: float = None,
                 ancillae_qubits: List[int] = None):
        self._name = name if name is not None else "circuit_" + unique_id_generator()
        self._gates: list = []
        self._gate_type = {}
        self._pointer = None
        self._precision = np.complex128
        self._qubits = wires if isinstance(wires, Qureg) else Qureg(int(wires))
        self._ancillae_qubits = list(ancillae_qubits or [])
        self.topology = topology
        self.fidelity = fidelity

    # --- properties --------------------------------------------------------
    @property
    def name(self) -> str:
        return self._name

    @name.setter
    def name(self, v: str):
        self._name = v

    @property
    def qubits(self) -> Qureg:
        return self._qubits

    @property
    def gates(self) -> list:
        return self._gates

    @gates.setter
    def gates(self, g):
        self._gates = g

    @property
    def precision(self):
        return self._precision

    @property
    def ancillae_qubits(self) -> List[int]:
        return self._ancillae_qubits

    def width(self) -> int:
        return len(self._qubits)

    def size(self) -> int:
        return len(self._gates)

    def depth(self) -> int:
        d = np.zeros(self.width(), dtype=int)
        for g in self._gates:
            a = g.cargs + g.targs
            if a:
                d[a] = np.max(d[a]) + 1
        return int(np.max(d)) if self.width() else 0

    def count_1qubit_gate(self) -> int:
        return sum(1 for g in self._gates if isinstance(g, BasicGate) and g.controls + g.targets == 1)

    def count_2qubit_gate(self) -> int:
        return sum(1 for g in self._gates if isinstance(g, BasicGate) and g.controls + g.targets == 2)

    def count_gate_by_gatetype(self, gate_type: GateType) -> int:
        return self._gate_type.get(gate_type, 0)

    def __str__(self):
        return str({"name": self.name, "width": self.width(), "size": self.size(), "depth": self.depth(),
                    "1-qubit gates": self.count_1qubit_gate(), "2-qubit gates": self.count_2qubit_gate()})

    # --- qubit selection syntax ---------------------------------------------
    def __call__(self, indexes):
        if isinstance(indexes, (int, np.integer)):
            indexes = [int(indexes)]
        if isinstance(indexes, (list, tuple)):
            for i in indexes:
                if not 0 <= int(i) < self.width():
                    raise qerr.IndexExceedError("Circuit qubit index out of range", i)
            self._pointer = [int(i) for i in indexes]
        elif isinstance(indexes, Qureg):
            self._pointer = [self._qubits.index(q) for q in indexes]
        else:
            raise qerr.TypeError("Circuit(...) takes int / list[int] / Qureg")
        return self

    def __getitem__(self, item):
        return self._qubits[item]

    # --- building -------------------------------------------------------------
    def _record(self, gate, insert_idx: int = -1):
        if insert_idx == -1:
            self._gates.append(gate)
        else:
            self._gates.insert(insert_idx, gate)
        if isinstance(gate, BasicGate):
            self._gate_type[gate.type] = self._gate_type.get(gate.type, 0) + 1

    def _add_gate(self, gate: BasicGate, pointer, insert_idx: int):
        nargs = gate.controls + gate.targets
        assigned = gate.cargs + gate.targs
        if pointer is None:
            if not assigned:
                if nargs == 1:
                    for q in range(self.width()):
                        g = gate.copy()
                        g.targs = [q]
                        self._record(g)
                    return
                if nargs == self.width():
                    args = list(range(nargs))
                else:
                    raise qerr.CircuitAppendError(f"{gate.type.name} needs qubits assigned")
            else:
                args = assigned
        else:
            if len(pointer) < nargs:
                raise qerr.CircuitAppendError("fewer qubits given than the gate acts on")
            args = [pointer[a] for a in assigned] if (len(pointer) > nargs and assigned) else pointer[:nargs]
        for a in args:
            if not 0 <= a < self.width():
                raise qerr.CircuitAppendError(f"qubit {a} outside a {self.width()}-qubit circuit")
        g = gate.copy()
        g._cargs = list(args[:gate.controls])
        g._targs = list(args[gate.controls:gate.controls + gate.targets])
        if self._precision != g.precision:
            g.convert_precision()
        self._record(g, insert_idx)

    def append(self, op, is_extend: bool = False, insert_idx: int = -1):
        pointer = self._pointer
        if not is_extend:
            self._pointer = None
        if isinstance(op, BasicGate):
            self._add_gate(op, pointer, insert_idx)
        elif isinstance(op, Trigger):
            if pointer is not None:
                if len(pointer) != op.targets:
                    raise qerr.CircuitAppendError("Trigger qubit count mismatch")
                op.targs = list(pointer)
            elif not op.targs:
                raise qerr.CircuitAppendError("Trigger needs qubits assigned")
            self._record(op)
        elif isinstance(op, NoiseGate):
            if pointer is not None:
                op.targs = list(pointer)
            self._record(op, insert_idx)
        elif isinstance(op, CompositeGate):
            self.extend(op)
        elif isinstance(op, Operator):
            self._record(op)
        else:
            raise qerr.TypeError("Circuit.append takes BasicGate / Trigger / NoiseGate, got", type(op))

    def extend(self, gates):
        pointer = self._pointer
        if isinstance(gates, CompositeGate):
            gates = gates.gates
        for g in gates:
            if pointer is not None and isinstance(g, BasicGate):
                g = g.copy()
                a = g.cargs + g.targs
                new = [pointer[x] for x in a]
                g._cargs, g._targs = new[:g.controls], new[g.controls:]
            self.append(g, is_extend=True)
        self._pointer = None

    def __or__(self, other):
        if not isinstance(other, Circuit):
            raise qerr.TypeError("circuit | circuit only")
        other.extend(self._gates)

    def random_append(self, rand_size: int = 10, typelist: list = None, random_params: bool = False,
                      probabilities: list = None):
        """Append ``rand_size`` random gates.

        Same RNG consumption as QuICT/core/circuit/circuit.py:442-488 per gate:
        ``np.random.choice(range(len(typelist)), p=probabilities)`` for the type,
        ``random.sample(range(n), arity)`` for the qubits (controls first), then
        ``np.random.uniform(0, 2*pi, params)`` when ``random_params``.
        """
        if typelist is None:
            typelist = list(_DEFAULT_RANDOM_TYPES)
        bad = set(typelist) & {GateType.unitary, GateType.perm, GateType.perm_fx}
        if bad:
            raise qerr.CircuitSpecialAppendError(f"{bad} not supported by random_append")
        if probabilities is not None:
            if not np.isclose(sum(probabilities), 1, atol=1e-6):
                raise qerr.ValueError("probabilities must sum to 1")
            if len(probabilities) != len(typelist):
                raise qerr.CircuitSpecialAppendError("probabilities and typelist lengths differ")
        idxs = list(range(len(typelist)))
        n = self.width()
        for _ in range(rand_size):
            t = np.random.choice(idxs, p=probabilities)
            self.append(build_random_gate(typelist[t], n, random_params))

    # --- conversions -----------------------------------------------------------
    def convert_precision(self):
        for g in self._gates:
            if hasattr(g, "convert_precision"):
                g.convert_precision()
        self._precision = np.complex64 if self._precision == np.complex128 else np.complex128

    def get_lastcall_for_each_qubits(self) -> List[GateType]:
        last = [None] * self.width()
        seen = 0
        for g in reversed(self._gates):
            for a in g.cargs + g.targs:
                if last[a] is None:
                    last[a] = g.type
                    seen += 1
            if seen == self.width():
                break
        return last

    def qasm(self, output_file: str = None) -> str:
        creg = min(self.count_gate_by_gatetype(GateType.measure), self.width()) or self.width()
        s = 'OPENQASM 2.0;\ninclude "qelib1.inc";\n' + f"qreg q[{self.width()}];\ncreg c[{creg}];\n"
        cbit = 0
        for g in self._gates:
            if isinstance(g, BasicGate) and g.type == GateType.measure:
                s += f"measure q[{g.targ}] -> c[{cbit}];\n"
                cbit = (cbit + 1) % creg
            elif isinstance(g, BasicGate):
                s += g.qasm()
        if output_file:
            with open(output_file, "w") as f:
                f.write(s)
        return s

    def matrix(self, device: str = "CPU"):
        """The circuit's unitary (QuICT/core/utils/circuit_info.py:177-200); gates
        of matrix type ``special`` are skipped.  device="GPU": built by the HIP
        library from the identity with fused tile passes and returned as a
        DeviceArray (``.get()`` -> numpy); device="CPU": host product of the
        expanded gate matrices (small circuits)."""
        if device == "GPU":
            from ..simulation.unitary import circuit_matrix
            return circuit_matrix(self, "double" if self._precision == np.complex128 else "single")
        if device != "CPU":
            raise qerr.ValueError("Circuit.matrix.device must be CPU or GPU, got", device)
        n = self.width()
        u = np.identity(1 << n, dtype=np.complex128)
        for g in self._gates:
            if isinstance(g, BasicGate) and g.matrix_type == MatrixType.special:
                continue
            u = matrix_product_to_circuit(g.matrix.astype(np.complex128), g.cargs + g.targs, n) @ u
        return u.astype(self._precision)

    def reset_qubits(self):
        self._qubits.reset_qubits()
