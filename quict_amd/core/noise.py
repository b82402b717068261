"""Noise channels and the noise model.

Channel definitions follow QuICT/core/noise/noise_error.py:
  * ``QuantumNoiseError(ops)``: Kraus operators ``sqrt(p_i) * M_i`` (:60-71),
    ``tensor``/``compose``/``expand`` (:90-130), ``apply_to_gate`` = ``K_i @ G`` (:132-138).
  * ``PauliError`` / ``BitflipError`` / ``PhaseflipError`` / ``PhaseBitflipError`` (:149-236).
  * ``DepolarizingError(p, n)``: probabilities ``[1 - p/max, p/4^n, ...]`` with
    ``max = 4^n / (4^n - 1)`` over Pauli strings in ``product('ixyz')`` order (:239-263).
  * ``DampingError(amp, phase, dissipation)``: the six operators of :266-338,
    keeping those with non-zero probability, each with weight 1.
``NoiseModel.transpile`` follows QuICT/core/noise/noise_model.py:152-185.
"""
from enum import Enum
from itertools import product
from typing import List, Tuple, Union

import numpy as np

from .gate import BasicGate, GATE_TYPE_TO_CLASS, Unitary, ID, X, Y, Z
from .gate_type import GateType
from .operator import NoiseGate
from . import exception as qerr


class NoiseChannel(Enum):
    unitary = "unitary channel"
    pauil = "pauli channel"
    depolarizing = "depolarizing channel"
    damping = "damping channel"
    readout = "readout channel"


def is_kraus_ops(kraus: List[np.ndarray]) -> bool:
    r, c = kraus[0].shape
    if r != c or (r & (r - 1)):
        return False
    if any(k.shape != (r, c) for k in kraus):
        return False
    s = sum(k.conj().T @ k for k in kraus)
    return bool(np.allclose(s, np.identity(r), rtol=1e-4))


class QuantumNoiseError:
    # single-qubit channels whose tensor product this channel is (set by ``tensor``)
    _factors = None

    def __init__(self, ops: List[Tuple[np.ndarray, float]]):
        if not isinstance(ops, list):
            raise qerr.TypeError("QuantumNoiseError expects list[(matrix, prob)]")
        self._set_ops(ops)
        self._type = NoiseChannel.unitary

    def _set_ops(self, ops):
        self._operators = ops
        self._kraus_operators = [np.sqrt(p) * m for m, p in ops]
        if not is_kraus_ops(self._kraus_operators):
            raise qerr.KrausError("operators do not form a Kraus set")
        self._kraus_conj_ops = [k.T.conj() for k in self._kraus_operators]
        self._qubits = int(np.log2(self._kraus_operators[0].shape[0]))
        self._precision = ops[0][0].dtype

    @property
    def kraus(self) -> List[np.ndarray]:
        return self._kraus_operators

    @property
    def kraus_ct(self) -> List[np.ndarray]:
        return self._kraus_conj_ops

    @property
    def operators(self):
        return self._operators

    @operators.setter
    def operators(self, ops):
        self._set_ops(ops)

    @property
    def qubits(self) -> int:
        return self._qubits

    @property
    def precision(self):
        return self._precision

    @property
    def type(self) -> NoiseChannel:
        return self._type

    @type.setter
    def type(self, t: NoiseChannel):
        self._type = t

    def expand(self, extend_qubits: int) -> "QuantumNoiseError":
        eye = np.identity(1 << extend_qubits, dtype=self._precision)
        return QuantumNoiseError([(np.kron(m, eye), p) for m, p in self._operators])

    def compose(self, other: "QuantumNoiseError") -> "QuantumNoiseError":
        left, right = self.operators, other.operators
        if self.qubits > other.qubits:
            right = other.expand(self.qubits - other.qubits).operators
        elif self.qubits < other.qubits:
            left = self.expand(other.qubits - self.qubits).operators
        return QuantumNoiseError([(a @ b, p * q) for a, p in left for b, q in right])

    def tensor(self, other: "QuantumNoiseError") -> "QuantumNoiseError":
        out = QuantumNoiseError([(np.kron(a, b), p * q) for a, p in self._operators for b, q in other.operators])
        out._factors = (self._factors or [self]) + (other._factors or [other])
        return out

    @property
    def factors(self) -> List["QuantumNoiseError"]:
        """The channels this one is the tensor product of (first = first qubit), or [self]."""
        return list(self._factors) if self._factors else [self]

    def apply_to_gate(self, matrix: np.ndarray) -> List[np.ndarray]:
        if matrix.shape != (1 << self._qubits, 1 << self._qubits):
            raise qerr.NoiseApplyError("gate size differs from the noise's qubit count")
        return [k @ matrix for k in self.kraus]

    def prob_mapping_operator(self, prob: float) -> np.ndarray:
        for m, p in self._operators:
            if prob < p:
                return m
            prob -= p
        return self._operators[-1][0]

    def superoperator(self) -> np.ndarray:
        return sum(np.kron(k, k.conj()) for k in self.kraus)


_PAULI = {"i": ID.matrix, "x": X.matrix, "y": Y.matrix, "z": Z.matrix}


class PauliError(QuantumNoiseError):
    def __init__(self, ops: List[Tuple[str, float]], num_qubits: int = 1):
        if not isinstance(ops, (list, tuple)):
            raise qerr.TypeError("PauliError ops must be list/tuple")
        out = []
        for name, p in ops:
            if p < 0 or p > 1:
                raise qerr.ValueError("PauliError probability outside [0, 1]", p)
            if len(name) != num_qubits and num_qubits > 1:
                raise qerr.PauliNoiseUnmatchedError("Pauli string length differs from qubit count")
            m = None
            for ch in name:
                if ch not in _PAULI:
                    raise qerr.ValueError("Pauli letters are i/x/y/z", ch)
                if m is None:
                    m = _PAULI[ch]
                elif num_qubits > 1:
                    m = np.kron(m, _PAULI[ch])
                else:
                    m = m @ _PAULI[ch]
            out.append((m, p))
        super().__init__(out)
        self._type = NoiseChannel.pauil


class BitflipError(PauliError):
    def __init__(self, prob: float):
        super().__init__([("i", 1 - prob), ("x", prob)])


class PhaseflipError(PauliError):
    def __init__(self, prob: float):
        super().__init__([("i", 1 - prob), ("z", prob)])


class PhaseBitflipError(PauliError):
    def __init__(self, prob: float):
        super().__init__([("i", 1 - prob), ("y", prob)])


class DepolarizingError(PauliError):
    def __init__(self, prob: float, num_qubits: int = 1):
        if not isinstance(prob, float):
            raise qerr.TypeError("DepolarizingError.prob must be float")
        if not isinstance(num_qubits, int) or num_qubits < 1:
            raise qerr.ValueError("DepolarizingError.num_qubits must be >= 1")
        nops = 4 ** num_qubits
        pmax = nops / (nops - 1)
        if prob < 0 or prob > pmax:
            raise qerr.ValueError(f"DepolarizingError.prob must be in [0, {pmax}]", prob)
        probs = [1 - prob / pmax] + [prob / nops] * (nops - 1)
        names = ["".join(t) for t in product("ixyz", repeat=num_qubits)]
        super().__init__(list(zip(names, probs)), num_qubits)
        self._type = NoiseChannel.depolarizing


class DampingError(QuantumNoiseError):
    def __init__(self, amplitude_prob: float, phase_prob: float, dissipation_state: float = 0.0):
        for v in (amplitude_prob, phase_prob, dissipation_state):
            if not 0 <= v <= 1:
                raise qerr.ValueError("DampingError probabilities must be in [0, 1]")
        if amplitude_prob + phase_prob > 1:
            raise qerr.DampingNoiseMixedProbExceedError("amplitude + phase damping must be <= 1")
        self.amplitude_prob = amplitude_prob
        self.phase_prob = phase_prob
        self.dissipation_state = dissipation_state
        m0, m1 = np.sqrt(1 - dissipation_state), np.sqrt(dissipation_state)
        sa, sp = np.sqrt(amplitude_prob), np.sqrt(phase_prob)
        sap = np.sqrt(1 - amplitude_prob - phase_prob)
        mats = [m0 * np.array([[1, 0], [0, sap]], dtype=np.complex128),
                m0 * np.array([[0, sa], [0, 0]], dtype=np.complex128),
                m0 * np.array([[0, 0], [0, sp]], dtype=np.complex128),
                m1 * np.array([[sap, 0], [0, 1]], dtype=np.complex128),
                m1 * np.array([[0, 0], [sa, 0]], dtype=np.complex128),
                m1 * np.array([[sp, 0], [0, 0]], dtype=np.complex128)]
        weights = [a * b for a, b in product([1 - dissipation_state, dissipation_state],
                                             [1 - amplitude_prob - phase_prob, amplitude_prob, phase_prob])]
        super().__init__([(m, 1) for m, w in zip(mats, weights) if w != 0])
        self._type = NoiseChannel.damping


class ReadoutError:
    """Classical readout flip matrix P[i, j] = Pr(read j | true i)."""

    def __init__(self, prob: np.ndarray):
        prob = np.asarray(prob, dtype=np.float64)
        if prob.ndim != 2 or prob.shape[0] != prob.shape[1]:
            raise qerr.ValueError("ReadoutError needs a square matrix")
        if not np.allclose(prob.sum(axis=1), 1):
            raise qerr.ValueError("ReadoutError rows must sum to 1")
        self._prob = prob
        self._qubits = int(np.log2(prob.shape[0]))

    @property
    def prob(self):
        return self._prob

    @property
    def qubits(self) -> int:
        return self._qubits

    def apply_to_qubits(self, state: int) -> int:
        u = np.random.random()
        row = self._prob[state]
        for j, p in enumerate(row):
            if u < p:
                return j
            u -= p
        return len(row) - 1


class NoiseModel:
    _BASE = ["h", "s", "sdg", "x", "y", "z", "rx", "ry", "rz", "u1", "t", "tdg", "cx", "cz", "crz", "cu1"]

    def __init__(self, name: str = "Noise Model", basic_gates: List[str] = None):
        self._name = name
        self._basic_gates = list(basic_gates) if basic_gates is not None else list(self._BASE)
        self._error_by_gate = {g: [] for g in self._basic_gates}
        self._readout_errors = []

    def _check_gates(self, noise: QuantumNoiseError, gates):
        if not isinstance(noise, QuantumNoiseError):
            raise qerr.TypeError("noise must be a QuantumNoiseError")
        if not gates:
            raise qerr.ValueError("gates must not be empty")
        gates = [gates] if isinstance(gates, str) else list(gates)
        for g in gates:
            if g not in self._basic_gates:
                raise qerr.ValueError("gate not in this model's basic gates", g)
            inst = GATE_TYPE_TO_CLASS[GateType.__members__[g]]()
            if inst.controls + inst.targets != noise.qubits:
                raise qerr.NoiseApplyError(f"{g} acts on {inst.controls + inst.targets} qubits, "
                                           f"noise on {noise.qubits}")
        return gates

    def add(self, noise: QuantumNoiseError, gates: Union[str, List[str]], qubits: Union[int, List[int]] = None):
        if qubits is None:
            return self.add_noise_for_all_qubits(noise, gates)
        qubits = [qubits] if isinstance(qubits, int) else list(qubits)
        for g in self._check_gates(noise, gates):
            self._error_by_gate[g].append((noise, qubits))

    def add_noise_for_all_qubits(self, noise: QuantumNoiseError, gates: Union[str, List[str]]):
        for g in self._check_gates(noise, gates):
            self._error_by_gate[g].append((noise, -1))

    def add_readout_error(self, noise: ReadoutError, qubits: Union[int, List[int]]):
        qubits = [qubits] if isinstance(qubits, int) else list(qubits)
        if noise.qubits > 1 and noise.qubits != len(qubits):
            raise qerr.NoiseApplyError("multi-qubit readout error needs exactly that many qubits")
        self._readout_errors.append((noise, qubits))

    def transpile(self, circuit, accumulated_mode: bool = False):
        """Insert channels after (Kraus, ``NoiseGate``) or before (sampled Pauli
        ``Unitary``) every listed gate (noise_model.py:152-185)."""
        from .circuit import Circuit
        out = Circuit(circuit.width())
        if circuit.precision != out.precision:
            out.convert_precision()
        for gate in circuit.gates:
            if isinstance(gate, BasicGate) and gate.type.name in self._basic_gates:
                args = gate.cargs + gate.targs
                keep = True
                for noise, qubits in self._error_by_gate[gate.type.name]:
                    if qubits == -1 or set(qubits) & set(args) == set(args):
                        if accumulated_mode or noise.type == NoiseChannel.damping:
                            ng = NoiseGate(noise.apply_to_gate(gate.matrix.astype(np.complex128)), len(args))
                            ng.targs = list(args)
                            ng.source = (gate, noise)
                            out.append(ng)
                            keep = False
                        else:
                            m = noise.prob_mapping_operator(np.random.random())
                            u = Unitary(m)
                            out(list(args))
                            out.append(u)
                if keep:
                    out.append(gate)
            else:
                out.append(gate)
        return out

    def apply_readout_error(self, qureg):
        for err, qubits in self._readout_errors:
            if err.qubits == 1:
                for q in qubits:
                    if qureg[q].measured is None:
                        continue
                    qureg[q].measured = err.apply_to_qubits(int(qureg[q]))
            else:
                if any(qureg[q].measured is None for q in qubits):
                    continue
                v = err.apply_to_qubits(int(qureg[qubits]))
                for q in qubits[::-1]:
                    qureg[q].measured = v & 1
                    v >>= 1
