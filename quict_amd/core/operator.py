"""Circuit operators that are not plain gates.

* ``NoiseGate`` — a Kraus list acting on ``targs`` (QuICT/core/operator/noise_gate.py:6-35).
* ``Trigger``  — mid-circuit measurement that splices a CompositeGate chosen
  by the measured value (QuICT/core/operator/trigger.py:9-79).
"""
from typing import Callable, Dict, List, Union

import numpy as np

from . import exception as qerr


class Operator:
    def __init__(self, targets: int):
        if not isinstance(targets, int) or targets < 0:
            raise qerr.ValueError("Operator.targets must be a non-negative int")
        self._targets = targets
        self._targs: List[int] = []

    @property
    def targets(self) -> int:
        return self._targets

    @property
    def targs(self) -> List[int]:
        return self._targs

    @targs.setter
    def targs(self, t):
        self._targs = [int(t)] if isinstance(t, (int, np.integer)) else [int(x) for x in t]

    @property
    def cargs(self) -> List[int]:
        return []

    @property
    def targ(self) -> int:
        return self._targs[0]

    @property
    def controls(self) -> int:
        return 0

    def __and__(self, targets):
        self.targs = targets
        return self

    def __or__(self, circuit):
        circuit.append(self)


class NoiseGate(Operator):
    """A quantum channel given by its Kraus operators (already multiplied by
    the gate they follow, when built by ``NoiseModel.transpile``)."""

    def __init__(self, noise_matrix: List[np.ndarray], args_num: int):
        super().__init__(args_num)
        self._noise_matrix = [np.asarray(m) for m in noise_matrix]
        self._precision = self._noise_matrix[0].dtype
        # (gate, QuantumNoiseError) this channel was built from, when known: the
        # density-matrix simulator then applies the gate and the channel's
        # single-qubit factors separately (the same map, fewer operations)
        self.source = None

    @property
    def noise_matrix(self) -> List[np.ndarray]:
        return self._noise_matrix

    @property
    def type(self) -> str:
        return "noise"

    @property
    def qasm_name(self) -> str:
        return "noise"

    def convert_precision(self):
        self._precision = np.complex128 if self._precision == np.complex64 else np.complex64
        self._noise_matrix = [m.astype(self._precision) for m in self._noise_matrix]

    def superoperator(self) -> np.ndarray:
        """sum_i K_i (x) conj(K_i): the channel on vec(rho) (row index major)."""
        d = self._noise_matrix[0].shape[0]
        s = np.zeros((d * d, d * d), dtype=np.complex128)
        for k in self._noise_matrix:
            k = k.astype(np.complex128)
            s += np.kron(k, k.conj())
        return s


class Trigger(Operator):
    """Measure ``targets`` qubits, then splice the CompositeGate mapped to the outcome."""

    def __init__(self, targets: int,
                 state_gate_mapping: Union[Dict[int, object], list, tuple, Callable]):
        super().__init__(targets)
        from .composite_gate import CompositeGate
        from .gate import BasicGate
        ok = (CompositeGate, BasicGate, type(None))
        if isinstance(state_gate_mapping, (list, tuple)):
            for g in state_gate_mapping:
                if not isinstance(g, ok):
                    raise qerr.TypeError("Trigger mapping entries must be CompositeGate/BasicGate/None")
            self._mapping = dict(enumerate(state_gate_mapping))
        elif isinstance(state_gate_mapping, dict):
            for k, v in state_gate_mapping.items():
                if not isinstance(k, int) or not isinstance(v, ok):
                    raise qerr.TypeError("Trigger mapping must be dict[int, CompositeGate/BasicGate/None]")
            self._mapping = dict(state_gate_mapping)
        elif callable(state_gate_mapping):
            for i in range(1 << targets):
                if not isinstance(state_gate_mapping(i), ok):
                    raise qerr.TypeError("Trigger mapping function must return CompositeGate/BasicGate/None")
            self._mapping = state_gate_mapping
        else:
            raise qerr.TypeError("Trigger mapping must be list/dict/function")

    def mapping(self, state: int):
        if not 0 <= state < (1 << self.targets):
            raise qerr.ValueError("Trigger.mapping state out of range", state)
        if callable(self._mapping) and not isinstance(self._mapping, dict):
            return self._mapping(state)
        return self._mapping.get(state)
