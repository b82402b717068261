"""CompositeGate: an ordered gate list with local qubit indices.

Supports the ``with CompositeGate() as cg: H & 0; CX & [0, 1]`` context
syntax, ``cg & [qubits]`` re-targeting, ``inverse()``, ``matrix()`` and being
spliced into a Circuit (QuICT/core/gate/composite_gate.py:19-273).
"""
from typing import List

import numpy as np

from .gate import BasicGate, CGATE_LIST
from . import exception as qerr


class CompositeGate:
    def __init__(self, name: str = None, gates: List[BasicGate] = None):
        self.name = name or "composite"
        self._gates: List[BasicGate] = []
        self.checkpoint = None
        for g in gates or []:
            self.append(g)

    def __enter__(self):
        CGATE_LIST.append(self)
        return self

    def __exit__(self, *exc):
        CGATE_LIST.remove(self)
        return False

    @property
    def gates(self) -> List[BasicGate]:
        return self._gates

    def size(self) -> int:
        return len(self._gates)

    def clean(self):
        self._gates = []

    def width(self) -> int:
        qs = set()
        for g in self._gates:
            qs.update(g.cargs + g.targs)
        return (max(qs) + 1) if qs else 0

    def append(self, gate: BasicGate):
        if isinstance(gate, CompositeGate):
            for g in gate.gates:
                self.append(g)
            return
        self._gates.append(gate.copy())

    def extend(self, gates):
        for g in gates:
            self.append(g)

    def __or__(self, target):
        """``cg | circuit`` or ``cg | circuit([q...])``."""
        target.extend(self)

    def __and__(self, targets):
        """Re-map local qubit i -> targets[i] for every gate."""
        if isinstance(targets, int):
            targets = [targets]
        for g in self._gates:
            args = g.cargs + g.targs
            new = [targets[a] for a in args]
            g._cargs = new[:g.controls]
            g._targs = new[g.controls:]
        if CGATE_LIST and CGATE_LIST[-1] is not self:
            CGATE_LIST[-1].append(self)
        return self

    def inverse(self) -> "CompositeGate":
        return CompositeGate(gates=[g.inverse() for g in reversed(self._gates)])

    def convert_precision(self):
        for g in self._gates:
            g.convert_precision()

    def matrix(self, *_args, **_kw) -> np.ndarray:
        from .utils import matrix_product_to_circuit
        n = self.width()
        u = np.identity(1 << n, dtype=np.complex128)
        for g in self._gates:
            u = matrix_product_to_circuit(g.matrix.astype(np.complex128), g.cargs + g.targs, n) @ u
        return u

    def qasm(self) -> str:
        return "".join(g.qasm() for g in self._gates)

    def __iter__(self):
        return iter(self._gates)

    def __len__(self):
        return len(self._gates)
