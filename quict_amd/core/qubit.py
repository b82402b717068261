"""Qubit / Qureg: measurement outcome holders.

Follows QuICT/core/qubit/qubit.py:13-306 where it matters to the simulator:
``Qubit.measured`` keeps the latest outcome, and ``int(Qureg)`` reads the
outcomes big-endian (the first qubit is the most significant bit,
qubit.py:192-213), which is also the state-vector amplitude index.
"""
import uuid
from typing import List, Union

import numpy as np

from . import exception as qerr


class Qubit:
    def __init__(self, fidelity: float = 1.0):
        self._id = uuid.uuid4().hex
        self._measured = None
        self._history: List[int] = []
        self.fidelity = fidelity

    @property
    def id(self) -> str:
        return self._id

    @property
    def measured(self):
        return self._measured

    @measured.setter
    def measured(self, v):
        self._measured = None if v is None else int(v)
        if v is not None:
            self._history.append(int(v))

    @property
    def historical_measured(self) -> List[int]:
        return self._history

    def reset(self):
        self._measured = None
        self._history = []

    def __int__(self):
        if self._measured is None:
            raise qerr.QubitMeasureError(f"qubit {self._id} has not been measured")
        return self._measured

    def __bool__(self):
        return True


class Qureg(list):
    def __init__(self, qubits: Union[int, "Qubit", list, None] = None):
        super().__init__()
        if qubits is None:
            return
        if isinstance(qubits, (int, np.integer)):
            for _ in range(int(qubits)):
                self.append(Qubit())
        elif isinstance(qubits, Qubit):
            self.append(qubits)
        else:
            for q in qubits:
                if isinstance(q, Qubit):
                    self.append(q)
                elif isinstance(q, Qureg):
                    self.extend(q)
                else:
                    raise qerr.TypeError("Qureg accepts int / Qubit / Qureg, got", type(q))

    def __call__(self, indexes):
        if isinstance(indexes, (int, np.integer)):
            return Qureg(self[int(indexes)])
        return Qureg([self[int(i)] for i in indexes])

    def __getitem__(self, item):
        if isinstance(item, (int, np.integer)):
            return super().__getitem__(int(item))
        if isinstance(item, slice):
            return Qureg(super().__getitem__(item))
        if isinstance(item, (list, tuple)):
            return Qureg([super(Qureg, self).__getitem__(int(i)) for i in item])
        raise qerr.TypeError("Qureg index must be int / slice / list, got", type(item))

    def __int__(self):
        v = 0
        for q in self:
            if q.measured is None:
                raise qerr.QubitMeasureError(f"qubit {q.id} has not been measured")
            v = (v << 1) | (1 if q.measured == 1 else 0)
        return v

    def __str__(self):
        return format(int(self), "b").zfill(len(self))

    def __eq__(self, other):
        if not isinstance(other, Qureg) or len(other) != len(self):
            return False
        ids = {q.id for q in self}
        return all(q.id in ids for q in other)

    __hash__ = object.__hash__

    def index(self, qubit, *args):
        if isinstance(qubit, Qubit):
            for i, q in enumerate(self):
                if q is qubit:
                    return i
            raise qerr.ValueError("qubit not in this Qureg")
        for i, q in enumerate(self):
            if q.id == qubit:
                return i
        raise qerr.ValueError("qubit id not in this Qureg")

    def reset_qubits(self):
        for q in self:
            q.reset()
