"""Quantum gates: the input contract of the simulation path.

API mirrors QuICT's ``BasicGate`` family (QuICT/core/gate/gate.py:21-2700):
singletons such as ``H``, ``CX``, parameterised factories such as ``Rx(theta)``,
``gate | circuit(q)`` / ``gate & [q0, q1]`` placement, ``cargs``/``targs``/
``pargs``, ``matrix``, ``matrix_type``, ``inverse()``, ``convert_precision()``.

Conventions that the simulator relies on (and that the reference fixes):
  * a gate's argument list is ``cargs + targs``; argument 0 is the MSB of the
    gate-matrix row/column index (QuICT/core/utils/utils.py:36-44);
  * matrices are built with numpy from the closed forms cited per gate below,
    so values agree with the reference to the last bit.

Implementation is table-driven: one ``_Spec`` per ``GateType`` holds the
arity, matrix type, parameter count, default parameters, matrix builder and
inverse rule; the named classes (``HGate`` ...) are thin subclasses that only
pick their spec.
"""
from __future__ import annotations

import copy as _copy
from typing import Callable, List, Optional, Union

import numpy as np

from .gate_type import (
    GateType, MatrixType, SPECIAL_GATE_SET, DIAGONAL_GATE_SET, PAULI_GATE_SET, CLIFFORD_GATE_SET,
)
from . import exception as qerr

# CompositeGate "with" context stack (gates created with ``&`` inside a
# ``with CompositeGate():`` block are appended to the innermost one).
CGATE_LIST: list = []

_S2 = 1 / np.sqrt(2)


class _Spec:
    __slots__ = ("controls", "targets", "params", "mtype", "defaults", "build", "inverse", "qasm_name")

    def __init__(self, controls, targets, params, mtype, build, defaults=(), inverse=None, qasm_name=None):
        self.controls = controls
        self.targets = targets
        self.params = params
        self.mtype = mtype
        self.build = build              # callable(pargs, dtype) -> ndarray, or None
        self.defaults = list(defaults)
        self.inverse = inverse          # callable(gate) -> gate, or None (self-inverse)
        self.qasm_name = qasm_name


def _const(rows):
    arr = np.array(rows, dtype=np.complex128)
    return lambda p, dt: arr.astype(dt)


# --- matrix builders: closed forms of QuICT/core/gate/gate.py ----------------
def _u1(p, dt):        # gate.py:788-793
    return np.array([[1, 0], [0, np.exp(1j * p[0])]], dtype=dt)


def _u2(p, dt):        # gate.py:839-846
    return np.array([[1 * _S2, -np.exp(1j * p[1]) * _S2],
                     [np.exp(1j * p[0]) * _S2, np.exp(1j * (p[0] + p[1])) * _S2]], dtype=dt)


def _u3(p, dt):        # gate.py:893-900
    return np.array([[np.cos(p[0] / 2), -np.exp(1j * p[2]) * np.sin(p[0] / 2)],
                     [np.exp(1j * p[1]) * np.sin(p[0] / 2), np.exp(1j * (p[1] + p[2])) * np.cos(p[0] / 2)]],
                    dtype=dt)


def _rx(p, dt):        # gate.py:986-991
    return np.array([[np.cos(p[0] / 2), 1j * -np.sin(p[0] / 2)],
                     [1j * -np.sin(p[0] / 2), np.cos(p[0] / 2)]], dtype=dt)


def _ry(p, dt):        # gate.py:1031-1036
    return np.array([[np.cos(p[0] / 2), -np.sin(p[0] / 2)],
                     [np.sin(p[0] / 2), np.cos(p[0] / 2)]], dtype=dt)


def _rz(p, dt):        # gate.py:1077-1082
    return np.array([[np.exp(-p[0] / 2 * 1j), 0], [0, np.exp(p[0] / 2 * 1j)]], dtype=dt)


def _phase(p, dt):     # gate.py:1185-1190
    return np.array([[1, 0], [0, np.exp(p[0] * 1j)]], dtype=dt)


def _gphase(p, dt):    # gate.py:1225-1230
    return np.array([[np.exp(p[0] * 1j), 0], [0, np.exp(p[0] * 1j)]], dtype=dt)


def _crz(p, dt):       # gate.py:1400-1407
    return np.array([[1, 0, 0, 0], [0, 1, 0, 0],
                     [0, 0, np.exp(-p[0] / 2 * 1j), 0], [0, 0, 0, np.exp(p[0] / 2 * 1j)]], dtype=dt)


def _cu1(p, dt):       # gate.py:1452-1459
    return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, np.exp(1j * p[0])]], dtype=dt)


def _cu3(p, dt):       # gate.py:1524-1532
    m = np.identity(4, dtype=np.complex128)
    m[2:, 2:] = _u3(p, np.complex128)
    return m.astype(dt)


def _fsim(p, dt):      # gate.py:1594-1604
    c, s, phi = np.cos(p[0]), np.sin(p[0]), p[1]
    return np.array([[1, 0, 0, 0], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [0, 0, 0, np.exp(-1j * phi)]], dtype=dt)


def _rxx(p, dt):       # gate.py:1642-1652
    c, s = np.cos(p[0] / 2), np.sin(p[0] / 2)
    return np.array([[c, 0, 0, -1j * s], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [-1j * s, 0, 0, c]], dtype=dt)


def _ryy(p, dt):       # gate.py:1719-1729
    c, s = np.cos(p[0] / 2), np.sin(p[0] / 2)
    return np.array([[c, 0, 0, 1j * s], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [1j * s, 0, 0, c]], dtype=dt)


def _rzz(p, dt):       # gate.py:1796-1806
    e, se = np.exp(0.5j * p[0]), np.exp(-0.5j * p[0])
    return np.array([[se, 0, 0, 0], [0, e, 0, 0], [0, 0, e, 0], [0, 0, 0, se]], dtype=dt)


def _rzx(p, dt):       # gate.py:1873-1883
    c, s = np.cos(p[0] / 2), np.sin(p[0] / 2)
    return np.array([[c, -1j * s, 0, 0], [-1j * s, c, 0, 0], [0, 0, c, 1j * s], [0, 0, 1j * s, c]], dtype=dt)


def _ccrz(p, dt):      # gate.py:2499-2510
    m = np.identity(8, dtype=np.complex128)
    m[6, 6] = np.exp(-p[0] / 2 * 1j)
    m[7, 7] = np.exp(p[0] / 2 * 1j)
    return m.astype(dt)


def _id_n(n):
    return np.identity(1 << n, dtype=np.complex128)


def _ccx_mat():
    m = _id_n(3)
    m[6:, 6:] = [[0, 1], [1, 0]]
    return m


def _ccz_mat():
    m = _id_n(3)
    m[7, 7] = -1
    return m


def _cswap_mat():
    m = _id_n(3)
    m[5, 5] = m[6, 6] = 0
    m[5, 6] = m[6, 5] = 1
    return m


# --- inverse rules (QuICT/core/gate/gate.py per-class ``inverse``) ------------
def _inv_neg_all(g):
    out = g.copy()
    out.pargs = [-x for x in g.pargs]
    return out


def _inv_type(other: GateType):
    def f(g):
        out = _build(other)
        out.targs = list(g.targs)
        out.cargs = list(g.cargs)
        if g.precision == np.complex64:
            out.convert_precision()
        return out
    return f


def _inv_u2(g):         # gate.py:850-854
    out = g.copy()
    out.pargs = [np.pi - g.pargs[1], np.pi - g.pargs[0]]
    return out


def _inv_u3(g):         # gate.py:953-957
    out = g.copy()
    out.pargs = [g.pargs[0], np.pi - g.pargs[2], np.pi - g.pargs[1]]
    return out


def _inv_dagger(g):
    """Exact inverse via the conjugate transpose, as a Unitary gate.

    Used where the reference's rule is only correct up to a global phase (SX,
    SY, SW) or missing (iSwap family): simulations that compare amplitudes need
    the exact inverse.
    """
    out = Unitary(np.ascontiguousarray(g.matrix.conj().T))
    out.targs = list(g.cargs) + list(g.targs)
    return out


_SPECS = {
    GateType.h: _Spec(0, 1, 0, MatrixType.normal, _const([[_S2, _S2], [_S2, -_S2]])),
    GateType.hy: _Spec(0, 1, 0, MatrixType.normal, _const([[_S2, -1j * _S2], [1j * _S2, -_S2]])),
    GateType.s: _Spec(0, 1, 0, MatrixType.control, _const([[1, 0], [0, 1j]]), inverse=_inv_type(GateType.sdg)),
    GateType.sdg: _Spec(0, 1, 0, MatrixType.control, _const([[1, 0], [0, -1j]]), inverse=_inv_type(GateType.s)),
    GateType.x: _Spec(0, 1, 0, MatrixType.swap, _const([[0, 1], [1, 0]])),
    GateType.y: _Spec(0, 1, 0, MatrixType.reverse, _const([[0, -1j], [1j, 0]])),
    GateType.z: _Spec(0, 1, 0, MatrixType.control, _const([[1, 0], [0, -1]])),
    GateType.sx: _Spec(0, 1, 0, MatrixType.normal, _const([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]]),
                       inverse=_inv_dagger),
    GateType.sy: _Spec(0, 1, 0, MatrixType.normal, _const([[_S2, -_S2], [_S2, _S2]]), inverse=_inv_dagger),
    GateType.sw: _Spec(0, 1, 0, MatrixType.normal,
                       _const([[_S2, -np.sqrt(1j / 2)], [np.sqrt(-1j / 2), _S2]]), inverse=_inv_dagger),
    GateType.id: _Spec(0, 1, 0, MatrixType.diagonal, _const([[1, 0], [0, 1]])),
    GateType.u1: _Spec(0, 1, 1, MatrixType.control, _u1, [np.pi / 2], _inv_neg_all),
    GateType.u2: _Spec(0, 1, 2, MatrixType.normal, _u2, [np.pi / 2, np.pi / 2], _inv_u2),
    GateType.u3: _Spec(0, 1, 3, MatrixType.normal, _u3, [0, 0, np.pi / 2], _inv_u3),
    GateType.rx: _Spec(0, 1, 1, MatrixType.normal, _rx, [np.pi / 2], _inv_neg_all),
    GateType.ry: _Spec(0, 1, 1, MatrixType.normal, _ry, [np.pi / 2], _inv_neg_all),
    GateType.rz: _Spec(0, 1, 1, MatrixType.diagonal, _rz, [np.pi / 2], _inv_neg_all),
    GateType.t: _Spec(0, 1, 0, MatrixType.control, _const([[1, 0], [0, _S2 + 1j * _S2]]),
                      inverse=_inv_type(GateType.tdg)),
    GateType.tdg: _Spec(0, 1, 0, MatrixType.control, _const([[1, 0], [0, _S2 + 1j * -_S2]]),
                        inverse=_inv_type(GateType.t)),
    GateType.phase: _Spec(0, 1, 1, MatrixType.control, _phase, [0], _inv_neg_all, qasm_name="p"),
    GateType.gphase: _Spec(0, 1, 1, MatrixType.diagonal, _gphase, [0], _inv_neg_all, qasm_name="phase"),
    GateType.cz: _Spec(1, 1, 0, MatrixType.control, _const(np.diag([1, 1, 1, -1]))),
    GateType.cx: _Spec(1, 1, 0, MatrixType.reverse,
                       _const([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])),
    GateType.cy: _Spec(1, 1, 0, MatrixType.reverse,
                       _const([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]])),
    GateType.ch: _Spec(1, 1, 0, MatrixType.normal,
                       _const([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, _S2, _S2], [0, 0, _S2, -_S2]])),
    GateType.crz: _Spec(1, 1, 1, MatrixType.diagonal, _crz, [np.pi / 2], _inv_neg_all),
    GateType.cu1: _Spec(1, 1, 1, MatrixType.control, _cu1, [np.pi / 2], _inv_neg_all),
    GateType.cu3: _Spec(1, 1, 3, MatrixType.normal, _cu3, [np.pi / 2, 0, 0], _inv_u3),
    GateType.fsim: _Spec(0, 2, 2, MatrixType.ctrl_normal, _fsim, [np.pi / 2, 0], _inv_neg_all),
    GateType.rxx: _Spec(0, 2, 1, MatrixType.normal_normal, _rxx, [0], _inv_neg_all),
    GateType.ryy: _Spec(0, 2, 1, MatrixType.normal_normal, _ryy, [np.pi / 2], _inv_neg_all),
    GateType.rzz: _Spec(0, 2, 1, MatrixType.diag_diag, _rzz, [np.pi / 2], _inv_neg_all),
    GateType.rzx: _Spec(0, 2, 1, MatrixType.diag_normal, _rzx, [np.pi / 2], _inv_neg_all),
    GateType.swap: _Spec(0, 2, 0, MatrixType.swap,
                         _const([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])),
    GateType.iswap: _Spec(0, 2, 0, MatrixType.swap,
                          _const([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]),
                          inverse=_inv_type(GateType.iswapdg)),
    GateType.iswapdg: _Spec(0, 2, 0, MatrixType.swap,
                            _const([[1, 0, 0, 0], [0, 0, -1j, 0], [0, -1j, 0, 0], [0, 0, 0, 1]]),
                            inverse=_inv_type(GateType.iswap)),
    GateType.sqiswap: _Spec(0, 2, 0, MatrixType.swap,
                            _const([[1, 0, 0, 0], [0, 0, (1 + 1j) * _S2, 0], [0, (1 + 1j) * _S2, 0, 0],
                                    [0, 0, 0, 1]]),
                            inverse=_inv_dagger),
    GateType.cswap: _Spec(1, 2, 0, MatrixType.swap, _const(_cswap_mat())),
    GateType.ccx: _Spec(2, 1, 0, MatrixType.reverse, _const(_ccx_mat())),
    GateType.ccz: _Spec(2, 1, 0, MatrixType.normal, _const(_ccz_mat())),
    GateType.ccrz: _Spec(2, 1, 1, MatrixType.diagonal, _ccrz, [0], _inv_neg_all),
    GateType.measure: _Spec(0, 1, 0, MatrixType.special, None),
    GateType.reset: _Spec(0, 1, 0, MatrixType.special, None),
    GateType.barrier: _Spec(0, 1, 0, MatrixType.special, None),
    GateType.unitary: _Spec(0, 0, 0, MatrixType.normal, None),
    GateType.perm: _Spec(0, 0, 0, MatrixType.special, None),
    GateType.perm_fx: _Spec(0, 0, 0, MatrixType.normal, None),
    GateType.qft: _Spec(0, 3, 0, MatrixType.normal, None),
    GateType.iqft: _Spec(0, 3, 0, MatrixType.normal, None),
}


def _is_number(x):
    return isinstance(x, (int, float, complex, np.integer, np.floating, np.complexfloating))


class BasicGate:
    """One quantum gate (see module docstring for the conventions)."""

    TYPE: GateType = None

    def __init__(self, params: list = None, _type: GateType = None):
        gtype = _type or self.TYPE
        spec = _SPECS[gtype]
        self._type = gtype
        self._spec = spec
        self._matrix_type = spec.mtype
        self._controls = spec.controls
        self._targets = spec.targets
        self._params = spec.params
        self._cargs: List[int] = []
        self._targs: List[int] = []
        self._pargs: list = []
        self._precision = np.complex128
        self._matrix = None
        self._qasm_name = spec.qasm_name or gtype.name
        self._name = "-".join([str(gtype), "", ""])
        self.assigned_qubits = []
        if spec.params:
            self.pargs = list(params) if params is not None else list(spec.defaults)

    # --- attributes --------------------------------------------------------
    @property
    def type(self) -> GateType:
        return self._type

    @property
    def matrix_type(self) -> MatrixType:
        return self._matrix_type

    @property
    def name(self) -> str:
        return self._name

    @name.setter
    def name(self, v):
        self._name = v

    @property
    def qasm_name(self) -> str:
        return self._qasm_name

    @property
    def precision(self):
        return self._precision

    @property
    def controls(self) -> int:
        return self._controls

    @controls.setter
    def controls(self, v: int):
        self._controls = int(v)

    @property
    def targets(self) -> int:
        return self._targets

    @targets.setter
    def targets(self, v: int):
        self._targets = int(v)

    @property
    def params(self) -> int:
        return self._params

    @params.setter
    def params(self, v: int):
        self._params = int(v)

    @property
    def cargs(self) -> List[int]:
        return self._cargs

    @cargs.setter
    def cargs(self, cargs):
        if isinstance(cargs, (int, np.integer)):
            cargs = [int(cargs)]
        cargs = list(cargs)
        if len(cargs) != len(set(cargs)):
            raise qerr.ValueError("BasicGate.cargs must not hold duplicates", cargs)
        self._cargs = cargs

    @property
    def targs(self) -> List[int]:
        return self._targs

    @targs.setter
    def targs(self, targs):
        if isinstance(targs, (int, np.integer)):
            targs = [int(targs)]
        targs = list(targs)
        if len(targs) != len(set(targs)):
            raise qerr.ValueError("BasicGate.targs must not hold duplicates", targs)
        if set(self._cargs) & set(targs):
            raise qerr.ValueError("BasicGate.targs overlap the control qubits", targs)
        self._targs = targs

    @property
    def pargs(self) -> list:
        return self._pargs

    @pargs.setter
    def pargs(self, pargs):
        pargs = list(pargs) if isinstance(pargs, (list, tuple, np.ndarray)) else [pargs]
        if len(pargs) != self._params:
            raise qerr.ValueError(f"BasicGate.pargs needs {self._params} values, got", len(pargs))
        self._pargs = pargs

    @property
    def carg(self) -> int:
        return self._cargs[0]

    @property
    def targ(self) -> int:
        return self._targs[0]

    @property
    def parg(self):
        return self._pargs[0]

    @property
    def matrix(self) -> np.ndarray:
        if self._type in (GateType.measure, GateType.reset, GateType.barrier, GateType.perm):
            raise qerr.GateMatrixError(f"{self._type.name} has no matrix")
        if self._matrix is not None:
            return self._matrix
        if self._spec.build is None:
            raise qerr.GateMatrixError(f"{self._type.name} has no matrix yet")
        if self._params == 0:
            self._matrix = self._spec.build(None, self._precision)
            return self._matrix
        return self._spec.build(self._pargs, self._precision)

    @matrix.setter
    def matrix(self, m):
        self._matrix = m

    @property
    def target_matrix(self) -> np.ndarray:
        """The matrix restricted to the target qubits (controls removed)."""
        m = self.matrix
        t = 1 << self._targets
        return m[-t:, -t:]

    # --- construction syntax -------------------------------------------------
    def __call__(self, *params):
        """``Rx(theta)`` style parameter binding; returns a fresh gate."""
        if self._params == 0:
            return self.copy()
        if len(params) == 1 and isinstance(params[0], (list, tuple)) and self._params != 1:
            params = tuple(params[0])
        for p in params:
            if not _is_number(p):
                raise qerr.TypeError(f"{self._type.name} parameter must be int/float/complex, got", type(p))
        g = self.copy()
        g.pargs = list(params)
        return g

    def __or__(self, targets):
        """``gate | circuit`` / ``gate | circuit(q)`` / ``gate | composite``."""
        try:
            targets.append(self)
        except qerr.QuICTException:
            raise
        except Exception as e:  # pragma: no cover - mirrors reference wrapping
            raise qerr.GateAppendError(f"cannot append {self._type.name}: {e}")

    def __and__(self, targets):
        """``gate & q`` / ``gate & [c, t]``: bind qubit indices (controls first)."""
        g = self.copy()
        if isinstance(targets, (int, np.integer)):
            if g.controls + g.targets != 1:
                raise qerr.GateQubitAssignedError("qubit count must equal the gate's arity")
            g.targs = [int(targets)]
        elif isinstance(targets, (list, tuple)):
            targets = [int(t) for t in targets]
            if len(targets) != g.controls + g.targets:
                raise qerr.GateQubitAssignedError("qubit count must equal the gate's arity")
            g.cargs = targets[:g.controls]
            g.targs = targets[g.controls:]
        else:
            raise qerr.TypeError("BasicGate.& expects int or list[int], got", type(targets))
        if CGATE_LIST:
            CGATE_LIST[-1].append(g)
            return None
        return g

    def __eq__(self, other):
        if not isinstance(other, BasicGate):
            return False
        if self.type != other.type or (self.cargs + self.targs) != (other.cargs + other.targs):
            return False
        if self.type in (GateType.measure, GateType.reset, GateType.barrier):
            return True
        return bool(np.allclose(self.matrix, other.matrix))

    __hash__ = object.__hash__

    def __str__(self):
        return str({"name": self.name, "controls": self.controls, "control_bit": self.cargs,
                    "targets": self.targets, "target_bit": self.targs, "parameters": self.pargs})

    __repr__ = __str__

    # --- predicates ----------------------------------------------------------
    def is_single(self) -> bool:
        return self.controls + self.targets == 1

    def is_control_single(self) -> bool:
        return self.controls == 1 and self.targets == 1

    def is_clifford(self) -> bool:
        return self.type in CLIFFORD_GATE_SET

    def is_pauli(self) -> bool:
        return self.type in PAULI_GATE_SET

    def is_special(self) -> bool:
        return self.type in SPECIAL_GATE_SET

    def is_diagonal(self) -> bool:
        if self.type in DIAGONAL_GATE_SET:
            return True
        if self.type == GateType.unitary:
            m = self.matrix
            return bool(np.allclose(np.diag(np.diag(m)), m))
        return False

    def is_identity(self) -> bool:
        if self.type in (GateType.reset, GateType.measure, GateType.barrier, GateType.perm):
            return False
        n = self.controls + self.targets
        return bool(np.allclose(self.matrix, np.identity(1 << n, dtype=self.precision)))

    # --- transforms ----------------------------------------------------------
    def copy(self) -> "BasicGate":
        g = object.__new__(type(self))
        g.__dict__.update(self.__dict__)
        g._cargs = list(self._cargs)
        g._targs = list(self._targs)
        g._pargs = list(self._pargs)
        g.assigned_qubits = list(self.assigned_qubits) if self.assigned_qubits else []
        if self._matrix is not None and self._type in (GateType.unitary, GateType.perm_fx,
                                                       GateType.qft, GateType.iqft):
            g._matrix = self._matrix.copy()
        return g

    def inverse(self) -> "BasicGate":
        rule = self._spec.inverse
        if rule is None:
            return self.copy()
        return rule(self)

    def convert_precision(self):
        """Toggle complex128 <-> complex64 (QuICT/core/gate/gate.py:329-337)."""
        if self.type in (GateType.measure, GateType.reset, GateType.barrier):
            return
        self._precision = np.complex64 if self._precision == np.complex128 else np.complex128
        if self._matrix is not None:
            self._matrix = self._matrix.astype(self._precision)

    def qasm(self) -> str:
        if self.type in (GateType.perm, GateType.perm_fx, GateType.unitary, GateType.qft, GateType.iqft):
            raise qerr.QASMError(f"{self.type.name} has no OpenQASM form")
        s = self.qasm_name
        if self.params > 0:
            s += "(" + ", ".join(str(p) for p in self.pargs) + ")"
        s += " " + ", ".join(f"q[{a}]" for a in self.cargs + self.targs) + ";\n"
        return s

    def expand(self, qubits: Union[int, list]) -> np.ndarray:
        """Full-register matrix of this gate (brute force; for tests and small n)."""
        from .utils import matrix_product_to_circuit
        if isinstance(qubits, int):
            qubits = list(range(qubits))
        args = self.cargs + self.targs
        if len(qubits) == len(args):
            return self.matrix
        return matrix_product_to_circuit(self.matrix, [qubits.index(a) for a in args], len(qubits))


def _build(gtype: GateType) -> BasicGate:
    return GATE_TYPE_TO_CLASS[gtype]()


# --- special gates with non-table behaviour ---------------------------------
class UnitaryGate(BasicGate):
    """Arbitrary unitary on ``targets`` qubits (QuICT/core/gate/gate.py:2265-2345)."""
    TYPE = GateType.unitary

    def __call__(self, params, matrix_type: MatrixType = MatrixType.normal):
        m = np.array(params, dtype=np.complex128) if isinstance(params, list) else np.asarray(params)
        if m.size == 0:
            raise qerr.GateMatrixError("empty unitary")
        if m.ndim != 2 or m.shape[0] != m.shape[1]:
            side = int(round(np.sqrt(m.size)))
            m = m.reshape(side, side)
        n = int(np.log2(m.shape[0]))
        if (1 << n) != m.shape[0]:
            raise qerr.GateMatrixError("unitary side must be a power of two")
        u = UnitaryGate()
        u._targets = n
        u._matrix = m.astype(self._precision)
        u._precision = self._precision
        if n <= 3:
            u._matrix_type = u._classify()
        else:
            u._matrix_type = matrix_type
        return u

    def _classify(self) -> MatrixType:
        """QuICT/core/gate/gate.py:2313-2325: diagonal / control / reverse detection."""
        m = self._matrix
        d = 1 << self._targets
        mt = MatrixType.normal
        if np.allclose(np.diag(np.diag(m)), m):
            is_ctrl = np.allclose(m[:-1, :-1], np.identity(d - 1, dtype=m.dtype))
            mt = MatrixType.control if is_ctrl else MatrixType.diagonal
        if (np.allclose(m[:-2, :-2], np.identity(d - 2, dtype=m.dtype))
                and np.sum(m[:-2, -1]) + m[-1, -1] == 0 and np.sum(m[-1, :-2]) + m[-2, -2] == 0):
            mt = MatrixType.reverse
        return mt

    def inverse(self):
        g = self.copy()
        g._matrix = np.ascontiguousarray(self._matrix.conj().T)
        return g


class PermGate(BasicGate):
    """Qubit permutation (QuICT/core/gate/gate.py:2110-2175)."""
    TYPE = GateType.perm

    def __call__(self, targets: int, params: list):
        if not isinstance(params, list) or not isinstance(targets, int):
            raise qerr.TypeError("PermGate(targets:int, params:list)")
        if len(params) != targets or sorted(params) != list(range(targets)):
            raise qerr.GateParametersAssignedError("params must be a permutation of range(targets)")
        g = PermGate()
        g._targets = targets
        g._params = targets
        g._pargs = list(params)
        return g

    def inverse(self):
        g = self.copy()
        inv = [0] * self.targets
        for i, p in enumerate(self.pargs):
            inv[p] = i
        g._pargs = inv
        return g


class PermFxGate(BasicGate):
    """Oracle-style permutation |x,y> -> |x, y ^ f(x)> (QuICT/core/gate/gate.py:2178-2230)."""
    TYPE = GateType.perm_fx

    def __call__(self, n: int, params: list):
        if not isinstance(params, list) or not isinstance(n, int):
            raise qerr.TypeError("PermFx(n:int, params:list)")
        N = 1 << n
        if any(p >= N for p in params):
            raise qerr.ValueError("PermFx params must be < 2^n")
        g = PermFxGate()
        g._targets = n + 1
        g._params = 1 << (n + 1)
        pset = set(params)
        g._pargs = [(i ^ 1) if (i >> 1) in pset else i for i in range(1 << (n + 1))]
        m = np.zeros((1 << (n + 1), 1 << (n + 1)), dtype=self._precision)
        for i, p in enumerate(g._pargs):
            m[i, p] = 1
        g._matrix = m
        return g


class QFTGate(BasicGate):
    """Quantum Fourier transform on ``targets`` qubits (QuICT/core/gate/gate.py:2523-2575)."""
    TYPE = GateType.qft

    def __init__(self, targets: int = 3):
        super().__init__()
        self._targets = targets

    def __call__(self, targets: int):
        return type(self)(targets)

    def _decomposition(self):
        n = self._targets
        seq = []
        for i in range(n):
            seq.append((GateType.h, [i], []))
            for j in range(i + 1, n):
                seq.append((GateType.cu1, [j, i], [2 * np.pi / (1 << (j - i + 1))]))
        return seq

    @property
    def matrix(self):
        if self._matrix is None:
            from .utils import matrix_product_to_circuit
            n = self._targets
            u = np.identity(1 << n, dtype=np.complex128)
            for gt, args, p in self._decomposition():
                g = GATE_TYPE_TO_CLASS[gt](p) if p else GATE_TYPE_TO_CLASS[gt]()
                u = matrix_product_to_circuit(g.matrix, args, n) @ u
            self._matrix = u.astype(self._precision)
        return self._matrix

    @matrix.setter
    def matrix(self, m):
        self._matrix = m

    def inverse(self):
        g = (IQFTGate if self.TYPE == GateType.qft else QFTGate)(self._targets)
        g.targs = list(self.targs)
        return g


class IQFTGate(QFTGate):
    TYPE = GateType.iqft

    def _decomposition(self):
        n = self._targets
        seq = []
        for i in range(n - 1, -1, -1):
            for j in range(n - 1, i, -1):
                seq.append((GateType.cu1, [j, i], [-2 * np.pi / (1 << (j - i + 1))]))
            seq.append((GateType.h, [i], []))
        return seq


def _make_class(name: str, gtype: GateType):
    return type(name, (BasicGate,), {"TYPE": gtype, "__doc__": f"{gtype.value} gate"})


HGate = _make_class("HGate", GateType.h)
HYGate = _make_class("HYGate", GateType.hy)
SGate = _make_class("SGate", GateType.s)
SDaggerGate = _make_class("SDaggerGate", GateType.sdg)
XGate = _make_class("XGate", GateType.x)
YGate = _make_class("YGate", GateType.y)
ZGate = _make_class("ZGate", GateType.z)
SXGate = _make_class("SXGate", GateType.sx)
SYGate = _make_class("SYGate", GateType.sy)
SWGate = _make_class("SWGate", GateType.sw)
IDGate = _make_class("IDGate", GateType.id)
U1Gate = _make_class("U1Gate", GateType.u1)
U2Gate = _make_class("U2Gate", GateType.u2)
U3Gate = _make_class("U3Gate", GateType.u3)
RxGate = _make_class("RxGate", GateType.rx)
RyGate = _make_class("RyGate", GateType.ry)
RzGate = _make_class("RzGate", GateType.rz)
TGate = _make_class("TGate", GateType.t)
TDaggerGate = _make_class("TDaggerGate", GateType.tdg)
PhaseGate = _make_class("PhaseGate", GateType.phase)
GlobalPhaseGate = _make_class("GlobalPhaseGate", GateType.gphase)
CZGate = _make_class("CZGate", GateType.cz)
CXGate = _make_class("CXGate", GateType.cx)
CYGate = _make_class("CYGate", GateType.cy)
CHGate = _make_class("CHGate", GateType.ch)
CRzGate = _make_class("CRzGate", GateType.crz)
CU1Gate = _make_class("CU1Gate", GateType.cu1)
CU3Gate = _make_class("CU3Gate", GateType.cu3)
FSimGate = _make_class("FSimGate", GateType.fsim)
RxxGate = _make_class("RxxGate", GateType.rxx)
RyyGate = _make_class("RyyGate", GateType.ryy)
RzzGate = _make_class("RzzGate", GateType.rzz)
RzxGate = _make_class("RzxGate", GateType.rzx)
SwapGate = _make_class("SwapGate", GateType.swap)
iSwapGate = _make_class("iSwapGate", GateType.iswap)
iSwapDaggerGate = _make_class("iSwapDaggerGate", GateType.iswapdg)
SquareRootiSwapGate = _make_class("SquareRootiSwapGate", GateType.sqiswap)
CSwapGate = _make_class("CSwapGate", GateType.cswap)
CCXGate = _make_class("CCXGate", GateType.ccx)
CCZGate = _make_class("CCZGate", GateType.ccz)
CCRzGate = _make_class("CCRzGate", GateType.ccrz)
MeasureGate = _make_class("MeasureGate", GateType.measure)
ResetGate = _make_class("ResetGate", GateType.reset)
BarrierGate = _make_class("BarrierGate", GateType.barrier)

GATE_TYPE_TO_CLASS = {
    GateType.h: HGate, GateType.hy: HYGate, GateType.s: SGate, GateType.sdg: SDaggerGate,
    GateType.x: XGate, GateType.y: YGate, GateType.z: ZGate, GateType.sx: SXGate,
    GateType.sy: SYGate, GateType.sw: SWGate, GateType.id: IDGate, GateType.u1: U1Gate,
    GateType.u2: U2Gate, GateType.u3: U3Gate, GateType.rx: RxGate, GateType.ry: RyGate,
    GateType.rz: RzGate, GateType.t: TGate, GateType.tdg: TDaggerGate, GateType.phase: PhaseGate,
    GateType.gphase: GlobalPhaseGate, GateType.cz: CZGate, GateType.cx: CXGate, GateType.cy: CYGate,
    GateType.ch: CHGate, GateType.crz: CRzGate, GateType.cu1: CU1Gate, GateType.cu3: CU3Gate,
    GateType.fsim: FSimGate, GateType.rxx: RxxGate, GateType.ryy: RyyGate, GateType.rzz: RzzGate,
    GateType.rzx: RzxGate, GateType.swap: SwapGate, GateType.iswap: iSwapGate,
    GateType.iswapdg: iSwapDaggerGate, GateType.sqiswap: SquareRootiSwapGate,
    GateType.cswap: CSwapGate, GateType.ccx: CCXGate, GateType.ccz: CCZGate, GateType.ccrz: CCRzGate,
    GateType.measure: MeasureGate, GateType.reset: ResetGate, GateType.barrier: BarrierGate,
    GateType.unitary: UnitaryGate, GateType.perm: PermGate, GateType.perm_fx: PermFxGate,
    GateType.qft: QFTGate, GateType.iqft: IQFTGate,
}

# Singletons, named as in the reference.
H = HGate()
Hy = HYGate()
S = SGate()
S_dagger = SDaggerGate()
X = XGate()
Y = YGate()
Z = ZGate()
SX = SXGate()
SY = SYGate()
SW = SWGate()
ID = IDGate()
U1 = U1Gate()
U2 = U2Gate()
U3 = U3Gate()
Rx = RxGate()
Ry = RyGate()
Rz = RzGate()
T = TGate()
T_dagger = TDaggerGate()
Phase = PhaseGate()
GPhase = GlobalPhaseGate()
CZ = CZGate()
CX = CXGate()
CY = CYGate()
CH = CHGate()
CRz = CRzGate()
CU1 = CU1Gate()
CU3 = CU3Gate()
FSim = FSimGate()
Rxx = RxxGate()
Ryy = RyyGate()
Rzz = RzzGate()
Rzx = RzxGate()
Swap = SwapGate()
iSwap = iSwapGate()
iSwap_dagger = iSwapDaggerGate()
sqiSwap = SquareRootiSwapGate()
CSwap = CSwapGate()
CCX = CCXGate()
CCZ = CCZGate()
CCRz = CCRzGate()
Measure = MeasureGate()
Reset = ResetGate()
Barrier = BarrierGate()
Unitary = UnitaryGate()
Perm = PermGate()
PermFx = PermFxGate()
QFT = QFTGate()
IQFT = IQFTGate()


def build_gate(gate_type: GateType, qubits, params: Optional[list] = None) -> BasicGate:
    """QuICT/core/gate/gate_builder.py:72-104."""
    g = GATE_TYPE_TO_CLASS[gate_type]()
    if params is not None:
        params = params if isinstance(params, list) else [params]
        g = g(*params)
    if isinstance(qubits, (int, np.integer)):
        qubits = [int(qubits)]
    if len(qubits) != g.controls + g.targets:
        raise qerr.GateQubitAssignedError("qubit count must equal the gate's arity")
    g.cargs = list(qubits[:g.controls])
    g.targs = list(qubits[g.controls:])
    return g


def build_random_gate(gate_type: GateType, qubits: int, random_params: bool = False) -> BasicGate:
    """Random placement and parameters; RNG draw order of
    QuICT/core/gate/gate_builder.py:107-131: ``random.sample`` for the qubits
    (controls first), then ``np.random.uniform(0, 2*pi, params)``."""
    import random
    g = GATE_TYPE_TO_CLASS[gate_type]()
    chosen = random.sample(range(qubits), g.controls + g.targets)
    g.cargs = chosen[:g.controls]
    g.targs = chosen[g.controls:]
    if random_params and g.params:
        g.pargs = list(np.random.uniform(0, 2 * np.pi, g.params))
    return g
