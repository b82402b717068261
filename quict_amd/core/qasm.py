"""Minimal OpenQASM 2.0 reader/writer for the simulation path.

Reads the subset used by QuICT's circuit fixtures: ``qreg``/``creg``
declarations, standard gates named as in QuICT's name table
(QuICT/tools/interface/qasm_interface.py:42-85), parameter expressions built
from numbers, ``pi`` and + - * / ( ), ``measure``, ``reset`` and ``barrier``.
Gate definitions (``gate ... { }``), ``if`` and ``opaque`` are rejected.
"""
import ast
import operator
import re

import numpy as np

from .circuit import Circuit
from .gate import GATE_TYPE_TO_CLASS
from .gate_type import GateType
from . import exception as qerr

NAME_TO_TYPE = {
    "u1": GateType.u1, "u2": GateType.u2, "u3": GateType.u3, "U": GateType.u3, "x": GateType.x,
    "y": GateType.y, "z": GateType.z, "t": GateType.t, "tdg": GateType.tdg, "s": GateType.s,
    "sdg": GateType.sdg, "sx": GateType.sx, "sy": GateType.sy, "sw": GateType.sw, "p": GateType.phase,
    "phase": GateType.gphase, "swap": GateType.swap, "iswap": GateType.iswap,
    "iswapdg": GateType.iswapdg, "sqiswap": GateType.sqiswap, "rx": GateType.rx, "ry": GateType.ry,
    "rz": GateType.rz, "id": GateType.id, "h": GateType.h, "hy": GateType.hy, "cx": GateType.cx,
    "CX": GateType.cx, "ccx": GateType.ccx, "ccz": GateType.ccz, "cy": GateType.cy, "cz": GateType.cz,
    "ch": GateType.ch, "crz": GateType.crz, "ccrz": GateType.ccrz, "rzz": GateType.rzz,
    "rxx": GateType.rxx, "ryy": GateType.ryy, "rzx": GateType.rzx, "cu1": GateType.cu1,
    "cu3": GateType.cu3, "cswap": GateType.cswap, "fsim": GateType.fsim,
    "measure": GateType.measure, "reset": GateType.reset, "barrier": GateType.barrier,
}

_BINOPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
           ast.Pow: operator.pow}


def _eval_param(text: str) -> float:
    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return node.value
        if isinstance(node, ast.Name) and node.id == "pi":
            return np.pi
        if isinstance(node, ast.BinOp) and type(node.op) in _BINOPS:
            return _BINOPS[type(node.op)](ev(node.left), ev(node.right))
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = ev(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        raise qerr.QASMError(f"unsupported parameter expression: {text}")
    return ev(ast.parse(text.strip(), mode="eval"))


_STMT = re.compile(r"^(?P<name>[A-Za-z_][A-Za-z0-9_]*)\s*(\((?P<params>[^)]*)\))?\s*(?P<args>.*)$")
_ARG = re.compile(r"^(?P<reg>[A-Za-z_][A-Za-z0-9_]*)\s*(\[\s*(?P<idx>\d+)\s*\])?$")


def loads(text: str) -> Circuit:
    text = re.sub(r"//[^\n]*", "", text)
    statements = [s.strip() for s in text.split(";") if s.strip()]
    qregs = {}
    total = 0
    body = []
    for st in statements:
        if st.startswith("OPENQASM") or st.startswith("include"):
            continue
        if st.startswith("qreg"):
            m = re.match(r"qreg\s+([A-Za-z_][A-Za-z0-9_]*)\s*\[\s*(\d+)\s*\]", st)
            if not m:
                raise qerr.QASMError(f"bad qreg: {st}")
            qregs[m.group(1)] = (total, int(m.group(2)))
            total += int(m.group(2))
            continue
        if st.startswith("creg"):
            continue
        if st.startswith(("gate ", "opaque ", "if")):
            raise qerr.QASMError(f"unsupported statement: {st}")
        body.append(st)
    circuit = Circuit(total)
    for st in body:
        st = st.split("->")[0].strip()
        m = _STMT.match(st)
        if not m or m.group("name") not in NAME_TO_TYPE:
            raise qerr.QASMError(f"unknown statement: {st}")
        gtype = NAME_TO_TYPE[m.group("name")]
        params = [_eval_param(p) for p in m.group("params").split(",")] if m.group("params") else []
        qubits = []
        for a in [x.strip() for x in m.group("args").split(",") if x.strip()]:
            am = _ARG.match(a)
            if not am or am.group("reg") not in qregs:
                raise qerr.QASMError(f"bad argument: {a}")
            base, size = qregs[am.group("reg")]
            if am.group("idx") is None:
                qubits.extend(range(base, base + size))
            else:
                qubits.append(base + int(am.group("idx")))
        g = GATE_TYPE_TO_CLASS[gtype]()
        if params:
            g = g(*params)
        if gtype in (GateType.barrier,):
            for q in qubits:
                g2 = g.copy()
                g2.targs = [q]
                circuit.append(g2)
            continue
        if gtype in (GateType.measure, GateType.reset) and len(qubits) > 1:
            for q in qubits:
                g | circuit(q)
            continue
        g | circuit(qubits)
    return circuit


def load_file(path: str) -> Circuit:
    with open(path) as f:
        return loads(f.read())
