"""Gradients of expectation values of parameterised circuits on MI355X.

No reference counterpart (QuICT's QNN/VQE code is in the external QuICT_ml
package).  Builds on the reference's gate definitions and inverse rules
(QuICT/core/gate/gate.py: Rx/Ry/Rz/CRz/Rxx/Ryy/Rzz/Rzx are exp(-i theta P / 2),
U1/Phase are diag(1, e^{i theta})).

* ``adjoint_gradient``: forward simulation, then one native call
  (qs_adjoint_gradient_pauli) that forms lam = H psi and the energy and runs
  the backward sweep, which recomputes earlier states by applying gate
  inverses — no per-layer checkpoints — and accumulates
  dE/dtheta_k = 2 Re <lam| dG_k |psi_before_k>.
* ``parameter_shift_gradient``: [E(theta + pi/2 e_k) - E(theta - pi/2 e_k)] / 2,
  exact for exp(-i theta P / 2) gates; every evaluation is a GPU simulation.
* ``finite_difference_gradient``: central differences (h = 1e-6), a coarse check.
* ``AdjointExpectation``: torch.autograd.Function wrapping the adjoint path.
"""
import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ..core import GateType, GATE_TYPE_TO_CLASS, Circuit
from ..core import exception as qerr
from .. import _native as nat
from ..simulation.gate_compiler import compile_gate

_PAULI = {
    "I": np.eye(2, dtype=np.complex128),
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}

# generator P of G(theta) = exp(-i theta P / 2) on the gate's *target* qubits
_ROTATION_GENERATORS = {
    GateType.rx: _PAULI["X"], GateType.ry: _PAULI["Y"], GateType.rz: _PAULI["Z"], GateType.crz: _PAULI["Z"],
    GateType.rxx: np.kron(_PAULI["X"], _PAULI["X"]), GateType.ryy: np.kron(_PAULI["Y"], _PAULI["Y"]),
    GateType.rzz: np.kron(_PAULI["Z"], _PAULI["Z"]), GateType.rzx: np.kron(_PAULI["Z"], _PAULI["X"]),
}
_PHASE_TYPES = (GateType.u1, GateType.phase, GateType.cu1)
SHIFT_RULE_TYPES = tuple(_ROTATION_GENERATORS)


class PauliSum:
    """sum_k c_k P_k with P_k a Pauli string over qubits 0..n-1 (qubit 0 first)."""

    def __init__(self, n: int, terms: Sequence[Tuple[float, str]]):
        self.n = n
        self.terms = [(float(c), s.upper()) for c, s in terms]
        for _, s in self.terms:
            if len(s) != n or any(ch not in "IXYZ" for ch in s):
                raise qerr.ValueError("Pauli strings must have one I/X/Y/Z letter per qubit")
        self._recs = None  # (terms they were built from, qs_pauli_term records)

    def records(self) -> np.ndarray:
        cached = getattr(self, "_recs", None)
        if cached is not None and cached[0] == self.terms:
            return cached[1]
        arr = (nat.QsPauliTerm * len(self.terms))()
        for k, (c, s) in enumerate(self.terms):
            x = z = 0
            for q, ch in enumerate(s):
                bit = 1 << (self.n - 1 - q)
                if ch in "XY":
                    x |= bit
                if ch in "ZY":
                    z |= bit
            arr[k].coef, arr[k].xmask, arr[k].zmask = c, x, z
        self._recs = (list(self.terms), arr)
        return arr

    def matrix(self) -> np.ndarray:
        d = 1 << self.n
        m = np.zeros((d, d), dtype=np.complex128)
        for c, s in self.terms:
            t = np.array([[1.0 + 0j]])
            for ch in s:
                t = np.kron(t, _PAULI[ch])
            m += c * t
        return m


class ParameterizedCircuit:
    """Gate list whose angles come from a parameter vector.

    ``add(GateType.ry, [q], param=k)`` binds the gate's single angle to
    theta[k]; ``add(GateType.cx, [c, t])`` adds a fixed gate;
    ``value=`` fixes the angle(s) of a parameterised type.
    """

    def __init__(self, n: int):
        self.n = n
        self.ops: List[Tuple[GateType, List[int], Optional[int], Optional[list]]] = []
        self.nparams = 0
        self._template = None   # record templates, rebuilt after every add()
        self._work = None       # working copies of the templates' records

    def add(self, gtype: GateType, qubits, param: Optional[int] = None, value=None):
        qubits = [qubits] if isinstance(qubits, int) else list(qubits)
        if param is not None:
            if gtype not in _ROTATION_GENERATORS and gtype not in _PHASE_TYPES:
                raise qerr.ValueError(f"{gtype.name} cannot carry a trainable angle")
            self.nparams = max(self.nparams, param + 1)
        vals = None if value is None else (list(value) if isinstance(value, (list, tuple)) else [value])
        self.ops.append((gtype, qubits, param, vals))
        self._template = None
        return self

    @staticmethod
    def hea(n: int, layers: int) -> "ParameterizedCircuit":
        """Ry + Rz on every qubit then a CX ladder, per layer; params ordered (layer, {Ry, Rz}, qubit)."""
        pc = ParameterizedCircuit(n)
        for l in range(layers):
            for q in range(n):
                pc.add(GateType.ry, [q], param=(l * 2 + 0) * n + q)
                pc.add(GateType.rz, [q], param=(l * 2 + 1) * n + q)
            for q in range(n - 1):
                pc.add(GateType.cx, [q, q + 1])
        return pc

    def gate(self, j: int, theta: np.ndarray):
        """Gate object of op j at angles theta."""
        gtype, qubits, param, vals = self.ops[j]
        g = GATE_TYPE_TO_CLASS[gtype]()
        if param is not None:
            g = g(float(theta[param]))
        elif vals is not None:
            g = g(*vals)
        c = g.controls
        g.cargs = qubits[:c]
        g.targs = qubits[c:]
        return g

    def gates(self, theta: np.ndarray):
        return [(self.gate(j, theta), self.ops[j][2]) for j in range(len(self.ops))]

    def bind(self, theta: np.ndarray) -> Circuit:
        c = Circuit(self.n)
        for g, _ in self.gates(theta):
            c.append(g)
        return c


def _derivative_target_block(g) -> np.ndarray:
    """d G / d theta restricted to the target block (controls are kept as controls)."""
    gt = g.type
    t = g.targets
    Gt = np.asarray(g.matrix, dtype=np.complex128)[-(1 << t):, -(1 << t):]
    if gt in _ROTATION_GENERATORS:
        return -0.5j * _ROTATION_GENERATORS[gt] @ Gt
    if gt in _PHASE_TYPES:
        D = np.zeros_like(Gt)
        D[-1, -1] = 1j * Gt[-1, -1]
        return D
    raise qerr.ValueError(f"no derivative rule for {gt.name}")


def _rotation_payloads(gtype: GateType, th: np.ndarray):
    """(forward, inverse, derivative) record payloads of Rx / Ry / Rz for the
    angles th, one row per gate: the target blocks of exp(-i th P / 2), of its
    conjugate transpose and of d/dth = -(i/2) P exp(-i th P / 2)
    (QuICT/core/gate/gate.py:963-1096).  Rz records are diagonal."""
    return _rotation_payloads_cs(gtype, np.cos(th / 2), np.sin(th / 2))


def _rotation_payloads_cs(gtype: GateType, c: np.ndarray, s: np.ndarray):
    """_rotation_payloads from c = cos(th / 2), s = sin(th / 2): every entry is
    0, +-c, +-s, +-c/2 or +-s/2 (real or imaginary part)."""
    if gtype == GateType.ry:
        fwd = np.stack([c, -s, s, c], 1).astype(np.complex128)
        inv = np.stack([c, s, -s, c], 1).astype(np.complex128)
        der = np.stack([-s / 2, -c / 2, c / 2, -s / 2], 1).astype(np.complex128)
    elif gtype == GateType.rx:
        fwd = np.stack([c, -1j * s, -1j * s, c], 1)
        inv = np.stack([c, 1j * s, 1j * s, c], 1)
        der = np.stack([-s / 2 + 0j, -0.5j * c, -0.5j * c, -s / 2 + 0j], 1)
    else:
        e0, e1 = c - 1j * s, c + 1j * s
        z = np.zeros_like(e0)
        fwd = np.stack([e0, e1], 1)
        inv = np.stack([e1, e0], 1)
        der = np.stack([-0.5j * e0, z, z, 0.5j * e1], 1)
    return fwd, inv, der


_FAST_TYPES = (GateType.rx, GateType.ry, GateType.rz)
_PROBE_ANGLE = 0.7317


def _template(pc: ParameterizedCircuit):
    """Records of pc built once per circuit structure by the generic compiler
    (at a probe angle), plus, per rotation type whose records the closed forms
    above reproduce exactly, the rows and parameter indices to refill per call."""
    if pc._template is not None:
        return pc._template
    probe = np.full(max(pc.nparams, 1), _PROBE_ANGLE)
    fwd, inv, der, slot, keep = _records_generic(pc, probe)
    fast = {}
    for gt in _FAST_TYPES:
        rows = [j for j, (t, _, p, _) in enumerate(pc.ops) if t == gt and p is not None]
        if not rows:
            continue
        rows = np.array(rows)
        f, i, d = _rotation_payloads(gt, np.full(len(rows), _PROBE_ANGLE))
        ok = True
        for arr, pay in ((fwd, f), (inv, i), (der, d)):
            got = arr["m"][rows][:, : 2 * pay.shape[1]].view(np.complex128)
            ok = ok and np.array_equal(got, pay) and not np.any(arr["m"][rows][:, 2 * pay.shape[1]:])
        if ok:
            fast[gt] = (rows, np.array([pc.ops[j][2] for j in rows]))
    generic = [j for j, (t, _, p, _) in enumerate(pc.ops)
               if p is not None and not any(t == gt for gt in fast)]
    # every fast payload entry is a * cos(th/2) + b * sin(th/2): per record
    # array, the (a, b) of each fast row's first 8 matrix words, for one
    # vectorised refill (columns past a payload get 0, as they hold already)
    lin = None
    if fast:
        rows_all = np.concatenate([fast[gt][0] for gt in fast])
        params_all = np.concatenate([fast[gt][1] for gt in fast])
        lin = [rows_all, params_all]
        for k in range(3):
            A, B = [], []
            for gt in fast:
                nr = len(fast[gt][0])
                for c_, s_, dst in ((1.0, 0.0, A), (0.0, 1.0, B)):
                    pay = _rotation_payloads_cs(gt, np.full(nr, c_), np.full(nr, s_))[k]
                    w = np.zeros((nr, 8))
                    w[:, : 2 * pay.shape[1]] = np.ascontiguousarray(pay).view(np.float64)
                    dst.append(w)
            lin.append((np.concatenate(A), np.concatenate(B)))
        # the linear form reproduces the compiler's records exactly at the probe angle
        c0 = np.full(len(rows_all), np.cos(_PROBE_ANGLE / 2))
        s0 = np.full(len(rows_all), np.sin(_PROBE_ANGLE / 2))
        for arr, (A, B) in zip((fwd, inv, der), lin[2:]):
            if not np.array_equal(arr["m"][rows_all, :8], c0[:, None] * A + s0[:, None] * B):
                fast, lin = {}, None
                generic = [j for j, (t, _, p, _) in enumerate(pc.ops) if p is not None]
                break
    pc._template = (fwd, inv, der, slot, keep, fast, generic, lin)
    return pc._template


def _records(pc: ParameterizedCircuit, theta: np.ndarray, reuse: bool = False):
    """Forward, inverse and derivative records of pc at theta: the structure's
    template with the rotation payloads refilled in closed form (vectorised);
    other parameterised gates go through the generic compiler.  reuse: refill
    the circuit's own working copies instead of new arrays (valid until the
    next such call; for callers that hand them to one native call)."""
    fwd0, inv0, der0, slot, keep0, fast, generic, lin = _template(pc)
    if reuse:
        if getattr(pc, "_work", None) is None or pc._work[0] is not fwd0:
            pc._work = (fwd0, fwd0.copy(), inv0.copy(), der0.copy())
        fwd, inv, der = pc._work[1:]
    else:
        fwd, inv, der = fwd0.copy(), inv0.copy(), der0.copy()
    keep = list(keep0)
    if lin is not None:
        half = np.asarray(theta, dtype=np.float64)[lin[1]] / 2
        c, s = np.cos(half)[:, None], np.sin(half)[:, None]
        for arr, (A, B) in zip((fwd, inv, der), lin[2:]):
            arr["m"][lin[0], :8] = c * A + s * B
    if generic:
        for j in generic:
            g = pc.gate(j, theta)
            compile_gate(g, pc.n, fwd[j], keep)
            compile_gate(g, pc.n, inv[j], keep, dagger=True)
            D = _derivative_target_block(g)
            der[j]["m"][:] = 0
            v = np.ascontiguousarray(D.ravel())
            der[j]["m"][: 2 * len(v)] = v.view(np.float64)
    return fwd, inv, der, slot, keep


def _records_generic(pc: ParameterizedCircuit, theta: np.ndarray):
    n = pc.n
    pairs = pc.gates(theta)
    fwd = np.zeros(len(pairs), dtype=nat.QS_GATE_DTYPE)
    inv = np.zeros(len(pairs), dtype=nat.QS_GATE_DTYPE)
    der = np.zeros(len(pairs), dtype=nat.QS_GATE_DTYPE)
    slot = np.full(len(pairs), -1, dtype=np.int32)
    keep: list = []
    for j, (g, param) in enumerate(pairs):
        compile_gate(g, n, fwd[j], keep)
        compile_gate(g, n, inv[j], keep, dagger=True)
        if param is not None:
            slot[j] = param
            D = _derivative_target_block(g)
            rec = der[j]
            rec["kind"] = nat.GATE_DENSE
            rec["ncontrols"] = g.controls
            rec["ntargets"] = g.targets
            args = list(g.cargs) + list(g.targs)
            rec["bits"][: len(args)] = [n - 1 - a for a in args]
            v = D.ravel()
            rec["m"][: 2 * len(v)] = np.column_stack([v.real, v.imag]).ravel()
            # the inverse record must carry the same control / target split
            if inv[j]["ncontrols"] != g.controls or inv[j]["ntargets"] != g.targets:
                raise qerr.GateAlgorithmNotImplementError(f"{g.type.name}: unexpected record shape")
    return fwd, inv, der, slot, keep


class _Engine:
    def __init__(self, n: int, precision: str, device: int):
        import torch
        self.torch = torch
        self.n = n
        self.device = device
        self.tdtype = torch.complex128 if precision == "double" else torch.complex64
        self.dcode = nat.QS_C128 if precision == "double" else nat.QS_C64
        self.lib = nat.load_library()
        self.ctx = nat.context(device)
        self.psi = torch.empty(1 << n, dtype=self.tdtype, device=f"cuda:{device}")
        self.lam = torch.empty_like(self.psi)

    def ptr(self, t):
        return ctypes.c_void_p(t.data_ptr())

    def stream(self):
        return nat.stream_ptr(self.device)

    def forward(self, fwd: np.ndarray, fuse: int = 1):
        nat.check(self.lib.qs_sv_init(self.ctx, self.ptr(self.psi), self.n, self.dcode, 0, self.stream()), "init")
        nat.check(self.lib.qs_sv_apply_gates(self.ctx, self.ptr(self.psi), self.n, self.dcode,
                                             nat.gate_array_ptr(fwd), len(fwd), fuse, None, self.stream()), "forward")

    def expectation(self, H: PauliSum) -> float:
        recs = H.records()
        nat.check(self.lib.qs_sv_apply_pauli_sum(self.ctx, self.ptr(self.psi), self.ptr(self.lam), self.n,
                                                 self.dcode, ctypes.cast(recs, ctypes.c_void_p), len(recs),
                                                 self.stream()), "pauli sum")
        re, im = ctypes.c_double(0), ctypes.c_double(0)
        nat.check(self.lib.qs_sv_inner(self.ctx, self.ptr(self.psi), self.ptr(self.lam), self.n, self.dcode,
                                       ctypes.byref(re), ctypes.byref(im), self.stream()), "inner")
        return re.value


_engines = {}


def _engine(n, precision, device) -> _Engine:
    key = (n, precision, device)
    if key not in _engines:
        _engines[key] = _Engine(n, precision, device)
    return _engines[key]


def expectation(pc: ParameterizedCircuit, theta: np.ndarray, H: PauliSum, precision: str = "double",
                device: int = 0) -> float:
    eng = _engine(pc.n, precision, device)
    fwd, _, _, _, keep = _records(pc, np.asarray(theta, dtype=np.float64))
    eng.forward(fwd)
    return eng.expectation(H)


def adjoint_gradient(pc: ParameterizedCircuit, theta: np.ndarray, H: PauliSum, precision: str = "double",
                     device: int = 0) -> Tuple[float, np.ndarray]:
    theta = np.asarray(theta, dtype=np.float64)
    eng = _engine(pc.n, precision, device)
    fwd, inv, der, slot, keep = _records(pc, theta, reuse=True)
    grads = np.zeros(max(pc.nparams, 1), dtype=np.float64)
    if not H.terms:
        return 0.0, grads[: pc.nparams]
    # the fused forward passes are launched while they are planned; the sweep
    # below is planned while they run (measured on MI355X at 20 qubits, 480
    # parameters: 2.52 ms against 2.56 ms with the forward computed inside the
    # adjoint call from the sweep's own tile passes, and the fused passes are
    # the faster forward on wider registers)
    eng.forward(fwd)
    recs = H.records()
    E = ctypes.c_double(0)
    nat.check(eng.lib.qs_adjoint_gradient_pauli(eng.ctx, eng.ptr(eng.psi), eng.ptr(eng.lam), pc.n, eng.dcode,
                                                ctypes.cast(recs, ctypes.c_void_p), len(recs), None,
                                                nat.gate_array_ptr(inv), len(inv), nat.gate_array_ptr(der),
                                                slot.ctypes.data_as(ctypes.c_void_p),
                                                grads.ctypes.data_as(ctypes.c_void_p), len(grads), ctypes.byref(E),
                                                eng.stream()),
              "qs_adjoint_gradient_pauli")
    return E.value, grads[: pc.nparams]


def parameter_shift_gradient(pc: ParameterizedCircuit, theta: np.ndarray, H: PauliSum, precision: str = "double",
                             device: int = 0, params: Optional[Sequence[int]] = None) -> np.ndarray:
    """Two-term rule for exp(-i t P/2) gates (P a Pauli string); the four-term
    rule c+ [E(t+pi/2) - E(t-pi/2)] - c- [E(t+3pi/2) - E(t-3pi/2)],
    c+- = (sqrt2 +- 1)/(4 sqrt2), for controlled rotations (CRz), whose generator
    has eigenvalues {0, 0, +-1/2}."""
    theta = np.asarray(theta, dtype=np.float64)
    types = {}
    for gtype, _, p, _ in pc.ops:
        if p is not None:
            if gtype not in SHIFT_RULE_TYPES:
                raise qerr.ValueError(f"no shift rule for {gtype.name}")
            types[p] = gtype
    idx = range(pc.nparams) if params is None else params
    g = np.zeros(pc.nparams)

    def shifted(k, s):
        t = theta.copy()
        t[k] += s
        return expectation(pc, t, H, precision, device)

    cp, cm = (np.sqrt(2) + 1) / (4 * np.sqrt(2)), (np.sqrt(2) - 1) / (4 * np.sqrt(2))
    for k in idx:
        if types.get(k) == GateType.crz:
            g[k] = cp * (shifted(k, np.pi / 2) - shifted(k, -np.pi / 2)) - \
                cm * (shifted(k, 3 * np.pi / 2) - shifted(k, -3 * np.pi / 2))
        else:
            g[k] = 0.5 * (shifted(k, np.pi / 2) - shifted(k, -np.pi / 2))
    return g


def finite_difference_gradient(pc: ParameterizedCircuit, theta: np.ndarray, H: PauliSum, h: float = 1e-6,
                               precision: str = "double", device: int = 0,
                               params: Optional[Sequence[int]] = None) -> np.ndarray:
    theta = np.asarray(theta, dtype=np.float64)
    idx = range(pc.nparams) if params is None else params
    g = np.zeros(pc.nparams)
    for k in idx:
        tp = theta.copy()
        tp[k] += h
        tm = theta.copy()
        tm[k] -= h
        g[k] = (expectation(pc, tp, H, precision, device) - expectation(pc, tm, H, precision, device)) / (2 * h)
    return g


def _autograd_function():
    import torch

    class AdjointExpectation(torch.autograd.Function):
        """E(theta) = <psi(theta)|H|psi(theta)>; backward by the adjoint method."""

        @staticmethod
        def forward(ctx, theta, pc, H, precision="double"):
            E, g = adjoint_gradient(pc, theta.detach().cpu().numpy(), H, precision, theta.device.index or 0
                                    if theta.is_cuda else 0)
            ctx.save_for_backward(torch.as_tensor(g, dtype=theta.dtype, device=theta.device))
            return torch.as_tensor(E, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g, None, None, None

    return AdjointExpectation


def __getattr__(name):
    if name == "AdjointExpectation":
        return _autograd_function()
    raise AttributeError(name)
