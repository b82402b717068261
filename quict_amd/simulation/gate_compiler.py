"""Translate QuICT-style gates into ``qs_gate`` records for the HIP library.

This replaces the dispatch of StateVectorSimulator.apply_gate
(QuICT/simulation/state_vector/statevector_simulator.py:202-496), which picks a
kernel by (MatrixType, #controls, #targets) and passes reversed qubit indices
(``n-1-q``).  Here the same decision produces a compact record instead:

  MatrixType            record kind   payload
  diagonal / diag_diag  DIAG          last 2^t diagonal entries (the all-controls-one block)
  control               PHASE         matrix[-1, -1]
  swap, 1 target (X)    XPERM         (1, 1)
  swap, 2 targets       SWAP          target block [1, 2], [2, 1]
  reverse               XPERM         matrix[-2, -1], matrix[-1, -2]
  everything else       DENSE         target block (controls folded in as controls)

Gate arguments are ``cargs + targs`` with argument 0 the most significant bit
of the gate-matrix index, mapped to amplitude bit ``n-1-q``.
"""
from typing import List, Tuple

import numpy as np

from ..core import BasicGate, GateType, MatrixType
from .. import _native as nat

_NON_UNITARY = (GateType.measure, GateType.reset, GateType.barrier, GateType.perm)


_OFFDIAG = {}


def _is_diag(m: np.ndarray) -> bool:
    d = m.shape[0]
    mask = _OFFDIAG.get(d)
    if mask is None:
        mask = _OFFDIAG[d] = ~np.eye(d, dtype=bool)
    return not m[mask].any()


def _controlled_split(mat: np.ndarray, nargs: int) -> Tuple[int, np.ndarray]:
    """Largest c such that mat = I on every block except the last 2^(nargs-c) one."""
    d = 1 << nargs
    for c in range(nargs - 1, 0, -1):
        t = nargs - c
        blk = 1 << t
        top = mat[: d - blk, : d - blk]
        if (np.array_equal(top, np.identity(d - blk, dtype=mat.dtype))
                and not np.any(mat[: d - blk, d - blk:]) and not np.any(mat[d - blk:, : d - blk])):
            return c, mat[d - blk:, d - blk:]
    return 0, mat


def _translate(gate: BasicGate, n: int, keep: list, row_offset: int, conj: bool, dagger: bool):
    """(kind, ncontrols, ntargets, bits, payload, big) of one gate, or None for
    gates with no effect.  ``payload`` is a complex128 vector (inline
    coefficients) or None; ``big`` the address of a wider matrix kept in ``keep``."""
    gt = gate.type
    if gt in (GateType.id, GateType.barrier):
        return None
    if gt in _NON_UNITARY:
        raise ValueError(f"{gt.name} is not a unitary gate record")
    args = list(gate.cargs) + list(gate.targs)
    bits = [row_offset + n - 1 - a for a in args]
    mat = np.asarray(gate.matrix, dtype=np.complex128)
    if conj:
        mat = mat.conj()
    if dagger:
        mat = mat.conj().T
    nargs = len(args)
    mt = gate.matrix_type
    c, t = gate.controls, gate.targets

    if gt == GateType.unitary and c == 0:
        # classify by structure (a Unitary carries no controls of its own)
        cc, blk = _controlled_split(mat, nargs)
        c, t = cc, nargs - cc
        mat_t = blk
        mt = MatrixType.diagonal if (_is_diag(mat_t) and t <= 4) else MatrixType.normal
        full_diag = np.diag(mat_t)
    else:
        mat_t = mat[-(1 << t):, -(1 << t):]
        full_diag = None

    if mt in (MatrixType.diagonal, MatrixType.diag_diag) and t <= 4:
        d = np.diag(mat)[-(1 << t):] if full_diag is None else full_diag
        if t == 1 and d[0] == 1:
            return nat.GATE_PHASE, c, t, bits, d[1:2], 0
        return nat.GATE_DIAG, c, t, bits, d, 0
    if mt == MatrixType.control:
        return nat.GATE_PHASE, c, t, bits, mat[-1:, -1], 0
    if mt == MatrixType.swap and t == 1:
        return nat.GATE_XPERM, c, t, bits, np.array([mat_t[0, 1], mat_t[1, 0]]), 0
    if mt == MatrixType.swap and t == 2:
        return nat.GATE_SWAP, c, t, bits, np.array([mat_t[1, 2], mat_t[2, 1]]), 0
    if mt == MatrixType.reverse:
        # reverse-classified gate (also a multi-target unitary): controls are all args but the last
        return nat.GATE_XPERM, nargs - 1, 1, bits, np.array([mat[-2, -1], mat[-1, -2]]), 0
    # dense target block; a diagonal block becomes DIAG / PHASE
    if t <= 4 and _is_diag(mat_t):
        d = np.diag(mat_t)
        if np.all(d[:-1] == 1):
            return nat.GATE_PHASE, c, t, bits, d[-1:], 0
        return nat.GATE_DIAG, c, t, bits, d, 0
    if t <= 2:
        return nat.GATE_DENSE, c, t, bits, mat_t.ravel(), 0
    big = np.ascontiguousarray(mat_t).view(np.float64).ravel().copy()
    keep.append(big)
    return nat.GATE_DENSE, c, t, bits, None, big.ctypes.data


def compile_gate(gate: BasicGate, n: int, rec: np.ndarray, keep: list, row_offset: int = 0,
                 conj: bool = False, dagger: bool = False) -> bool:
    """Fill one QS_GATE_DTYPE record.  Returns False for gates with no effect (id, barrier).

    ``row_offset`` shifts the amplitude bits (used for the row half of a
    density matrix); ``conj`` conjugates the matrix (column half);
    ``dagger`` writes the inverse (conjugate transpose).
    """
    r = _translate(gate, n, keep, row_offset, conj, dagger)
    if r is None:
        return False
    kind, c, t, bits, payload, big = r
    rec["kind"] = kind
    rec["ncontrols"] = c
    rec["ntargets"] = t
    rec["bits"][:] = 0
    rec["bits"][: len(bits)] = bits
    rec["m"][:] = 0
    if payload is not None:
        v = np.ascontiguousarray(payload, dtype=np.complex128).view(np.float64)
        rec["m"][: len(v)] = v
    rec["big"] = big
    return True


# Parameter-free gate types of fixed shape: the record (kind, controls, targets,
# payload) depends only on the type and precision, so it is translated once and
# reused; only the bit positions differ between instances.
_STATIC_TYPES = frozenset([
    GateType.h, GateType.hy, GateType.s, GateType.sdg, GateType.x, GateType.y, GateType.z, GateType.sx,
    GateType.sy, GateType.sw, GateType.t, GateType.tdg, GateType.cx, GateType.cy, GateType.cz, GateType.ch,
    GateType.swap, GateType.iswap, GateType.iswapdg, GateType.sqiswap, GateType.cswap, GateType.ccx,
    GateType.ccz])
_STATIC_RECORDS = {}
# One-angle / three-angle rotations whose payloads are computed for all gates of
# a call at once with the closed forms of the gate module (core/gate.py).
_ROTATION_TYPES = frozenset([GateType.rx, GateType.ry, GateType.rz, GateType.crz, GateType.u1, GateType.u3])


def _rotation_records(gt: GateType, P: np.ndarray, dt):
    """(kind per gate, payload rows, degenerate mask) for rotations of type gt with
    parameter rows P, matching _translate: the matrix in the gate's precision dt,
    the same record-kind decisions.  Degenerate gates (a dense matrix that is
    exactly diagonal) go back to the generic path."""
    G = P.shape[0]
    cast = (lambda a: a.astype(dt).astype(np.complex128))
    if gt in (GateType.rx, GateType.ry, GateType.u3):
        th = P[:, 0]
        c, s = np.cos(th / 2), np.sin(th / 2)
        if gt == GateType.rx:
            pay = np.stack([c + 0j, 1j * -s, 1j * -s, c + 0j], 1)
        elif gt == GateType.ry:
            pay = np.stack([c + 0j, -s + 0j, s + 0j, c + 0j], 1)
        else:
            pay = np.stack([np.cos(th / 2) + 0j, -np.exp(1j * P[:, 2]) * np.sin(th / 2),
                            np.exp(1j * P[:, 1]) * np.sin(th / 2), np.exp(1j * (P[:, 1] + P[:, 2])) * np.cos(th / 2)],
                           1)
        pay = cast(pay)
        degenerate = (pay[:, 1] == 0) & (pay[:, 2] == 0)
        return np.full(G, nat.GATE_DENSE, dtype=np.int32), pay, degenerate
    if gt == GateType.u1:
        pay = cast(np.exp(1j * P[:, 0]))[:, None]
        return np.full(G, nat.GATE_PHASE, dtype=np.int32), pay, np.zeros(G, dtype=bool)
    # rz, crz: diagonal (e^{-i th/2}, e^{i th/2}); PHASE when the first entry is 1
    d = cast(np.stack([np.exp(-P[:, 0] / 2 * 1j), np.exp(P[:, 0] / 2 * 1j)], 1))
    phase = d[:, 0] == 1
    kind = np.where(phase, nat.GATE_PHASE, nat.GATE_DIAG).astype(np.int32)
    pay = np.where(phase[:, None], np.stack([d[:, 1], np.zeros(G, dtype=np.complex128)], 1), d)
    return kind, pay, np.zeros(G, dtype=bool)


def compile_gates(gates: List[BasicGate], n: int, row_offset: int = 0, conj: bool = False,
                  dagger: bool = False) -> Tuple[np.ndarray, list]:
    """Records for a run of unitary gates; ``keep`` holds arrays referenced by pointer.
    The fields are gathered in plain arrays and written column-wise.  Fixed gate
    types reuse one translated record per type; Rx / Ry / Rz / CRz / U1 / U3
    payloads are computed for the whole run at once (plain records only: no row
    offset, conjugation or inverse); everything else goes through _translate."""
    keep: list = []
    G = len(gates)
    kind = np.zeros(G, dtype=np.int32)
    nc = np.zeros(G, dtype=np.int32)
    nt = np.zeros(G, dtype=np.int32)
    bits = np.zeros((G, nat.QS_MAX_GATE_BITS), dtype=np.int32)
    m = np.zeros((G, 32), dtype=np.float64)
    big = np.zeros(G, dtype=np.uint64)
    fast = row_offset == 0 and not conj and not dagger
    rot = {}   # (type, precision) -> ([slot], [gate])
    static = {}  # (type, precision) -> [slot]
    b1k, b1 = [], []              # one-argument gates: slot, argument
    b2k, b2a, b2b = [], [], []    # two-argument gates: slot, arguments
    k = 0
    for g in gates:
        gt = g.type
        if fast and (gt in _STATIC_TYPES or gt in _ROTATION_TYPES):
            args = g.cargs + g.targs
            if len(args) == 1:
                b1k.append(k)
                b1.append(args[0])
            elif len(args) == 2:
                b2k.append(k)
                b2a.append(args[0])
                b2b.append(args[1])
            else:
                bits[k, : len(args)] = [n - 1 - a for a in args]
            key = (gt, g.precision)
            if gt in _STATIC_TYPES:
                lst = static.get(key)
                if lst is None:
                    lst = static[key] = [g]  # the first entry: a gate to translate once
                lst.append(k)
            else:
                lst = rot.setdefault(key, ([], []))
                lst[0].append(k)
                lst[1].append(g)
            k += 1
            continue
        r = _translate(g, n, keep, row_offset, conj, dagger)
        if r is None:
            continue
        kind[k], nc[k], nt[k] = r[0], r[1], r[2]
        bits[k, : len(r[3])] = r[3]
        if r[4] is not None:
            v = np.ascontiguousarray(r[4], dtype=np.complex128).view(np.float64)
            m[k, : len(v)] = v
        big[k] = r[5]
        k += 1
    if b1k:
        bits[b1k, 0] = n - 1 - np.asarray(b1)
    if b2k:
        bits[b2k, 0] = n - 1 - np.asarray(b2a)
        bits[b2k, 1] = n - 1 - np.asarray(b2b)
    for key, lst in static.items():
        rec = _STATIC_RECORDS.get(key)
        if rec is None:
            r = _translate(lst[0], n, keep, 0, False, False)
            rec = _STATIC_RECORDS[key] = (r[0], r[1], r[2], None if r[4] is None else
                                          np.ascontiguousarray(r[4], dtype=np.complex128).view(np.float64))
        idx = np.asarray(lst[1:])
        kind[idx], nc[idx], nt[idx] = rec[0], rec[1], rec[2]
        if rec[3] is not None:
            m[idx, : len(rec[3])] = rec[3]
    for (gt, dt), (slots, gl) in rot.items():
        nc[slots], nt[slots] = gl[0].controls, gl[0].targets
    for (gt, dt), (slots, gl) in rot.items():
        slots = np.asarray(slots)
        P = np.array([g.pargs for g in gl], dtype=np.float64)
        kk, pay, degenerate = _rotation_records(gt, P, dt)
        kind[slots] = kk
        m[slots, : 2 * pay.shape[1]] = np.ascontiguousarray(pay).view(np.float64)
        for i in np.nonzero(degenerate)[0]:
            r = _translate(gl[i], n, keep, 0, False, False)
            j = slots[i]
            kind[j], nc[j], nt[j] = r[0], r[1], r[2]
            bits[j, :] = 0
            bits[j, : len(r[3])] = r[3]
            m[j, :] = 0
            v = np.ascontiguousarray(r[4], dtype=np.complex128).view(np.float64)
            m[j, : len(v)] = v
    arr = np.zeros(k, dtype=nat.QS_GATE_DTYPE)
    arr["kind"] = kind[:k]
    arr["ncontrols"] = nc[:k]
    arr["ntargets"] = nt[:k]
    arr["bits"] = bits[:k]
    arr["m"] = m[:k]
    arr["big"] = big[:k]
    return arr, keep


def superop_record(kraus: List[np.ndarray], args: List[int], n: int, rec, keep: list):
    """Record for a channel sum_k K (.) K^dagger on density-matrix qubits ``args``:
    dense superoperator sum_k K (x) conj(K) on [row bits..., column bits...]."""
    m = len(args)
    d = 1 << m
    Ks = np.asarray(kraus, dtype=np.complex128).reshape(-1, d, d)
    # sum_k kron(K_k, conj(K_k)): S[(a, c), (b, e)] = sum_k K_k[a, b] conj(K_k[c, e])
    S = np.einsum("kab,kce->acbe", Ks, Ks.conj()).reshape(d * d, d * d)
    row_bits = [2 * n - 1 - a for a in args]
    col_bits = [n - 1 - a for a in args]
    rec["kind"] = nat.GATE_DENSE
    rec["ncontrols"] = 0
    rec["ntargets"] = 2 * m
    rec["bits"][:] = 0
    rec["bits"][: 2 * m] = row_bits + col_bits
    rec["m"][:] = 0
    rec["big"] = 0
    v = np.ascontiguousarray(S).view(np.float64).ravel()
    if 2 * m <= 2:
        rec["m"][: len(v)] = v
    else:
        big = v.copy()
        keep.append(big)
        rec["big"] = big.ctypes.data
