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


def _is_diag(m: np.ndarray) -> bool:
    return bool(np.all(m[~np.eye(m.shape[0], dtype=bool)] == 0))


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


def compile_gate(gate: BasicGate, n: int, rec: np.ndarray, keep: list, row_offset: int = 0,
                 conj: bool = False, dagger: bool = False) -> bool:
    """Fill one QS_GATE_DTYPE record.  Returns False for gates with no effect (id, barrier).

    ``row_offset`` shifts the amplitude bits (used for the row half of a
    density matrix); ``conj`` conjugates the matrix (column half);
    ``dagger`` writes the inverse (conjugate transpose).
    """
    gt = gate.type
    if gt in (GateType.id, GateType.barrier):
        return False
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
    rec["bits"][:] = 0
    rec["m"][:] = 0
    rec["big"] = 0

    def put(vals):
        v = np.asarray(vals, dtype=np.complex128).ravel()
        rec["m"][: 2 * len(v)] = np.column_stack([v.real, v.imag]).ravel()

    if gt == GateType.unitary and c == 0:
        # classify by structure (a Unitary carries no controls of its own)
        cc, blk = _controlled_split(mat, nargs)
        c, t = cc, nargs - cc
        mat_t = blk
        if _is_diag(mat_t) and t <= 4:
            mt = MatrixType.diagonal
        else:
            mt = MatrixType.normal
        full_diag = np.diag(mat_t)
    else:
        mat_t = mat[-(1 << t):, -(1 << t):]
        full_diag = None

    rec["ncontrols"] = c
    rec["ntargets"] = t
    rec["bits"][: nargs] = bits

    if mt in (MatrixType.diagonal, MatrixType.diag_diag) and t <= 4:
        d = np.diag(mat)[-(1 << t):] if full_diag is None else full_diag
        if t == 1 and d[0] == 1:
            rec["kind"] = nat.GATE_PHASE
            put([d[1]])
        else:
            rec["kind"] = nat.GATE_DIAG
            put(d)
    elif mt == MatrixType.control:
        rec["kind"] = nat.GATE_PHASE
        put([mat[-1, -1]])
    elif mt == MatrixType.swap and t == 1:
        rec["kind"] = nat.GATE_XPERM
        put([mat_t[0, 1], mat_t[1, 0]])
    elif mt == MatrixType.swap and t == 2:
        rec["kind"] = nat.GATE_SWAP
        put([mat_t[1, 2], mat_t[2, 1]])
    elif mt == MatrixType.reverse and t == 1:
        rec["kind"] = nat.GATE_XPERM
        put([mat[-2, -1], mat[-1, -2]])
    elif mt == MatrixType.reverse:
        # reverse-classified multi-target unitary: controls are all args but the last
        rec["ncontrols"] = nargs - 1
        rec["ntargets"] = 1
        rec["kind"] = nat.GATE_XPERM
        put([mat[-2, -1], mat[-1, -2]])
    else:
        # dense target block; a diagonal block becomes DIAG / PHASE
        if _is_diag(mat_t) and t <= 4:
            d = np.diag(mat_t)
            if np.all(d[:-1] == 1):
                rec["kind"] = nat.GATE_PHASE
                put([d[-1]])
            else:
                rec["kind"] = nat.GATE_DIAG
                put(d)
        elif t <= 2:
            rec["kind"] = nat.GATE_DENSE
            put(mat_t)
        else:
            big = np.ascontiguousarray(np.column_stack([mat_t.real.ravel(), mat_t.imag.ravel()]).ravel())
            keep.append(big)
            rec["kind"] = nat.GATE_DENSE
            rec["big"] = big.ctypes.data
    return True


def compile_gates(gates: List[BasicGate], n: int, row_offset: int = 0, conj: bool = False,
                  dagger: bool = False) -> Tuple[np.ndarray, list]:
    """Records for a run of unitary gates; ``keep`` holds arrays referenced by pointer."""
    arr = np.zeros(len(gates), dtype=nat.QS_GATE_DTYPE)
    keep: list = []
    k = 0
    for g in gates:
        if compile_gate(g, n, arr[k], keep, row_offset, conj, dagger):
            k += 1
    return arr[:k].copy() if k != len(arr) else arr, keep


def superop_record(kraus: List[np.ndarray], args: List[int], n: int, rec, keep: list):
    """Record for a channel sum_k K (.) K^dagger on density-matrix qubits ``args``:
    dense superoperator sum_k K (x) conj(K) on [row bits..., column bits...]."""
    m = len(args)
    S = np.zeros((1 << (2 * m), 1 << (2 * m)), dtype=np.complex128)
    for K in kraus:
        K = np.asarray(K, dtype=np.complex128)
        S += np.kron(K, K.conj())
    row_bits = [2 * n - 1 - a for a in args]
    col_bits = [n - 1 - a for a in args]
    rec["kind"] = nat.GATE_DENSE
    rec["ncontrols"] = 0
    rec["ntargets"] = 2 * m
    rec["bits"][:] = 0
    rec["bits"][: 2 * m] = row_bits + col_bits
    rec["m"][:] = 0
    rec["big"] = 0
    if 2 * m <= 2:
        v = S.ravel()
        rec["m"][: 2 * len(v)] = np.column_stack([v.real, v.imag]).ravel()
    else:
        big = np.ascontiguousarray(np.column_stack([S.real.ravel(), S.imag.ravel()]).ravel())
        keep.append(big)
        rec["big"] = big.ctypes.data
