"""Gate and matrix-type enumerations.

Mirrors the names and members of QuICT's ``GateType`` / ``MatrixType``
(QuICT/core/utils/gate_type.py:4-143) so that user code written against the
reference enumerations keeps working.  The enum *values* are this package's
own short descriptions.
"""
from enum import Enum


class GateType(Enum):
    h = "hadamard"
    hy = "self-inverse y-hadamard"
    s = "s"
    sdg = "s dagger"
    x = "pauli x"
    y = "pauli y"
    z = "pauli z"
    sx = "sqrt x"
    sy = "sqrt y"
    sw = "sqrt w"
    id = "identity"
    u1 = "u1"
    u2 = "u2"
    u3 = "u3"
    rx = "rx"
    ry = "ry"
    rz = "rz"
    ri = "ri"
    t = "t"
    tdg = "t dagger"
    phase = "phase"
    gphase = "global phase"
    cz = "controlled z"
    cx = "controlled x"
    cy = "controlled y"
    ch = "controlled hadamard"
    crz = "controlled rz"
    cu1 = "controlled u1"
    cu3 = "controlled u3"
    fsim = "fsim"
    rxx = "rxx"
    ryy = "ryy"
    rzz = "rzz"
    rzx = "rzx"
    swap = "swap"
    cswap = "controlled swap"
    iswap = "iswap"
    iswapdg = "iswap dagger"
    sqiswap = "sqrt iswap"
    ccx = "toffoli"
    ccz = "doubly controlled z"
    ccrz = "doubly controlled rz"
    measure = "measure"
    reset = "reset"
    barrier = "barrier"
    unitary = "unitary"
    perm = "permutation"
    perm_fx = "permutation fx"
    qft = "qft"
    iqft = "iqft"


class MatrixType(Enum):
    """How a gate's matrix is structured; drives kernel dispatch.

    normal        dense (2x2, or controlled 2x2 block)
    diagonal      diagonal
    control       diag(1, ..., 1, a)
    swap          permutation that swaps |01>,|10> (or |0>,|1> for X)
    reverse       anti-diagonal block [[0, a], [b, 0]] on the target
    special       no matrix (measure / reset / barrier / perm)
    diag_diag     two-target diagonal (Rzz)
    ctrl_normal   diag corners + middle 2x2 block (FSim)
    normal_normal two interleaved 2x2 blocks (Rxx, Ryy)
    diag_normal   block diagonal of two 2x2 blocks (Rzx)
    """
    normal = "normal"
    diagonal = "diagonal"
    control = "control"
    swap = "swap"
    reverse = "reverse"
    special = "special"
    diag_diag = "diag_diag"
    ctrl_normal = "ctrl_normal"
    normal_normal = "normal_normal"
    diag_normal = "diag_normal"


SPECIAL_GATE_SET = [
    GateType.measure, GateType.reset, GateType.barrier, GateType.unitary,
    GateType.perm, GateType.perm_fx, GateType.qft, GateType.iqft,
]

DIAGONAL_GATE_SET = [
    GateType.s, GateType.sdg, GateType.z, GateType.id, GateType.u1, GateType.rz,
    GateType.t, GateType.tdg, GateType.phase, GateType.gphase, GateType.cz,
    GateType.crz, GateType.cu1, GateType.rzz, GateType.ccrz,
]

PAULI_GATE_SET = [GateType.x, GateType.y, GateType.z, GateType.id]

CLIFFORD_GATE_SET = [
    GateType.x, GateType.y, GateType.z, GateType.h, GateType.s, GateType.sdg, GateType.cx,
]
