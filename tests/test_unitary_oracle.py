"""Pins the circuit-matrix oracle (CPU only, no GPU).

``oracle.unitary.circuit_matrix`` restates the reference's CircuitMatrix
(depth layers, matrix groups, adjacent-pair merging with tensor / expansion /
qubit permutation).  It is checked against an independent brute-force product
of expanded gate matrices, that product against the reference's golden
10-qubit state and against numpy's kron, and both against unitarity and
circuit * inverse = identity — so a wrong merge order, a transposed operand,
a reversed qubit order or a dropped gate fails a check.
"""
import os
import random

import numpy as np
import pytest

from quict_amd.core import (Circuit, GateType, H, CX, Swap, Measure, Barrier, Unitary, CCX)
from quict_amd.core import qasm
from oracle import unitary as ou
from oracle import statevector as osv

TWO_Q = [GateType.cx, GateType.cy, GateType.cz, GateType.ch, GateType.crz, GateType.cu1, GateType.cu3,
         GateType.fsim, GateType.rxx, GateType.ryy, GateType.rzz, GateType.rzx, GateType.swap, GateType.iswap,
         GateType.sqiswap]
ONE_Q = [GateType.h, GateType.x, GateType.y, GateType.z, GateType.s, GateType.t, GateType.sx, GateType.u3,
         GateType.rx, GateType.ry, GateType.rz, GateType.phase]


def random_unitary(k, rng):
    a = rng.randn(1 << k, 1 << k) + 1j * rng.randn(1 << k, 1 << k)
    q, r = np.linalg.qr(a)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_circuit(n, size, seed, specials=False, dense2=False):
    c = Circuit(n)
    random.seed(seed)
    np.random.seed(seed)
    types = ONE_Q + (TWO_Q if n >= 2 else [])
    c.random_append(size, typelist=types, random_params=True)
    rng = np.random.RandomState(seed + 100)
    if dense2 and n >= 2:
        for _ in range(3):
            a, b = rng.choice(n, 2, replace=False)
            Unitary(random_unitary(2, rng)) | c([int(a), int(b)])
            H | c(int(rng.randint(n)))
    if specials:
        Barrier | c(0)
        Measure | c(n - 1)
        H | c(0)
    return c


def test_mapping_augment_and_permutation_known_values():
    # swapping the two qubits of a 2-qubit index: 01 <-> 10
    assert list(ou.mapping_augment([1, 0])) == [0, 2, 1, 3]
    assert list(ou.mapping_augment([0, 1])) == [0, 1, 2, 3]
    # 3 qubits, cyclic: bit k of res[i] is bit mapping[k] of i (MSB first)
    res = ou.mapping_augment([2, 0, 1])
    for i in range(8):
        b = [(i >> 2) & 1, (i >> 1) & 1, i & 1]
        assert res[i] == (b[2] << 2) | (b[0] << 1) | b[1]
    # CX with control/target exchanged = control on the second qubit
    cx = np.asarray(CX.matrix)
    rev = ou.matrix_permutation(cx, [1, 0])
    want = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]])
    assert np.array_equal(rev, want)


def test_matrix_tensor_i_matches_the_block_loop():
    rng = np.random.RandomState(1)
    A = rng.randn(2, 2) + 1j * rng.randn(2, 2)
    n, m = 4, 2
    out = ou.matrix_tensor_i(A, n, m)
    # the reference's loop: block (k, i, j) holds A[i, j] * I_m
    ref = np.zeros((n * m * 2, n * m * 2), dtype=complex)
    for i in range(2):
        for j in range(2):
            for k in range(n):
                r0, c0 = k * m * 2 + i * m, k * m * 2 + j * m
                ref[r0:r0 + m, c0:c0 + m] = A[i, j] * np.identity(m)
    assert np.array_equal(out, ref)


def test_expand_matches_kron_and_reversed_order():
    rng = np.random.RandomState(2)
    u1 = random_unitary(1, rng)
    u2 = random_unitary(2, rng)
    n = 4
    assert np.allclose(ou.expand(u1, [1], n), np.kron(np.kron(np.identity(2), u1), np.identity(4)), atol=0)
    assert np.allclose(ou.expand(u2, [1, 2], n), np.kron(np.kron(np.identity(2), u2), np.identity(2)), atol=0)
    # args in reverse order = the qubit-swapped matrix on sorted args
    sw = np.asarray(Swap.matrix)
    assert np.allclose(ou.expand(u2, [2, 1], n), ou.expand(sw @ u2 @ sw, [1, 2], n), atol=1e-15)


@pytest.mark.parametrize("n,size,seed", [(1, 6, 0), (2, 12, 1), (3, 30, 2), (4, 40, 3), (5, 60, 4), (6, 80, 5),
                                         (6, 80, 6), (7, 50, 7)])
def test_circuit_matrix_restatement_equals_brute_force(n, size, seed):
    c = random_circuit(n, size, seed, specials=(seed % 2 == 0), dense2=True)
    U = ou.circuit_matrix(c.gates, n)
    B = ou.bruteforce_matrix(c.gates, n)
    assert np.max(np.abs(U - B)) < 1e-12
    assert np.max(np.abs(U @ U.conj().T - np.identity(1 << n))) < 1e-12


def test_circuit_matrix_idle_qubits_and_empty():
    c = Circuit(4)
    H | c(2)
    CX | c([2, 1])
    U = ou.circuit_matrix(c.gates, 4)
    assert np.max(np.abs(U - ou.bruteforce_matrix(c.gates, 4))) < 1e-15
    assert np.array_equal(ou.circuit_matrix([], 3), np.identity(8))


def test_circuit_times_inverse_is_identity():
    c = random_circuit(5, 60, 11)
    inv = Circuit(5)
    for g in reversed(c.gates):
        g.inverse() | inv(list(g.cargs + g.targs))
    both = list(c.gates) + list(inv.gates)
    U = ou.circuit_matrix(both, 5)
    assert np.max(np.abs(U - np.identity(32))) < 1e-12


def test_three_qubit_gates_rejected_like_the_reference():
    c = Circuit(3)
    CCX | c([0, 1, 2])
    with pytest.raises(TypeError):
        ou.circuit_matrix(c.gates, 3)


def test_bruteforce_reproduces_the_golden_vector(golden_dir):
    """U|0> of the reference's golden circuit (3-qubit gates included) equals
    its golden state vector."""
    c = qasm.load_file(os.path.join(golden_dir, "random_circuit_for_correction.qasm"))
    sv = np.load(os.path.join(golden_dir, "state_vector.npy"))
    U = ou.bruteforce_matrix(c.gates, c.width())
    assert np.max(np.abs(U[:, 0] - sv)) < 1e-12
    assert np.max(np.abs(ou.run(U) - sv)) < 1e-12


def test_run_equals_the_statevector_oracle():
    n = 6
    c = random_circuit(n, 70, 21, dense2=True)
    U = ou.circuit_matrix(c.gates, n)
    rng = np.random.RandomState(3)
    x = rng.randn(1 << n) + 1j * rng.randn(1 << n)
    x /= np.linalg.norm(x)
    assert np.max(np.abs(ou.run(U, x) - osv.run(c, state_vector=x))) < 1e-12
    # the reference returns x untouched when U is (close to) the identity
    assert ou.run(np.identity(1 << n, dtype=complex), x) is not None
    assert np.array_equal(ou.run(np.identity(1 << n, dtype=complex), x), x)
