"""CPU oracle of the circuit-matrix / unitary-simulator path — TEST INFRASTRUCTURE ONLY.

Restates, in numpy, the reference's algorithm:

* ``circuit_matrix``          CircuitMatrix.get_unitary_matrix
                              (QuICT/core/utils/circuit_matrix.py:11-226): gates
                              layered by depth, collected into MatrixGroups,
                              each group and then the group list combined by
                              merging the adjacent pair with the fewest
                              arguments (tensor for disjoint pairs, otherwise
                              expand both to the union and multiply);
* ``matrix_tensor_i``         MatrixTensorI, I_n (x) A (x) I_m
                              (QuICT/ops/linalg/cpu_calculator.py:13-38);
* ``matrix_permutation``      MatrixPermutation with mapping_augment
                              (cpu_calculator.py:41-66, QuICT/ops/utils/utils.py:45-52);
* ``run``                     UnitarySimulator.run (QuICT/simulation/unitary/
                              unitary_simulator.py:59-106): U x unless allclose(U, I);
* ``bruteforce_matrix``       product of every gate expanded by
                              matrix_product_to_circuit (QuICT/core/utils/utils.py:8-63),
                              the independent check of ``circuit_matrix``.

Pinned by: brute force on random 1/2-qubit circuits (every merge branch, both
argument orders, idle qubits), the reference's golden 10-qubit state
(U|0> == unit_test/simulation/data/state_vector.npy) through
``bruteforce_matrix``, unitarity, and circuit * inverse = identity.
Like the reference, ``circuit_matrix`` rejects gates on 3 or more qubits
(Circuit.matrix decomposes them first, circuit_info.py:197-198).
"""
from typing import List, Tuple

import numpy as np

from quict_amd.core import MatrixType


def mapping_augment(mapping: List[int]) -> np.ndarray:
    """Index map of a qubit permutation (ops/utils/utils.py:45-52): bit k (MSB
    first) of res[i] is bit mapping[k] (MSB first) of i."""
    n = len(mapping)
    i = np.arange(1 << n, dtype=np.int64)
    res = np.zeros(1 << n, dtype=np.int64)
    for k, mk in enumerate(mapping):
        res |= ((i >> (n - 1 - mk)) & 1) << (n - 1 - k)
    return res


def matrix_permutation(A: np.ndarray, mapping: List[int]) -> np.ndarray:
    """A'[i][j] = A[res[i]][res[j]] (cpu_calculator.py:41-66)."""
    if A.shape[0] != (1 << len(mapping)):
        raise IndexError("Indices do not match!")
    idx = mapping_augment(list(mapping))
    return A[idx, :][:, idx]


def matrix_tensor_i(A: np.ndarray, n: int, m: int) -> np.ndarray:
    """I_n (x) A (x) I_m (cpu_calculator.py:13-38)."""
    return np.kron(np.kron(np.identity(n, dtype=A.dtype), A), np.identity(m, dtype=A.dtype))


def get_gates_order_by_depth(gates: list) -> List[list]:
    """circuit_matrix.py:11-35: each gate goes one layer after the deepest layer
    whose qubits it shares, or to layer 0."""
    layers = [[gates[0]]]
    layer_args = [set(gates[0].cargs + gates[0].targs)]
    for g in gates[1:]:
        ga = set(g.cargs + g.targs)
        for i in range(len(layer_args) - 1, -1, -1):
            if ga & layer_args[i]:
                if i == len(layer_args) - 1:
                    layers.append([g])
                    layer_args.append(ga)
                else:
                    layers[i + 1].append(g)
                    layer_args[i + 1] |= ga
                break
            if i == 0:
                layers[0].append(g)
                layer_args[0] |= ga
    return layers


class _Group:
    """MatrixGroup (circuit_matrix.py:38-54)."""

    def __init__(self, matrix, args, blocked=frozenset()):
        self.args = set(args)
        self.value = [(matrix, list(args))]
        self.blocked = set(blocked)

    def append(self, matrix, args):
        self.args |= set(args)
        self.value.append((matrix, list(args)))


def _find_related(groups: List[List[_Group]], args: List[int]):
    """circuit_matrix.py:127-149: (hits, blocked_layer).  Scans layers from the
    last; returns at the first layer with any shared qubit; stops after a
    layer whose blocked set meets the gate (a layer index of 0 does not stop)."""
    blocked = 0
    hits: List[Tuple[int, int, int]] = []
    for li in range(len(groups) - 1, -1, -1):
        for pi, g in enumerate(groups[li]):
            shared = set(args) & g.args
            for a in shared:
                hits.append((a, li, pi))
            if set(args) & g.blocked:
                blocked = li
        if hits:
            return hits, 0
        if blocked:
            break
    return hits, blocked


def _tensor_unitary(u: np.ndarray, uargs: List[int], ext: List[int]) -> np.ndarray:
    """circuit_matrix.py:174-198: expand u (on uargs) to the ordered qubit list ext."""
    pos = [ext.index(a) for a in uargs]
    d = np.diff(pos)
    if np.allclose(np.abs(d), 1):
        if np.allclose(d, -1):
            u = matrix_permutation(u, list(range(len(uargs) - 1, -1, -1)))
        return matrix_tensor_i(u, 1 << min(pos), 1 << (len(ext) - 1 - max(pos)))
    t = matrix_tensor_i(u, 1, 1 << (len(ext) - len(uargs)))
    targs = list(uargs) + [a for a in ext if a not in uargs]
    return matrix_permutation(t, [ext.index(a) for a in targs])


def merge_gates(u1, a1, u2, a2):
    """circuit_matrix.py:151-172: u2 applied after u1."""
    s1, s2 = set(a1), set(a2)
    if not (s1 & s2):
        return np.kron(u1, u2), list(a1) + list(a2)
    if s1 == s2:
        u2 = matrix_permutation(u2, [a1.index(a) for a in a2])
    if len(a1) >= len(a2):
        union = list(a1) + [a for a in a2 if a not in a1]
    else:
        union = list(a2) + [a for a in a1 if a not in a2]
    if len(union) != len(s1):
        u1 = _tensor_unitary(u1, a1, union)
    if len(union) != len(s2):
        u2 = _tensor_unitary(u2, a2, union)
    return u2 @ u1, union


def _smallest_adjacent(nums: List[int]) -> Tuple[int, int]:
    """circuit_matrix.py:217-226: first adjacent pair with the smallest sum."""
    best, left = sum(nums) * 2, -1
    for i in range(1, len(nums)):
        if nums[i] + nums[i - 1] < best:
            best, left = nums[i] + nums[i - 1], i - 1
    return left, left + 1


def combine(items: List[Tuple[np.ndarray, List[int]]]):
    """circuit_matrix.py:200-215 (the count of a merged pair is the sum of the two counts)."""
    items = list(items)
    nums = [len(a) for _, a in items]
    while len(nums) > 1:
        l, r = _smallest_adjacent(nums)
        items[l] = merge_gates(items[l][0], items[l][1], items[r][0], items[r][1])
        del items[r]
        nums[l] += nums[r]
        del nums[r]
    return items[0]


def circuit_matrix(gates: list, n: int, mini_arg: int = 0) -> np.ndarray:
    """CircuitMatrix.get_unitary_matrix (circuit_matrix.py:69-125), complex128."""
    if not gates:
        return np.identity(1 << n, dtype=np.complex128)
    groups: List[List[_Group]] = [[]]
    for layer in get_gates_order_by_depth(gates):
        for g in layer:
            if g.controls + g.targets >= 3:
                raise TypeError("CircuitMatrix takes 1- or 2-qubit gates")
            if g.matrix_type == MatrixType.special:
                continue
            args = list(g.cargs + g.targs)
            m = np.asarray(g.matrix, dtype=np.complex128)
            if len(args) == 2 and args[0] > args[1]:
                args.sort()
                m = matrix_permutation(m, [1, 0])
            hits, blocked = _find_related(groups, args)
            if not hits:
                groups[blocked].append(_Group(m, args))
            elif len(hits) == 2:
                (_, l0, p0), (_, l1, p1) = hits
                if p0 == p1:
                    groups[l0][p0].append(m, args)
                else:
                    blk = groups[l0][p0].args | groups[l1][p1].args
                    ng = _Group(m, args, blk)
                    if l0 == len(groups) - 1:
                        groups.append([ng])
                    else:
                        groups[l0 + 1].append(ng)
            else:
                _, l0, p0 = hits[0]
                groups[l0][p0].append(m, args)
    merged = [combine(g.value) for layer in groups for g in layer]
    u, uargs = combine(merged)
    base = list(range(mini_arg, n))
    if uargs != base:
        u = _tensor_unitary(u, uargs, base)
    return u


def expand(m: np.ndarray, args: List[int], n: int) -> np.ndarray:
    """matrix_product_to_circuit (utils.py:8-63): the 2^n x 2^n matrix of m acting
    on qubits args (args[0] = MSB of m's index; qubit q = index bit n-1-q)."""
    D = 1 << n
    idx = np.arange(D, dtype=np.int64)
    sub = np.zeros(D, dtype=np.int64)
    rest_mask = D - 1
    for k, a in enumerate(args):
        sub |= ((idx >> (n - 1 - a)) & 1) << (len(args) - 1 - k)
        rest_mask ^= 1 << (n - 1 - a)
    rest = idx & rest_mask
    same = rest[:, None] == rest[None, :]
    return np.where(same, np.asarray(m, dtype=np.complex128)[sub[:, None], sub[None, :]], 0)


def bruteforce_matrix(gates: list, n: int) -> np.ndarray:
    """Product of every non-special gate expanded to n qubits."""
    u = np.identity(1 << n, dtype=np.complex128)
    for g in gates:
        if g.matrix_type == MatrixType.special:
            continue
        u = expand(g.matrix, list(g.cargs + g.targs), n) @ u
    return u


def run(U: np.ndarray, state_vector: np.ndarray = None) -> np.ndarray:
    """unitary_simulator.py:59-106: |0..0> (or the given state), then U x unless U == I (allclose)."""
    D = U.shape[0]
    if state_vector is None:
        x = np.zeros(D, dtype=U.dtype)
        x[0] = 1
    else:
        x = np.asarray(state_vector, dtype=U.dtype)
    if np.allclose(U, np.identity(D, dtype=U.dtype)):
        return x
    return U @ x
