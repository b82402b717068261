"""Seeded circuit workloads used by the benchmark and the parity tests.

Each builder fixes its seeds *after* constructing the Circuit (so no object
construction can shift the RNG streams) and draws gates with
``Circuit.random_append`` (numpy's legacy global RandomState for the gate type
and angles, CPython ``random`` for the qubits).
"""
import random

import numpy as np

from .core import Circuit, GateType, CX, Ry, Rz

RANDOM_GATE_MIX = [
    GateType.h, GateType.x, GateType.y, GateType.z, GateType.s, GateType.t, GateType.rx, GateType.ry,
    GateType.rz, GateType.u3, GateType.cx, GateType.cz, GateType.crz, GateType.swap,
]
# 60% one-qubit (10 types x 0.06) / 40% two-qubit (4 types x 0.10)
RANDOM_GATE_PROBS_60_40 = [0.06] * 10 + [0.10] * 4


def random_mix_circuit(n: int, ngates: int, seed: int, probabilities=None) -> Circuit:
    """Random circuit over {H,X,Y,Z,S,T,Rx,Ry,Rz,U3,CX,CZ,CRz,Swap}, angles U(0, 2pi)."""
    c = Circuit(n)
    random.seed(seed)
    np.random.seed(seed)
    c.random_append(ngates, typelist=list(RANDOM_GATE_MIX), random_params=True, probabilities=probabilities)
    return c


def sv20_circuit() -> Circuit:
    """20 qubits, 400 uniform gates from the mix, seed 1."""
    return random_mix_circuit(20, 400, 1)


def sv30_circuit(n: int = 30, ngates: int = 2000) -> Circuit:
    """30 qubits, 2000 gates, 60/40 one-/two-qubit mix, seed 2."""
    return random_mix_circuit(n, ngates, 2, RANDOM_GATE_PROBS_60_40)


def dm_noise_model():
    """Depolarizing(p=1e-3) after every one-qubit gate of the mix, amplitude
    damping (gamma=1e-2) tensored with itself after every two-qubit gate."""
    from .core.noise import DepolarizingError, DampingError, NoiseModel
    names = [t.name for t in RANDOM_GATE_MIX]
    nm = NoiseModel(basic_gates=names)
    nm.add_noise_for_all_qubits(DepolarizingError(1e-3), names[:10])
    d = DampingError(0.01, 0, 0)
    nm.add_noise_for_all_qubits(d.tensor(d), names[10:])
    return nm


def dm14_circuit(n: int = 14, ngates: int = 300) -> Circuit:
    """14 qubits, 300 uniform gates from the mix, seed 5 (noisy density-matrix workload)."""
    return random_mix_circuit(n, ngates, 5)


def hea_params(n: int, layers: int, seed: int) -> np.ndarray:
    rng = np.random.RandomState(seed)
    return rng.uniform(-np.pi, np.pi, size=(layers, 2, n))


def hea_circuit(n: int, theta: np.ndarray) -> Circuit:
    """Hardware-efficient ansatz: per layer Ry(theta[l,0,q]) and Rz(theta[l,1,q]) on every
    qubit, then CX(i, i+1) for i = 0..n-2."""
    c = Circuit(n)
    for layer in theta:
        for q in range(n):
            Ry(float(layer[0, q])) | c(q)
            Rz(float(layer[1, q])) | c(q)
        for q in range(n - 1):
            CX | c([q, q + 1])
    return c


def zz_x_hamiltonian(n: int, x_coef: float = 0.5):
    """H = sum_i Z_i Z_{i+1} + x_coef * sum_i X_i as [(coef, pauli_string)], qubit 0 first."""
    terms = []
    for i in range(n - 1):
        s = ["I"] * n
        s[i] = s[i + 1] = "Z"
        terms.append((1.0, "".join(s)))
    for i in range(n):
        s = ["I"] * n
        s[i] = "X"
        terms.append((x_coef, "".join(s)))
    return terms
