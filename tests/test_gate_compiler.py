"""Gate records (CPU only): the run compiler's shortcuts (one translated
record per fixed gate type, rotation payloads computed for a whole run at
once) must produce exactly the records of the gate-by-gate translator, in both
precisions and at degenerate angles where a rotation becomes diagonal or the
identity."""
import random

import numpy as np
import pytest

from quict_amd import _native as nat
from quict_amd.core import Circuit, Rx, Ry, Rz, CRz, U1, U3, H, CX
from quict_amd.simulation import gate_compiler as gc
from test_oracle import ALL_TYPES


def one_by_one(gates, n):
    arr = np.zeros(len(gates), dtype=nat.QS_GATE_DTYPE)
    keep, k = [], 0
    for g in gates:
        if gc.compile_gate(g, n, arr[k], keep):
            k += 1
    return arr[:k]


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("single", [False, True])
def test_run_compiler_equals_gate_by_gate(seed, single):
    n = 8
    c = Circuit(n)
    random.seed(seed)
    np.random.seed(seed)
    c.random_append(300, typelist=ALL_TYPES, random_params=True)
    if single:
        c.convert_precision()
    recs, _ = gc.compile_gates(c.gates, n)
    assert np.array_equal(recs, one_by_one(c.gates, n))


def test_degenerate_angles():
    n = 4
    c = Circuit(n)
    for th in (0.0, 2 * np.pi, 4 * np.pi, -4 * np.pi, np.pi, 1e-300):
        Rx(th) | c(0)
        Ry(th) | c(1)
        Rz(th) | c(2)
        CRz(th) | c([3, 2])
        U1(th) | c(1)
        U3(th, 0.3, -0.2) | c(0)
        U3(0.0, th, 0.5) | c(3)
        H | c(2)
        CX | c([0, 1])
    recs, _ = gc.compile_gates(c.gates, n)
    assert np.array_equal(recs, one_by_one(c.gates, n))
