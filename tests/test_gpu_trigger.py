"""Dynamic circuits: Trigger (mid-circuit measurement that splices the gates
mapped to the outcome, statevector_simulator.py:498-523) on the GPU against
the oracle, with the same seeded measurement draws."""
import random

import numpy as np
import pytest

from quict_amd.core import Circuit, CompositeGate, H, X, CX, Ry, Trigger, Measure
from oracle import statevector as osv

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def circuit(kind: int) -> Circuit:
    n = 6
    c = Circuit(n)
    for q in range(3):
        H | c(q)
    Ry(0.7) | c(3)
    if kind == 0:      # dict mapping, one measured qubit
        with CompositeGate() as g0:
            X & 4
        with CompositeGate() as g1:
            H & 4
            CX & [4, 5]
        Trigger(1, {0: g0, 1: g1}) | c([0])
    elif kind == 1:    # list mapping over two measured qubits, one empty branch
        maps = []
        for s in range(4):
            if s == 2:
                maps.append(None)
                continue
            with CompositeGate() as g:
                Ry(0.3 * (s + 1)) & 5
                CX & [5, 3]
            maps.append(g)
        Trigger(2, maps) | c([1, 2])
    else:              # function mapping, then more gates and a measurement
        def f(state):
            with CompositeGate() as g:
                for q in range(state + 1):
                    X & (3 + q % 3)
            return g
        Trigger(2, f) | c([0, 2])
        CX | c([3, 1])
        Measure | c(5)
    H | c(4)
    return c


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_trigger_matches_oracle(kind, seed):
    say("trigger", kind, seed)
    from quict_amd.simulation import StateVectorSimulator
    c = circuit(kind)
    random.seed(seed)
    out = StateVectorSimulator().run(c).get()
    bits_gpu = [int(q.measured) if q.measured is not None else -1 for q in c.qubits]
    c2 = circuit(kind)
    random.seed(seed)
    ref = osv.run(c2)
    bits_ref = [int(q.measured) if q.measured is not None else -1 for q in c2.qubits]
    assert bits_gpu == bits_ref
    assert np.max(np.abs(out - ref)) < 1e-12


def test_trigger_with_checkpoint_is_refused():
    say("trigger checkpoint")
    from quict_amd.simulation import StateVectorSimulator
    from quict_amd.core import exception as qerr
    c = Circuit(3)
    H | c(0)
    with CompositeGate() as g:
        X & 1
    g.checkpoint = object()
    Trigger(1, [g, g]) | c([0])
    with pytest.raises(qerr.GateTypeNotImplementError):
        StateVectorSimulator().run(c)
