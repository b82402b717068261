"""OpenQASM 2 writer/reader round trip (CPU only).

Circuit.qasm() followed by qasm.loads() must give back the same gate list
(type, control and target qubits, parameters bit for bit) and therefore the
same state; checked on random circuits over every gate type with a QASM name
(the reference's name table, QuICT/tools/interface/qasm_interface.py:42-85),
with measure / reset / barrier, and on the reference's golden fixture.
"""
import os
import random

import numpy as np
import pytest

from quict_amd.core import Circuit, GateType, Measure, Reset, Barrier, H
from quict_amd.core import qasm
from oracle import statevector as osv
from test_oracle import ALL_TYPES


def same_gates(a: Circuit, b: Circuit):
    assert a.width() == b.width()
    assert len(a.gates) == len(b.gates)
    for g, h in zip(a.gates, b.gates):
        assert g.type == h.type
        assert list(g.cargs) == list(h.cargs) and list(g.targs) == list(h.targs)
        assert list(g.pargs) == list(h.pargs)


@pytest.mark.parametrize("seed", range(4))
def test_random_circuits_round_trip(seed):
    n = 7
    c = Circuit(n)
    random.seed(seed)
    np.random.seed(seed)
    c.random_append(120, typelist=ALL_TYPES, random_params=True)
    Barrier | c(2)
    Reset | c(1)
    H | c(1)
    Measure | c(4)
    text = c.qasm()
    assert text.startswith("OPENQASM 2.0;")
    back = qasm.loads(text)
    same_gates(c, back)
    # identical simulation (measurement draws reseeded identically)
    random.seed(11)
    a = osv.run(c)
    random.seed(11)
    b = osv.run(back)
    assert np.array_equal(a, b)


def test_golden_fixture_round_trip(golden_dir):
    c = qasm.load_file(os.path.join(golden_dir, "random_circuit_for_correction.qasm"))
    back = qasm.loads(c.qasm())
    same_gates(c, back)
    sv = np.load(os.path.join(golden_dir, "state_vector.npy"))
    assert np.max(np.abs(osv.run(back) - sv)) < 1e-12


def test_rejected_forms():
    from quict_amd.core import exception as qerr
    with pytest.raises(qerr.QASMError):
        qasm.loads('OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[2];\ngate foo a { h a; }\nfoo q[0];\n')
    with pytest.raises(Exception):
        qasm.loads('OPENQASM 2.0;\nqreg q[2];\nrx(import_os) q[0];\n')
