"""The C-ABI library loads and exports every symbol include/qsim.h declares;
the ctypes record layouts match the C compiler's (CPU only, no compute)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from conftest import ROOT
from quict_amd import _native as nat

HEADER = os.path.join(ROOT, "include", "qsim.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qs_[a-z0-9_]+)\s*\(", text)))


def test_library_built():
    assert os.path.exists(nat.LIB_PATH), "build libqsim_hip.so first (__graft_entry__.build())"


def test_every_declared_symbol_is_exported_and_bound():
    lib = nat.load_library()
    decl = declared_functions()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(lib, name), f"{name} declared in qsim.h but not exported"
        assert name in nat.SIGNATURES, f"{name} has no ctypes signature"
    # and nothing bound that the header does not declare
    assert sorted(nat.SIGNATURES) == decl


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "qsim.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(qs_gate), offsetof(qs_gate, bits),"
        " offsetof(qs_gate, m), offsetof(qs_gate, big), sizeof(qs_pauli_term), offsetof(qs_pauli_term, zmask),"
        " sizeof(qs_qnn_op), offsetof(qs_qnn_op, value), sizeof(qs_qnn_config));"
        "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    vals = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    assert vals == [ctypes.sizeof(nat.QsGate), nat.QsGate.bits.offset, nat.QsGate.m.offset, nat.QsGate.big.offset,
                    ctypes.sizeof(nat.QsPauliTerm), nat.QsPauliTerm.zmask.offset, ctypes.sizeof(nat.QsQnnOp),
                    nat.QsQnnOp.value.offset, ctypes.sizeof(nat.QsQnnConfig)]
    assert nat.QS_GATE_DTYPE.itemsize == ctypes.sizeof(nat.QsGate)


def test_no_device_means_loud_failure():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from quict_amd.core import exception as qerr
    from quict_amd.simulation import StateVectorSimulator
    with pytest.raises(qerr.NativeRuntimeError):
        StateVectorSimulator()


def test_error_translation_on_bad_arguments():
    lib = nat.load_library()
    # null context/pointer arguments are rejected in C before any device call
    rc = lib.qs_sv_init(None, None, 4, nat.QS_C128, 0, None)
    assert rc == nat.QS_ERR_ARG
    assert b"null" in lib.qs_last_error()
    rc = lib.qs_set_fusion(None, 12, 4)
    assert rc == nat.QS_ERR_ARG


def test_qnn_program_records():
    """QNN programs: gate lists become qs_qnn_op records (encodings first,
    trainable / fixed rotations, CX control then target, CZ); other gates are
    refused; the record layout matches include/qsim.h."""
    import ctypes
    from quict_amd import _native as nat
    from quict_amd.core import GateType
    from quict_amd.core import exception as qerr
    from quict_amd.gradient.adjoint import ParameterizedCircuit
    from quict_amd.gradient.qnn import program_ops
    assert ctypes.sizeof(nat.QsQnnOp) == 24 and nat.QsQnnOp.value.offset == 16
    assert ctypes.sizeof(nat.QsQnnConfig) == 20
    pc = ParameterizedCircuit(3)
    pc.add(GateType.rx, [0], param=2).add(GateType.cx, [2, 1]).add(GateType.rz, [1], value=0.25)
    pc.add(GateType.cz, [0, 2]).add(GateType.ry, [2], param=0)
    ops = program_ops(pc, [(1, 4)])
    assert ops == [(nat.QS_QNN_ENC_RY, 1, 0, 4, 0.0), (nat.QS_QNN_RX, 0, 0, 2, 0.0), (nat.QS_QNN_CX, 2, 1, -1, 0.0),
                   (nat.QS_QNN_RZ, 1, 0, -1, 0.25), (nat.QS_QNN_CZ, 0, 2, -1, 0.0), (nat.QS_QNN_RY, 2, 0, 0, 0.0)]
    bad = ParameterizedCircuit(2).add(GateType.h, [0])
    with pytest.raises(qerr.GateAlgorithmNotImplementError):
        program_ops(bad)
