"""The reference's own call shapes through ``quict_amd.ops.LinAlgLoader``.

``dispatch`` below restates the GPU branch of
QuICT/simulation/state_vector/statevector_simulator.py:202-315 (apply_gate and
apply_normal / diagonal / swap / reverse / control_matrix): per gate it picks
a loader function by MatrixType and arity and calls it with reversed qubit
indices, the gate matrix and ``(vec, n, sync)``.  Replaying circuits gate by
gate through it must give the reference's golden 10-qubit vector and the
oracle's amplitudes."""
import os
import random

import numpy as np
import pytest

from quict_amd import workloads
from quict_amd.core import Circuit, GateType, MatrixType, qasm
from oracle import statevector as osv

import reference_fixtures as rf

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def dispatch(alg, gate, vec, n, sync=False):
    """statevector_simulator.py:222-315 (GPU branch), gate by gate."""
    mt = gate.matrix_type
    gtype = gate.type
    dp = (vec, n, sync)
    if gtype in (GateType.id, GateType.barrier) or (
            gtype not in (GateType.measure, GateType.reset, GateType.perm) and gate.is_identity()):
        return vec
    if gtype in (GateType.unitary, GateType.qft, GateType.iqft, GateType.perm_fx) and gate.targets >= 3:
        k = gate.controls + gate.targets
        return alg.matrix_dot_vector(vec, n, np.asarray(gate.matrix).reshape(1 << k, 1 << k),
                                     list(gate.cargs) + list(gate.targs), sync)
    args = list(gate.cargs) + list(gate.targs)
    special = gtype in (GateType.measure, GateType.reset, GateType.perm)
    mat = None if special else np.asarray(gate.matrix)
    if mt == MatrixType.normal:
        if len(args) == 1:
            alg.normal_targ(n - 1 - args[0], mat, *dp)
        elif len(args) == 2 and gate.controls == gate.targets:
            alg.normal_ctargs(n - 1 - gate.cargs[0], n - 1 - gate.targs[0], mat, *dp)
        elif len(args) == 2 and gate.targets == 2:
            alg.normal_targs([n - 1 - a for a in args], mat, *dp)
        # three or more arguments: no branch (statevector_simulator.py:317-350), e.g. CCZ
    elif mt in (MatrixType.diagonal, MatrixType.diag_diag):
        if len(args) == 1:
            alg.diagonal_targ(n - 1 - gate.targs[0], mat, *dp)
        elif len(args) == 2 and gate.controls == gate.targets:
            alg.diagonal_ctargs(n - 1 - gate.cargs[0], n - 1 - gate.targs[0], mat, *dp)
        elif len(args) == 2:
            alg.diagonal_targs([n - 1 - a for a in args], mat, *dp)
        else:
            assert gtype == GateType.ccrz
            alg.diagonal_more([n - 1 - c for c in gate.cargs], n - 1 - gate.targs[0], mat, *dp)
    elif mt == MatrixType.swap:
        if len(args) == 1:
            alg.swap_targ(n - 1 - gate.targs[0], *dp)
        elif len(args) == 2:
            alg.swap_targs([n - 1 - a for a in args], mat, *dp)
        else:
            alg.swap_tmore([n - 1 - t for t in gate.targs], n - 1 - gate.cargs[0], *dp)
    elif mt == MatrixType.reverse:
        if len(args) == 1:
            alg.reverse_targ(n - 1 - args[0], mat, *dp)
        elif len(args) == 2:
            alg.reverse_ctargs(n - 1 - args[0], n - 1 - args[1], mat, *dp)
        else:
            alg.reverse_more([n - 1 - c for c in gate.cargs], n - 1 - gate.targs[0], *dp)
    elif mt == MatrixType.control:
        if len(args) == 1:
            alg.control_targ(n - 1 - args[0], mat[1, 1], *dp)
        else:
            alg.control_ctargs(n - 1 - args[0], n - 1 - args[1], mat[3, 3], *dp)
    elif mt == MatrixType.ctrl_normal:
        alg.ctrl_normal_targs([n - 1 - t for t in gate.targs], mat, *dp)
    elif mt == MatrixType.normal_normal:
        alg.normal_normal_targs([n - 1 - t for t in gate.targs], mat, *dp)
    elif mt == MatrixType.diag_normal:
        alg.diagonal_normal_targs([n - 1 - t for t in gate.targs], mat, *dp)
    elif gtype in (GateType.measure, GateType.reset):
        index = n - 1 - gate.targs[0]
        prob = alg.measured_prob_calculate(index, vec, n).get()
        if gtype == GateType.measure:
            alg.apply_measuregate(index, vec, n, prob, sync)
        else:
            alg.apply_resetgate(index, vec, n, prob, sync)
    elif gtype == GateType.perm:
        if len(args) == n:
            mapping = np.array(gate.pargs, dtype=np.int32)
        else:
            mapping = np.arange(n, dtype=np.int32)
            for idx, parg in enumerate(gate.pargs):
                mapping[args[idx]] = args[parg]
        alg.VectorPermutation(vec, mapping, changeInput=True, gpu_out=False, sync=sync)
    else:
        raise NotImplementedError(f"{gtype} {mt}")
    return vec


def run_through_loader(c: Circuit, precision="double", psi0=None):
    import torch
    from quict_amd.ops import LinAlgLoader
    alg = LinAlgLoader(device="GPU", enable_gate_kernel=True)
    n = c.width()
    dt = torch.complex128 if precision == "double" else torch.complex64
    if psi0 is None:
        vec = torch.zeros(1 << n, dtype=dt, device="cuda:0")
        vec[0] = 1
    else:
        vec = torch.as_tensor(psi0.astype(np.complex128 if precision == "double" else np.complex64),
                              device="cuda:0")
    for g in c.gates:
        vec = dispatch(alg, g, vec, n)
    return vec.cpu().numpy()


def test_loader_binds_the_reference_names():
    from quict_amd.ops import LinAlgLoader
    from quict_amd.ops.linalg_loader import GATE_KERNEL_FUNCTIONS, BASED_LINALG
    alg = LinAlgLoader(device="GPU", enable_gate_kernel=True)
    assert len(GATE_KERNEL_FUNCTIONS) == 21
    for name in GATE_KERNEL_FUNCTIONS + BASED_LINALG:
        assert callable(getattr(alg, name)), name
    assert not hasattr(LinAlgLoader(device="GPU"), "normal_targ")  # gate kernels only when enabled


@pytest.mark.parametrize("precision", ["double", "single"])
def test_golden_qasm_through_the_reference_dispatch(golden_dir, precision):
    say("golden QASM through LinAlgLoader", precision)
    c = qasm.load_file(os.path.join(golden_dir, "random_circuit_for_correction.qasm"))
    ref = np.load(os.path.join(golden_dir, "state_vector.npy"))
    out = run_through_loader(c, precision)
    err = float(np.max(np.abs(out - ref)))
    print(f"  max-abs error vs the golden vector: {err:.2e}")
    assert err < (1e-12 if precision == "double" else 1e-5)


# (the all-types fixtures hold CCZ, which this dispatch table drops: reference defect q2)
@pytest.mark.parametrize("name", ["mix12_6040", "mix13_6040", "rz_heavy9"])
def test_reference_fixtures_through_the_reference_dispatch(name):
    say("reference fixture through LinAlgLoader", name)
    meta, arrays = rf.sv_cases()
    m = meta[name]
    c = rf.circuit_from(m["n"], m["gates"])
    out = run_through_loader(c, psi0=arrays.get(name + "__psi0"))
    assert np.max(np.abs(out - arrays[name])) < 1e-12


def test_config_mix_with_measure_reset_perm_and_wide_unitaries():
    """12 qubits: the 60/40 mix, a 3-qubit and a 4-qubit Unitary (matrix_dot_vector),
    a Perm (VectorPermutation), mid-circuit measure / reset with seeded draws;
    against the oracle."""
    say("config mix + unitary + perm + measure through LinAlgLoader")
    from quict_amd.core import Unitary, Perm, Measure, Reset
    n = 12

    def build():
        c = workloads.random_mix_circuit(n, 120, 51, workloads.RANDOM_GATE_PROBS_60_40)
        for k, args in [(3, [4, 0, 9]), (4, [11, 2, 7, 5])]:
            q, _ = np.linalg.qr(np.random.RandomState(k).randn(1 << k, 1 << k)
                                + 1j * np.random.RandomState(k + 10).randn(1 << k, 1 << k))
            Unitary(q) | c(args)
        Perm(4, [2, 0, 3, 1]) | c([1, 6, 3, 10])
        Measure | c(3)
        Reset | c(8)
        for g in workloads.random_mix_circuit(n, 60, 52).gates:
            c.append(g)
        return c

    random.seed(11)
    out = run_through_loader(build())
    random.seed(11)
    ref = osv.run(build())
    assert np.max(np.abs(out - ref)) < 1e-12


def _gpu_table_rewrite(c: Circuit) -> Circuit:
    """What the reference's GPU dispatch table computes, as a circuit the oracle
    runs: q1 — normal_targs / diagonal_normal_targs take the first reversed
    index as the matrix LSB, so a 2-target Unitary or Rzx acts with its
    arguments reversed; q2 — CCZ (MatrixType.normal on 3 arguments) matches no
    branch and is dropped."""
    from quict_amd.core import build_gate, Unitary
    out = Circuit(c.width())
    for g in c.gates:
        args = list(g.cargs) + list(g.targs)
        if g.type == GateType.ccz:
            continue
        if g.type == GateType.rzx:
            out.append(build_gate(GateType.rzx, args[::-1], list(g.pargs)))
        elif g.type == GateType.unitary and len(args) == 2:
            Unitary(np.asarray(g.matrix)) | out(args[::-1])
        else:
            out.append(g)
    return out


def test_every_gate_kernel_function_through_the_reference_dispatch():
    """A 10-qubit circuit that reaches all 18 gate functions of the table
    (incl. ctrl_normal_targs / FSim, normal_normal_targs / Rxx, Ryy,
    diagonal_normal_targs / Rzx, normal_targs / 2-target Unitary,
    diagonal_more / CCRz, reverse_more / CCX, swap_tmore / CSwap) from a random
    state, against the oracle on the circuit the reference table computes."""
    say("every LinAlgLoader gate function")
    from quict_amd.core import Unitary, build_gate
    from test_oracle import ALL_TYPES
    n = 10
    c = Circuit(n)
    random.seed(71)
    np.random.seed(71)
    c.random_append(150, typelist=ALL_TYPES, random_params=True)
    q, _ = np.linalg.qr(np.random.RandomState(3).randn(4, 4) + 1j * np.random.RandomState(4).randn(4, 4))
    Unitary(q) | c([7, 2])
    used = {g.type for g in c.gates}
    assert {GateType.fsim, GateType.rxx, GateType.rzx, GateType.ccrz, GateType.ccx, GateType.cswap} <= used
    rng = np.random.RandomState(72)
    psi0 = rng.randn(1 << n) + 1j * rng.randn(1 << n)
    psi0 /= np.linalg.norm(psi0)
    out = run_through_loader(c, psi0=psi0)
    ref = osv.run(_gpu_table_rewrite(c), state_vector=psi0.copy())
    assert np.max(np.abs(out - ref)) < 1e-12
