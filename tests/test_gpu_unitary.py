"""Circuit matrices and the unitary simulator on the GPU against the oracle.

Bars: complex128 max-abs <= 1e-12 against the restated CircuitMatrix (small
registers) and the reference's golden 10-qubit state; complex64 within the
error printed per case (<= 2e-5 here); the matrix-vector product, transpose
and identity through the C ABI against numpy at every size class (tiles
smaller than, equal to and larger than the 32 x 32 transpose tile).  At 13
qubits, size-independent checks: rows of the column layout equal the
state-vector simulator run from basis states, U x equals the state-vector
simulator run from x.
"""
import ctypes
import os
import random

import numpy as np
import pytest

from quict_amd.core import Circuit, GateType, H, CX, Measure
from quict_amd.core import qasm
from quict_amd import _native as nat
from quict_amd import workloads
from oracle import unitary as ou

pytestmark = pytest.mark.gpu

ONE_Q = [GateType.h, GateType.x, GateType.y, GateType.z, GateType.s, GateType.t, GateType.u3, GateType.rx,
         GateType.ry, GateType.rz]
TWO_Q = [GateType.cx, GateType.cz, GateType.crz, GateType.ch, GateType.cu3, GateType.fsim, GateType.rxx,
         GateType.swap]


def say(*a):
    print("[case]", *a, flush=True)


def random_circuit(n, size, seed):
    c = Circuit(n)
    random.seed(seed)
    np.random.seed(seed)
    c.random_append(size, typelist=ONE_Q + (TWO_Q if n >= 2 else []), random_params=True)
    return c


def dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a).ravel(), device="cuda:0")


def lib_ctx():
    return nat.load_library(), nat.context(0), nat.stream_ptr(0)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [1, 2, 4, 5, 6, 9, 11])
def test_mat_abi_identity_transpose_matvec(n, dtype):
    say("mat abi", n, dtype.__name__)
    import torch
    lib, ctx, s = lib_ctx()
    code = nat.QS_C128 if dtype == np.complex128 else nat.QS_C64
    D = 1 << n
    rng = np.random.RandomState(n)
    I = torch.empty(D * D, dtype=torch.complex128 if code == nat.QS_C128 else torch.complex64, device="cuda:0")
    nat.check(lib.qs_mat_identity(ctx, ctypes.c_void_p(I.data_ptr()), n, code, s))
    assert np.array_equal(I.cpu().numpy().reshape(D, D), np.identity(D, dtype=dtype))

    M = (rng.randn(D, D) + 1j * rng.randn(D, D)).astype(dtype)
    x = (rng.randn(D) + 1j * rng.randn(D)).astype(dtype)
    Md, xd = dev(M), dev(x)
    y = torch.empty_like(xd)
    tol = 1e-12 * D if dtype == np.complex128 else 2e-5 * D
    for trans, ref in ((0, M.astype(np.complex128) @ x), (1, M.T.astype(np.complex128) @ x)):
        nat.check(lib.qs_mat_vec(ctx, ctypes.c_void_p(Md.data_ptr()), n, code, trans, ctypes.c_void_p(xd.data_ptr()),
                                 ctypes.c_void_p(y.data_ptr()), s))
        err = np.max(np.abs(y.cpu().numpy() - ref))
        print(f"  matvec trans={trans} max err {err:.2e}")
        assert err < tol
    nat.check(lib.qs_mat_transpose(ctx, ctypes.c_void_p(Md.data_ptr()), n, code, s))
    assert np.array_equal(Md.cpu().numpy().reshape(D, D), M.T)
    # out of place only
    rc = lib.qs_mat_vec(ctx, ctypes.c_void_p(Md.data_ptr()), n, code, 0, ctypes.c_void_p(xd.data_ptr()),
                        ctypes.c_void_p(xd.data_ptr()), s)
    assert rc == nat.QS_ERR_ARG


@pytest.mark.parametrize("fusion", [False, True])
@pytest.mark.parametrize("n,size,seed", [(1, 8, 0), (2, 20, 1), (3, 40, 2), (5, 80, 3), (7, 120, 4), (8, 200, 5)])
def test_circuit_matrix_vs_restated_circuit_matrix(n, size, seed, fusion):
    say("circuit matrix", n, size, seed, fusion)
    from quict_amd.simulation import circuit_matrix
    c = random_circuit(n, size, seed)
    ref = ou.circuit_matrix(c.gates, n)
    U = circuit_matrix(c, "double", fusion=fusion).get()
    assert np.max(np.abs(U - ref)) < 1e-12
    U64 = circuit_matrix(c, "single", fusion=fusion).get()
    err = np.max(np.abs(U64 - ref))
    print(f"  complex64 max err {err:.2e}")
    assert err < 2e-5


def test_circuit_matrix_method_and_three_qubit_gates(golden_dir):
    say("Circuit.matrix GPU golden")
    c = qasm.load_file(os.path.join(golden_dir, "random_circuit_for_correction.qasm"))
    n = c.width()
    U = c.matrix(device="GPU").get()
    B = ou.bruteforce_matrix(c.gates, n)
    assert np.max(np.abs(U - B)) < 1e-12


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("single", 1e-6)])
def test_unitary_simulator_golden(golden_dir, precision, tol):
    say("unitary sim golden", precision)
    from quict_amd.simulation import UnitarySimulator, Simulator
    c = qasm.load_file(os.path.join(golden_dir, "random_circuit_for_correction.qasm"))
    sv = np.load(os.path.join(golden_dir, "state_vector.npy"))
    out = UnitarySimulator("GPU", precision=precision).run(c).get()
    assert np.max(np.abs(out - sv)) < tol
    res = Simulator(device="GPU", backend="unitary", precision=precision).run(
        qasm.load_file(os.path.join(golden_dir, "random_circuit_for_correction.qasm")))
    assert np.max(np.abs(res["data"]["state_vector"] - sv)) < tol
    assert res["shots"] == 1


def test_unitary_simulator_matrix_input_state_and_previous():
    say("unitary sim matrix input")
    from quict_amd.simulation import UnitarySimulator
    n = 6
    c = random_circuit(n, 60, 7)
    U = ou.circuit_matrix(c.gates, n)
    rng = np.random.RandomState(8)
    x = rng.randn(1 << n) + 1j * rng.randn(1 << n)
    x /= np.linalg.norm(x)
    sim = UnitarySimulator()
    out = sim.run(U, state_vector=x).get()
    assert np.max(np.abs(out - U @ x)) < 1e-12
    # use_previous: apply the circuit to the previous result
    out2 = sim.run(c, use_previous=True).get()
    assert np.max(np.abs(out2 - U @ (U @ x))) < 1e-12
    assert np.max(np.abs(sim.unitary_matrix.get() - U)) < 1e-12
    # identity input returns the state bit for bit
    out3 = sim.run(np.identity(1 << n), state_vector=x).get()
    assert np.array_equal(out3, x)
    with pytest.raises(Exception):
        sim.run(np.ones((3, 3)))


def test_unitary_sample_ghz_and_seeded_like_statevector():
    say("unitary sample")
    from quict_amd.simulation import UnitarySimulator, StateVectorSimulator
    n = 4
    c = Circuit(n)
    H | c(0)
    for i in range(n - 1):
        CX | c([i, i + 1])
    sim = UnitarySimulator()
    sim.run(c)
    b = sim.sample(100)
    assert b[0] + b[-1] == 100
    c2 = random_circuit(8, 80, 9)
    usim, ssim = UnitarySimulator(), StateVectorSimulator()
    usim.run(c2)
    ssim.run(random_circuit(8, 80, 9))
    random.seed(1234)
    cu = usim.sample(500)
    random.seed(1234)
    cs = ssim.sample(500)
    assert cu == cs


def test_special_gates_are_skipped_in_the_matrix():
    say("special gates skipped")
    from quict_amd.simulation import circuit_matrix
    c = random_circuit(4, 30, 10)
    Measure | c(1)
    H | c(1)
    ref = ou.circuit_matrix(c.gates, 4)
    assert np.max(np.abs(circuit_matrix(c).get() - ref)) < 1e-12


def test_13q_columns_and_product_match_the_statevector_simulator():
    """Size-independent checks at 13 qubits (a 1 GiB complex128 unitary)."""
    say("unitary 13q")
    import torch
    from quict_amd.simulation import StateVectorSimulator, UnitarySimulator
    from quict_amd.simulation.unitary import build_columns
    n = 13
    c = workloads.random_mix_circuit(n, 300, 5)
    W, passes = build_columns(c.gates, n, nat.QS_C128)
    Wm = W.view(1 << n, 1 << n)
    sv = StateVectorSimulator()
    for col in (0, 1, 4095, (1 << n) - 1):
        e = np.zeros(1 << n, dtype=np.complex128)
        e[col] = 1
        ref = sv.run(workloads.random_mix_circuit(n, 300, 5), state_vector=e).get()
        assert np.max(np.abs(Wm[col].cpu().numpy() - ref)) < 1e-12
    rng = np.random.RandomState(11)
    x = rng.randn(1 << n) + 1j * rng.randn(1 << n)
    x /= np.linalg.norm(x)
    usim = UnitarySimulator()
    out = usim.run(workloads.random_mix_circuit(n, 300, 5), state_vector=x).get()
    ref = sv.run(workloads.random_mix_circuit(n, 300, 5), state_vector=x).get()
    assert np.max(np.abs(out - ref)) < 1e-12
    print(f"  {passes} fused passes for 300 gates")
    del W, Wm
    torch.cuda.empty_cache()
