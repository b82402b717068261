"""Fusion planner on the CPU.

The planner (quict_amd/csrc/qsim_fused_plan.hpp) turns a gate list into pass
programs: tile bits, register/thread layouts, relabelings (SWAP), flips (X, Y),
LDS relayouts with XOR-swizzled slot offsets, merged ops and coefficient
tables.  tests/planner/emu.cpp runs those exact programs with the semantics of
the device interpreter (per tile, per lane and register, LDS image with range
and collision checks).  Its results must match the C restatement of the
reference kernels (oracle/csrc/cpu_ref_sim.c, itself pinned against the numpy
oracle and the reference's golden vector) to 1e-12, starting from a random
state so that every amplitude position matters.  The GPU tests run the same
programs on the device.
"""
import ctypes
import os
import random
import subprocess

import numpy as np
import pytest

from quict_amd import _native as nat
from quict_amd import workloads
from quict_amd.core import (Circuit, H, Unitary, QFT, IQFT, CCX, CSwap, Rzx, FSim, X, Y, Swap, CX, CY)
from quict_amd.simulation.gate_compiler import compile_gates
from oracle import statevector as osv
from test_oracle import ALL_TYPES

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "planner", "_build", "libplanner_emu.so")
C128 = (4, 8, 16)  # register bits, thread bits, LDS bytes per amplitude
C64 = (4, 8, 8)
C64_K13 = (4, 9, 8)


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "planner")], check=True)
    lib = ctypes.CDLL(LIB)
    lib.emu_apply.restype = ctypes.c_int
    lib.emu_apply.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.cpu_apply_gates.restype = ctypes.c_int
    lib.emu_last_error.restype = ctypes.c_char_p
    return lib


def run_both(lib, c, n, shape, L=4, seed=0):
    recs, keep = compile_gates(c.gates, n)
    rng = np.random.RandomState(seed)
    v = rng.randn(1 << n) + 1j * rng.randn(1 << n)
    v /= np.linalg.norm(v)
    ref = v.copy()
    lib.cpu_apply_gates(ref.ctypes.data_as(ctypes.c_void_p), n, nat.gate_array_ptr(recs), len(recs), 1)
    out = v.copy()
    st = np.zeros(18, dtype=np.int32)
    RB, TB, S = shape
    rc = lib.emu_apply(out.ctypes.data_as(ctypes.c_void_p), n, nat.gate_array_ptr(recs), len(recs), RB, TB, L, S,
                       st.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, lib.emu_last_error()
    return out, ref, st


def all_types(n, ngates, seed):
    c = Circuit(n)
    random.seed(seed)
    np.random.seed(seed)
    c.random_append(ngates, typelist=ALL_TYPES, random_params=True)
    return c


def flips_and_swaps(n, ngates, seed):
    rng = random.Random(seed)
    np.random.seed(seed)
    c = Circuit(n)
    for _ in range(ngates):
        k = rng.random()
        if k < 0.25:
            X | c(rng.randrange(n))
        elif k < 0.4:
            Y | c(rng.randrange(n))
        elif k < 0.6:
            Swap | c(rng.sample(range(n), 2))
        elif k < 0.7:
            CX | c(rng.sample(range(n), 2))
        elif k < 0.75:
            CY | c(rng.sample(range(n), 2))
        elif k < 0.8:
            CSwap | c(rng.sample(range(n), 3))
        else:
            c.random_append(1, typelist=ALL_TYPES, random_params=True)
    return c


def dense_multi(n=16):
    rng = np.random.RandomState(4)
    c = Circuit(n)
    for q in range(n):
        H | c(q)
    for k, args in [(3, [5, 1, 12]), (4, [0, 15, 7, 3]), (5, [2, 9, 14, 6, 11]), (3, [13, 14, 15])]:
        q, _ = np.linalg.qr(rng.randn(1 << k, 1 << k) + 1j * rng.randn(1 << k, 1 << k))
        Unitary(q) | c(args)
        Rzx(0.3 * k) | c([args[0], args[1]])
    QFT(4) | c([6, 2, 0, 4])
    IQFT(3) | c([8, 10, 1])
    CCX | c([3, 7, 11])
    CSwap | c([0, 12, 4])
    FSim(0.4, 0.9) | c([9, 2])
    return c


@pytest.mark.parametrize("shape", [C128, C64, C64_K13], ids=["c128-tile", "c64-tile", "c64-tile-k13"])
@pytest.mark.parametrize("n,ngates,seed", [(14, 200, 1), (15, 300, 2), (16, 400, 3)])
def test_all_gate_types(emu, shape, n, ngates, seed):
    out, ref, st = run_both(emu, all_types(n, ngates, seed), n, shape)
    assert st[1] > 0  # fused passes were planned
    assert np.max(np.abs(out - ref)) < 1e-12


@pytest.mark.parametrize("shape", [C128, C64], ids=["c128-tile", "c64-tile"])
@pytest.mark.parametrize("n,ngates,seed", [(14, 300, 5), (16, 500, 6)])
def test_relabelings_and_flips(emu, shape, n, ngates, seed):
    out, ref, st = run_both(emu, flips_and_swaps(n, ngates, seed), n, shape)
    assert np.max(np.abs(out - ref)) < 1e-12


@pytest.mark.parametrize("shape", [C128, C64], ids=["c128-tile", "c64-tile"])
def test_multi_target_dense(emu, shape):
    out, ref, st = run_both(emu, dense_multi(), 16, shape)
    assert st[3 + 6] + st[3 + 7] > 0  # 3- and 4-target in-register ops were used
    assert np.max(np.abs(out - ref)) < 1e-12


@pytest.mark.parametrize("L", [2, 5, 6])
def test_low_bit_counts(emu, L):
    c = workloads.random_mix_circuit(16, 300, 21, workloads.RANDOM_GATE_PROBS_60_40)
    out, ref, st = run_both(emu, c, 16, C128, L=L)
    assert np.max(np.abs(out - ref)) < 1e-12


@pytest.mark.parametrize("lr", [1, 2])
def test_low_bits_in_registers_at_load_and_store(emu, lr, monkeypatch):
    """Planner variant that leaves up to `lr` low bits in registers at the load
    and store (fewer relayouts, slower tile accesses; not the default)."""
    monkeypatch.setenv("QS_FUSION_LOWREGS", str(lr))
    c = workloads.random_mix_circuit(16, 300, 21, workloads.RANDOM_GATE_PROBS_60_40)
    out, ref, st = run_both(emu, c, 16, C128)
    assert np.max(np.abs(out - ref)) < 1e-12


def test_low_bits_up_to_the_thread_bits(emu):
    """An 11-bit tile has 7 thread slots: 7 low bits plan and run, 8 are refused
    (the planner keeps the low bits on thread slots at the load and store)."""
    c = workloads.random_mix_circuit(16, 300, 21, workloads.RANDOM_GATE_PROBS_60_40)
    out, ref, st = run_both(emu, c, 16, (4, 7, 16), L=7)
    assert np.max(np.abs(out - ref)) < 1e-12
    recs, _ = compile_gates(c.gates, 16)
    v = np.zeros(1 << 16, dtype=np.complex128)
    st = np.zeros(18, dtype=np.int32)
    rc = emu.emu_apply(v.ctypes.data_as(ctypes.c_void_p), 16, nat.gate_array_ptr(recs), len(recs), 4, 7, 8, 16,
                       st.ctypes.data_as(ctypes.c_void_p))
    assert rc != 0 and b"low_bits" in emu.emu_last_error()


def wide_controlled(n=16):
    """A 9-controlled X as a 10-qubit Unitary on qubits 0..9 between two random
    150-gate runs."""
    m = np.eye(1 << 10, dtype=np.complex128)
    m[-2:, -2:] = [[0, 1], [1, 0]]
    c = workloads.random_mix_circuit(n, 150, 41, workloads.RANDOM_GATE_PROBS_60_40)
    Unitary(m) | c(list(range(10)))
    c2 = workloads.random_mix_circuit(n, 150, 42, workloads.RANDOM_GATE_PROBS_60_40)
    for g in c2.gates:
        c.append(g)
    return c


@pytest.mark.parametrize("remap", ["0", "1"])
def test_wide_controlled_gate_on_high_bits(emu, remap, monkeypatch):
    """A 9-controlled X as a 10-qubit Unitary on qubits 0..9 (index bits 15..6):
    more bits above the 4 low bits than a 12-bit tile has room for, so it runs
    as its own per-gate pass between fused passes."""
    monkeypatch.setenv("QS_FUSION_REMAP", remap)
    n = 16
    c = wide_controlled(n)
    out, ref, st = run_both(emu, c, n, C128)
    assert st[0] > st[1]  # one per-gate pass besides the fused ones
    assert np.max(np.abs(out - ref)) < 1e-12


@pytest.mark.parametrize("case", ["all_types", "flips", "wide_controlled", "dense_multi", "hea20"])
def test_streamed_chunks_equal_the_whole_program(emu, case, monkeypatch):
    """qs_sv_apply_gates(fuse=1) uploads and launches the passes in chunks of
    1, 2, 4, ... while it plans the rest: the rebased chunk programs, with the
    per-gate passes in between, give the reference kernels' result and the same
    plan (passes, ops, ops per kind) as the whole program."""
    from quict_amd.gradient.adjoint import ParameterizedCircuit
    monkeypatch.setenv("QS_EMU_FUSE", "1")
    if case == "hea20":  # the 20-qubit 12-layer gradient workload's forward circuit, at 16 qubits
        pc = ParameterizedCircuit.hea(16, 12)
        c = Circuit(16)
        for g, _ in pc.gates(np.random.RandomState(3).uniform(-np.pi, np.pi, pc.nparams)):
            c.append(g)
        n = 16
    else:
        c, n = {"all_types": (all_types(15, 300, 2), 15), "flips": (flips_and_swaps(16, 500, 6), 16),
                "wide_controlled": (wide_controlled(16), 16), "dense_multi": (dense_multi(), 16)}[case]
    out0, ref, st0 = run_both(emu, c, n, C128)
    monkeypatch.setenv("QS_EMU_STREAM", "1")
    out, _, st = run_both(emu, c, n, C128)
    assert np.max(np.abs(out - ref)) < 1e-12
    assert np.array_equal(st[:11], st0[:11]) and st[1] >= 3  # two chunks or more
    assert np.max(np.abs(out - out0)) < 1e-13
    if case == "wide_controlled":
        assert st[0] > st[1]


def test_headline_generator_at_18_qubits(emu):
    """The 2000-gate 60/40 generator of the 30-qubit workload, at 18 qubits."""
    out, ref, st = run_both(emu, workloads.sv30_circuit(n=18), 18, C128)
    assert np.max(np.abs(out - ref)) < 1e-12
    assert st[1] < 120  # about twenty gates per pass


@pytest.mark.parametrize("shape", [C128, C64], ids=["c128-tile", "c64-tile"])
@pytest.mark.parametrize("n,ngates,seed", [(19, 500, 33)])
def test_qubits_moved_onto_low_bits(emu, shape, n, ngates, seed, monkeypatch):
    """Plans whose stores move the qubits needed next onto the low index bits
    (and whose last passes put every qubit back; an option, QS_FUSION_REMAP=1)
    against the reference kernels, and against the plan that moves nothing."""
    c = workloads.random_mix_circuit(n, ngates, seed, workloads.RANDOM_GATE_PROBS_60_40)
    monkeypatch.setenv("QS_FUSION_REMAP", "1")
    out, ref, st = run_both(emu, c, n, shape)
    assert np.max(np.abs(out - ref)) < 1e-12
    assert st[17] > 0  # the moving plan was the shorter one here
    monkeypatch.setenv("QS_FUSION_REMAP", "0")
    out0, _, st0 = run_both(emu, c, n, shape)
    assert st0[17] == 0
    assert st[0] < st0[0]
    assert np.max(np.abs(out - out0)) < 1e-12


@pytest.mark.parametrize("n,ngates,seed,moves", [(20, 900, 33, True), (21, 1000, 36, False)])
def test_fuse2_keeps_moves_only_when_they_save_passes(emu, n, ngates, seed, moves, monkeypatch):
    """The planner mode of qs_sv_apply_gates(fuse=2): the moving plan is kept
    only when it saves at least 8% of the passes (else the fuse=1 plan runs);
    either way the result equals the reference kernels."""
    c = workloads.random_mix_circuit(n, ngates, seed, workloads.RANDOM_GATE_PROBS_60_40)
    recs, keep = compile_gates(c.gates, n)
    monkeypatch.setenv("QS_EMU_FUSE", "1")
    st1 = np.zeros(18, dtype=np.int32)
    assert emu.emu_apply(None, n, nat.gate_array_ptr(recs), len(recs), 4, 8, 4, 16,
                         st1.ctypes.data_as(ctypes.c_void_p)) == 0
    monkeypatch.setenv("QS_EMU_FUSE", "2")
    out, ref, st2 = run_both(emu, c, n, C128)
    assert np.max(np.abs(out - ref)) < 1e-12
    if moves:
        assert st2[17] > 0 and 100 * st2[0] <= 92 * st1[0]
    else:
        assert st2[17] == 0 and st2[0] == st1[0]


@pytest.mark.parametrize("lr", [1])
def test_qubits_moved_with_low_bits_in_registers(emu, lr, monkeypatch):
    monkeypatch.setenv("QS_FUSION_LOWREGS", str(lr))
    monkeypatch.setenv("QS_FUSION_REMAP", "1")
    c = workloads.random_mix_circuit(20, 900, 41, workloads.RANDOM_GATE_PROBS_60_40)
    out, ref, st = run_both(emu, c, 20, C128)
    assert np.max(np.abs(out - ref)) < 1e-12
    assert st[17] > 0


def test_from_zero_state_vs_numpy_oracle(emu):
    n = 14
    c = all_types(n, 250, 9)
    recs, keep = compile_gates(c.gates, n)
    v = np.zeros(1 << n, dtype=np.complex128)
    v[0] = 1
    st = np.zeros(18, dtype=np.int32)
    rc = emu.emu_apply(v.ctypes.data_as(ctypes.c_void_p), n, nat.gate_array_ptr(recs), len(recs), 4, 8, 4, 16,
                       st.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    ref = osv.run(all_types(n, 250, 9))
    assert np.max(np.abs(v - ref)) < 1e-12


def dm_records(circuit, n, noise_model):
    """The density-matrix simulator's record program for a noisy circuit."""
    from quict_amd.simulation.density_matrix import program_records
    return program_records(noise_model.transpile(circuit, True).gates, n)


@pytest.mark.parametrize("shape", [C128, C64])
def test_noise_superoperators_and_xx_ops(emu, shape):
    """Density-matrix programs (noise superoperators with real payloads) and
    Rxx / Ryy / FSim gates (complex payloads) use the diagonal + anti-diagonal
    two-target op; results must still match the reference kernels."""
    n = 7
    c = workloads.random_mix_circuit(n, 60, 12)
    recs, keep = dm_records(c, n, workloads.dm_noise_model())
    rng = np.random.RandomState(5)
    v = rng.randn(1 << (2 * n)) + 1j * rng.randn(1 << (2 * n))
    v /= np.linalg.norm(v)
    ref = v.copy()
    emu.cpu_apply_gates(ref.ctypes.data_as(ctypes.c_void_p), 2 * n, nat.gate_array_ptr(recs), len(recs), 1)
    out = v.copy()
    st = np.zeros(18, dtype=np.int32)
    RB, TB, S = shape
    rc = emu.emu_apply(out.ctypes.data_as(ctypes.c_void_p), 2 * n, nat.gate_array_ptr(recs), len(recs), RB, TB, 4, S,
                       st.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, emu.emu_last_error()
    assert np.max(np.abs(out - ref)) < 1e-12
    assert st[16] > 0  # diagonal + anti-diagonal ops were used

    from quict_amd.core import Rxx, Ryy
    c2 = Circuit(14)
    random.seed(3)
    for k in range(40):
        a, b = random.sample(range(14), 2)
        [Rxx(0.1 * k + 0.2), Ryy(0.3 * k + 0.1), FSim(0.2 * k, 0.7)][k % 3] | c2([a, b])
        H | c2(random.randrange(14))
        X | c2(random.randrange(14))
    out2, ref2, st2 = run_both(emu, c2, 14, shape, seed=2)
    assert np.max(np.abs(out2 - ref2)) < 1e-12
    assert st2[16] > 0
