"""State-vector parity on the GPU: the HIP path (through the C-ABI) against the oracle.

complex128 must match within 1e-12 max-abs; complex64 error against the
complex128 oracle is asserted per case (and printed).
"""
import os
import random

import numpy as np
import pytest

from quict_amd.core import (Circuit, GateType, H, CX, Ry, Rz, X, Measure, Reset, Perm, Unitary, QFT, IQFT, CCX,
                            CSwap, Rzx, FSim, Swap)
from quict_amd.core import qasm
from quict_amd import workloads
from oracle import statevector as osv
from test_oracle import ALL_TYPES

pytestmark = pytest.mark.gpu


def sim(precision="double", fusion=True):
    from quict_amd.simulation import StateVectorSimulator
    return StateVectorSimulator("GPU", precision=precision, fusion=fusion)


def say(*a):
    print("[case]", *a, flush=True)


def random_all_types(n, ngates, seed):
    c = Circuit(n)
    random.seed(seed)
    np.random.seed(seed)
    c.random_append(ngates, typelist=ALL_TYPES, random_params=True)
    return c


@pytest.mark.parametrize("fusion", [False, True])
def test_golden_vector_c128(golden_dir, fusion):
    say("golden c128 fusion", fusion)
    c = qasm.load_file(os.path.join(golden_dir, "random_circuit_for_correction.qasm"))
    ref = np.load(os.path.join(golden_dir, "state_vector.npy"))
    out = sim("double", fusion).run(c).get()
    assert np.max(np.abs(out - ref)) < 1e-12


def test_golden_vector_c64(golden_dir):
    say("golden c64")
    c = qasm.load_file(os.path.join(golden_dir, "random_circuit_for_correction.qasm"))
    ref = np.load(os.path.join(golden_dir, "state_vector.npy"))
    out = sim("single").run(c).get()
    err = np.max(np.abs(out - ref))
    print("complex64 max-abs error vs complex128 golden:", err)
    assert err < 1e-6


@pytest.mark.parametrize("n,ngates,seed", [(3, 60, 1), (6, 120, 2), (10, 200, 3), (13, 200, 4), (14, 300, 5),
                                           (16, 300, 6), (18, 300, 7), (20, 400, 8)])
@pytest.mark.parametrize("fusion", [False, True])
def test_random_all_gate_types_vs_oracle(n, ngates, seed, fusion):
    say("random all types", n, ngates, seed, "fusion", fusion)
    c = random_all_types(n, ngates, seed)
    ref = osv.run(c)
    s = sim("double", fusion)
    out = s.run(c).get()
    assert np.max(np.abs(out - ref)) < 1e-12


@pytest.mark.parametrize("n,seed", [(14, 11), (20, 12)])
def test_random_c64_error_vs_c128_oracle(n, seed):
    say("c64 random", n, seed)
    c = random_all_types(n, 200, seed)
    ref = osv.run(c)
    out = sim("single").run(c).get()
    err = np.max(np.abs(out.astype(np.complex128) - ref))
    print(f"complex64 max-abs error vs complex128 oracle (n={n}, 200 gates): {err:.3e}")
    assert err < 1e-5


@pytest.mark.parametrize("fusion", [False, True])
def test_dense_multi_target_gates(fusion):
    say("dense multi-target fusion", fusion)
    n = 16
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
    ref = osv.run(c)
    out = sim("double", fusion).run(c).get()
    assert np.max(np.abs(out - ref)) < 1e-12


def test_fused_equals_unfused_26q_and_inverse_roundtrip():
    say("26q fused vs unfused + inverse round trip")
    n = 26
    c = workloads.random_mix_circuit(n, 300, 21, workloads.RANDOM_GATE_PROBS_60_40)
    a = sim("double", True)
    va = a.run(c).tensor.clone()
    b = sim("double", False)
    vb = b.run(c).tensor
    diff = (va - vb).abs().max().item()
    assert diff < 1e-12
    assert abs((va.abs() ** 2).sum().item() - 1) < 1e-12
    assert a.last_passes < b.last_passes
    print("passes fused / unfused:", a.last_passes, b.last_passes)
    inv = Circuit(n)
    for g in reversed(c.gates):
        inv.append(g.inverse())
    back = a.run(inv, state_vector=va).get()
    e0 = np.zeros(1 << n, dtype=np.complex128)
    e0[0] = 1
    assert np.max(np.abs(back - e0)) < 1e-11


@pytest.mark.parametrize("tile_log2,low_bits,prefetch,precision", [
    (11, 4, 0, "double"), (13, 4, 0, "double"), (12, 2, 0, "double"), (12, 6, 0, "double"),
    (12, 4, 1, "double"), (13, 4, 0, "single"), (12, 3, 1, "single")])
def test_fusion_settings_match_unfused(tile_log2, low_bits, prefetch, precision):
    """Every tile shape / low-bit count / launch variant reachable through
    qs_set_fusion and qs_set_fusion_launch gives the per-gate result."""
    say("fusion settings", tile_log2, low_bits, prefetch, precision)
    import ctypes
    from quict_amd import _native as nat
    lib = nat.load_library()
    ctx = nat.context(0)
    n = 20
    c = workloads.random_mix_circuit(n, 400, 31, workloads.RANDOM_GATE_PROBS_60_40)
    ref = sim(precision, False).run(c).get()
    try:
        nat.check(lib.qs_set_fusion(ctx, tile_log2, low_bits))
        nat.check(lib.qs_set_fusion_launch(ctx, prefetch, 2))
        out = sim(precision, True).run(c).get()
    finally:
        lib.qs_set_fusion(ctx, 0, 0)
        lib.qs_set_fusion_launch(ctx, 0, 2)
    tol = 1e-12 if precision == "double" else 1e-5
    assert np.max(np.abs(out - ref)) < tol


def test_low_bits_beyond_the_thread_slots_are_refused():
    """An 11-bit tile has 7 thread slots: qs_set_fusion refuses 8 low bits
    before anything runs; 7 are accepted."""
    say("qs_set_fusion refuses low_bits > tile_log2 - 4")
    from quict_amd import _native as nat
    lib = nat.load_library()
    ctx = nat.context(0)
    try:
        assert lib.qs_set_fusion(ctx, 11, 8) != 0
        assert b"low_bits" in lib.qs_last_error()
        nat.check(lib.qs_set_fusion(ctx, 11, 7))
    finally:
        lib.qs_set_fusion(ctx, 0, 0)


def test_wide_controlled_gate_between_fused_passes():
    """A 9-controlled X (10-qubit Unitary on qubits 0..9) inside a fused
    20-qubit circuit: it has more bits above the low bits than a tile holds,
    so it runs as its own per-gate pass; fused equals per-gate to 1e-12."""
    say("wide controlled gate in a fused circuit 20q")
    from quict_amd.core import Unitary
    n = 20
    m = np.eye(1 << 10, dtype=np.complex128)
    m[-2:, -2:] = [[0, 1], [1, 0]]
    c = workloads.random_mix_circuit(n, 200, 41, workloads.RANDOM_GATE_PROBS_60_40)
    Unitary(m) | c(list(range(10)))
    for g in workloads.random_mix_circuit(n, 200, 42, workloads.RANDOM_GATE_PROBS_60_40).gates:
        c.append(g)
    ref = sim("double", False).run(c).get()
    out = sim("double", True).run(c).get()
    assert np.max(np.abs(out - ref)) < 1e-12


@pytest.mark.parametrize("precision", ["double", "single"])
def test_qubits_moved_onto_low_bits_match_unfused(precision, monkeypatch):
    """Planner option QS_FUSION_REMAP=1: stores that move the qubits needed next
    onto the low index bits (the same tile addresses, permuted index bits), and
    passes without gates that put every qubit back; 22 qubits, 1200 gates,
    where the option takes fewer passes than the default plan."""
    say("remapped plan 22q", precision)
    n = 22
    c = workloads.random_mix_circuit(n, 1200, 37, workloads.RANDOM_GATE_PROBS_60_40)
    ref = sim(precision, False).run(c).get()
    monkeypatch.setenv("QS_FUSION_REMAP", "0")
    plain = sim(precision, True)
    out0 = plain.run(c).get()
    monkeypatch.setenv("QS_FUSION_REMAP", "1")
    moved = sim(precision, True)
    out1 = moved.run(c).get()
    print(f"  passes: default {plain.last_passes}, moving qubits {moved.last_passes}")
    assert moved.last_passes < plain.last_passes
    tol = 1e-12 if precision == "double" else 1e-5
    assert np.max(np.abs(out1 - ref)) < tol
    assert np.max(np.abs(out1 - out0)) < tol


def test_relabel_and_flip_passes_are_exact():
    """Uncontrolled X and Swap gates compile to relabelings of the tile layout
    (no arithmetic): the load and store layouts of such passes differ.  Every
    amplitude of the prepared state has its own phase, so a misplaced amplitude
    shows as an O(1e-4) error.  Many workgroups per pass (24 qubits) so that the
    waves of one workgroup run out of step."""
    say("relabel / flip passes 24q")
    n = 24
    rng = random.Random(17)
    c = Circuit(n)
    for q in range(n):
        H | c(q)
        Rz(0.05 + 0.11 * q) | c(q)
    for _ in range(200):
        if rng.random() < 0.5:
            X | c(rng.randrange(n))
        else:
            Swap | c(rng.sample(range(n), 2))
    va = sim("double", True).run(c).get()
    vb = sim("double", False).run(c).get()
    assert np.max(np.abs(va - vb)) < 1e-12


def test_measure_seeded_matches_oracle():
    say("measure seeded")
    n = 10
    c = workloads.random_mix_circuit(n, 100, 31)
    for q in [3, 0, 7]:
        Measure | c(q)
    H | c(3)
    Measure | c(9)
    c2 = workloads.random_mix_circuit(n, 100, 31)
    for q in [3, 0, 7]:
        Measure | c2(q)
    H | c2(3)
    Measure | c2(9)
    random.seed(77)
    ref = osv.run(c2)
    ref_meas = [c2.qubits[q].measured for q in range(n)]
    random.seed(77)
    out = sim("double").run(c).get()
    got = [c.qubits[q].measured for q in range(n)]
    assert got == ref_meas
    assert np.max(np.abs(out - ref)) < 1e-12


def test_prob_zero_values():
    say("prob zero")
    from quict_amd import _native as nat
    n = 12
    c = workloads.random_mix_circuit(n, 80, 41)
    s = sim("double")
    v = s.run(c).get()
    for bit in [0, 5, 11]:
        p = s.get_measured_prob(bit)
        assert abs(p - osv.prob_zero(v, n, bit)) < 1e-14


def test_reset_and_perm():
    say("reset + perm")
    n = 9
    c = workloads.random_mix_circuit(n, 60, 51)
    Reset | c(4)
    Perm(4, [2, 0, 3, 1]) | c([1, 5, 7, 8])
    Reset | c(0)
    c2 = workloads.random_mix_circuit(n, 60, 51)
    Reset | c2(4)
    Perm(4, [2, 0, 3, 1]) | c2([1, 5, 7, 8])
    Reset | c2(0)
    ref = osv.run(c2)
    out = sim("double").run(c).get()
    assert np.max(np.abs(out - ref)) < 1e-12


@pytest.mark.parametrize("precision", ["double", "single"])
def test_sample_seeded_exact_vs_sequential_oracle(precision):
    say("sample seeded", precision)
    n = 8
    c = workloads.random_mix_circuit(n, 80, 61)
    ref_state = osv.run(c)
    random.seed(1001)
    ref_counts = osv.sample(ref_state, n, 300)
    s = sim(precision)
    s.run(workloads.random_mix_circuit(n, 80, 61))
    random.seed(1001)
    counts = s.sample(300)
    if precision == "double":
        assert counts == ref_counts
    else:
        mism = sum(abs(a - b) for a, b in zip(counts, ref_counts)) // 2
        print("complex64 sampling: shots differing from the complex128 sequential oracle:", mism)
        assert mism <= 3


def test_sample_config_20q_matches_conditional_oracle():
    say("20q config sample")
    c = workloads.sv20_circuit()
    ref = osv.run(workloads.sv20_circuit())
    s = sim("double")
    out = s.run(c).get()
    assert np.max(np.abs(out - ref)) < 1e-12
    random.seed(1001)
    u = np.array([[random.random() for _ in range(20)] for _ in range(1024)])
    ref_idx = osv.sample_outcomes(ref, 20, u)
    random.seed(1001)
    idx = s.sample_indices(1024)
    assert np.array_equal(idx.astype(np.int64), ref_idx)


def test_ghz_support():
    say("ghz")
    n = 5
    c = Circuit(n)
    H | c(0)
    for i in range(n - 1):
        CX | c([i, i + 1])
    for prec in ["double", "single"]:
        s = sim(prec)
        s.run(c)
        counts = s.sample(100)
        assert counts[0] + counts[-1] == 100


def test_fused_calls_return_before_their_passes_finish():
    """qs_sv_apply_gates queues its fused passes and returns: right after a
    29-qubit 300-gate call the stream still has work; three calls in a row
    (program slots reused while earlier passes may still run) give the result
    of one call on the concatenated gates and the oracle's, to 1e-12."""
    say("async fused calls")
    import ctypes
    import torch
    from quict_amd import _native as nat
    from quict_amd.simulation.gate_compiler import compile_gates
    lib = nat.load_library()
    ctx = nat.context(0)
    n = 29
    v = torch.zeros(1 << n, dtype=torch.complex64, device="cuda:0")
    v[0] = 1
    c = workloads.random_mix_circuit(n, 300, 81, workloads.RANDOM_GATE_PROBS_60_40)
    recs, keep = compile_gates(c.gates, n)
    torch.cuda.synchronize()
    nat.check(lib.qs_sv_apply_gates(ctx, ctypes.c_void_p(v.data_ptr()), n, nat.QS_C64, nat.gate_array_ptr(recs),
                                    len(recs), 2, None, nat.stream_ptr(0)))
    pending = not torch.cuda.current_stream().query()
    torch.cuda.synchronize()
    assert pending, "the call waited for its passes"
    del v
    n = 20
    parts = [workloads.random_mix_circuit(n, 300, 82 + k, workloads.RANDOM_GATE_PROBS_60_40) for k in range(3)]
    whole = Circuit(n)
    for p in parts:
        for g in p.gates:
            whole.append(g)
    outs = []
    for chunks in ([whole], parts):
        v = torch.zeros(1 << n, dtype=torch.complex128, device="cuda:0")
        v[0] = 1
        torch.cuda.synchronize()
        for p in chunks:
            r, kp = compile_gates(p.gates, n)
            nat.check(lib.qs_sv_apply_gates(ctx, ctypes.c_void_p(v.data_ptr()), n, nat.QS_C128,
                                            nat.gate_array_ptr(r), len(r), 1, None, nat.stream_ptr(0)))
            del r, kp  # the records are read during the call
        outs.append(v.cpu().numpy())
    assert np.max(np.abs(outs[0] - outs[1])) < 1e-12
    assert np.max(np.abs(outs[1] - osv.run(whole))) < 1e-12


def test_sample_above_24_qubits_returns_nonzero_counts():
    """Above 24 qubits sample() returns {basis index: count} of the outcomes
    that occurred (a 2^n list is out of reach at 30 qubits: 2^30 Python ints);
    its entries equal the seeded per-shot outcomes, and the Simulator result
    keys them by n-character bit strings as the reference's Result does."""
    say("sample dict above 24 qubits")
    from quict_amd.simulation import StateVectorSimulator, Simulator
    n = 25
    c = workloads.random_mix_circuit(n, 60, 91)
    sim = StateVectorSimulator("GPU", precision="single")
    sim.run(c)
    random.seed(17)
    counts = sim.sample(500)
    assert isinstance(counts, dict) and sum(counts.values()) == 500
    random.seed(17)
    idx = sim.sample_indices(500)
    vals, cnt = np.unique(idx, return_counts=True)
    assert counts == {int(v): int(k) for v, k in zip(vals, cnt)}
    random.seed(17)
    res = Simulator(device="GPU", backend="state_vector", precision="single").run(c, shots=500)
    assert res["shots"] == 500
    assert res["data"]["counts"] == {format(k, "b").zfill(n): v for k, v in sorted(counts.items())}
