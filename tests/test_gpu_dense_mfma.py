"""Dense gates on 4..6 targets through the matrix-core kernel (k_dense_mfma,
v_mfma_f64_16x16x4 / v_mfma_f32_16x16x4) against the oracle's mat-vec
(cpu_calculator.py:139-186 restated) and against the per-amplitude kernel it
replaces (QS_DENSE_MFMA=0): random unitaries, QFT / IQFT blocks, controlled
blocks, targets on the low bits, the high bits and scattered, both
precisions."""
import numpy as np
import pytest

from quict_amd import workloads
from quict_amd.core import Circuit, H, Unitary, QFT, IQFT
from oracle import statevector as osv

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def rand_unitary(k, seed):
    rng = np.random.RandomState(seed)
    q, r = np.linalg.qr(rng.randn(1 << k, 1 << k) + 1j * rng.randn(1 << k, 1 << k))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def circuit_with(n, blocks, seed):
    c = Circuit(n)
    for q in range(n):  # every amplitude non-zero, controlled subspaces included
        H | c(q)
    for g in workloads.random_mix_circuit(n, 40, seed).gates:
        c.append(g)
    for g, args in blocks:
        g | c(args)
        for q in range(0, n, 3):
            H | c(q)
    return c


CASES = {
    "u5_scattered": (16, lambda: [(Unitary(rand_unitary(5, 1)), [3, 11, 0, 7, 14])]),
    "u6_low_bits": (16, lambda: [(Unitary(rand_unitary(6, 2)), [15, 14, 13, 12, 11, 10])]),
    "u6_high_bits": (17, lambda: [(Unitary(rand_unitary(6, 3)), [0, 1, 2, 3, 4, 5])]),
    "qft6_iqft5": (15, lambda: [(QFT(6), [2, 9, 4, 13, 0, 7]), (IQFT(5), [14, 1, 8, 5, 10])]),
    "controlled_u5": (16, lambda: [(Unitary(np.block([[np.eye(96), np.zeros((96, 32))],
                                                     [np.zeros((32, 96)), rand_unitary(5, 4)]])),
                                    [6, 12, 1, 9, 3, 15, 0])]),
    "u4_per_gate": (14, lambda: [(Unitary(rand_unitary(4, 5)), [8, 2, 13, 5])]),
}


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_matrix_core_dense_matches_the_oracle(name, precision, monkeypatch):
    say("dense mfma", name, precision)
    from quict_amd.simulation import StateVectorSimulator
    n, blocks = CASES[name]
    c = circuit_with(n, blocks(), 60)
    ref = osv.run(circuit_with(n, blocks(), 60))
    fused = name != "u4_per_gate"  # 4-target blocks fuse into the tile pass; per gate they take this kernel
    out = StateVectorSimulator("GPU", precision=precision, fusion=fused).run(c).get()
    err = float(np.max(np.abs(out - ref)))
    print(f"  max-abs error vs the oracle: {err:.2e}")
    assert err < (1e-12 if precision == "double" else 2e-5)
    monkeypatch.setenv("QS_DENSE_MFMA", "0")
    naive = StateVectorSimulator("GPU", precision=precision, fusion=fused).run(c).get()
    assert np.max(np.abs(naive - ref)) < (1e-12 if precision == "double" else 2e-5)


def test_matrix_core_dense_at_24_qubits_norm_and_inverse():
    """24 qubits: a 6-target Unitary then its inverse restores the state; the
    norm stays 1."""
    say("dense mfma 24q round trip")
    import torch
    from quict_amd.simulation import StateVectorSimulator
    n = 24
    u = rand_unitary(6, 9)
    c = workloads.random_mix_circuit(n, 60, 61)
    sim = StateVectorSimulator("GPU", precision="double")
    before = sim.run(c).get()
    c2 = Circuit(n)
    Unitary(u) | c2([20, 3, 11, 7, 0, 16])
    Unitary(u.conj().T) | c2([20, 3, 11, 7, 0, 16])
    mid = sim.run(c2, state_vector=torch.as_tensor(before, device="cuda:0")).get()
    assert abs(np.linalg.norm(mid) - 1) < 1e-12
    assert np.max(np.abs(mid - before)) < 1e-12


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("n", [9, 15])
def test_three_target_dense_in_place_per_gate(n, precision):
    """Per-gate passes (fusion off, and every gate of a register below the
    tile size) of three-target dense gates, uncontrolled and with one or two
    controls, on low, high and scattered bits: the in-place k_dense3 against
    the oracle."""
    say("dense3 per gate", n, precision)
    from quict_amd.simulation import StateVectorSimulator
    ctl2 = np.block([[np.eye(24), np.zeros((24, 8))], [np.zeros((8, 24)), rand_unitary(3, 8)]])
    ctl1 = np.block([[np.eye(8), np.zeros((8, 8))], [np.zeros((8, 8)), rand_unitary(3, 9)]])
    blocks = lambda: [(Unitary(rand_unitary(3, 6)), [n - 1, n - 2, n - 3]), (Unitary(rand_unitary(3, 7)), [0, 1, 2]),
                      (QFT(3), [4, 1, n - 2]), (Unitary(ctl2), [3, n - 1, 0, 5, 2]),
                      (Unitary(ctl1), [n - 3, 1, 7, 4])]
    c = circuit_with(n, blocks(), 61)
    ref = osv.run(circuit_with(n, blocks(), 61))
    out = StateVectorSimulator("GPU", precision=precision, fusion=False).run(c).get()
    err = float(np.max(np.abs(out - ref)))
    print(f"  max-abs error vs the oracle: {err:.2e}")
    assert err < (1e-12 if precision == "double" else 2e-5)
