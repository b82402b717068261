"""Pins the CPU oracle (CPU only, no GPU).

Checks chosen so that a dropped term, wrong sign, wrong bit order or a
transposed operand in the oracle fails at least one of them:
  * the reference's own golden 10-qubit vector (unit_test/simulation/data/
    state_vector.npy for random_circuit_for_correction.qasm) at 1e-12;
  * brute force: the dense circuit unitary from per-gate kron expansion;
  * closed forms for measurement probabilities and channels;
  * the literal dense density-matrix algorithm against the row/column one.
"""
import os
import random

import numpy as np
import pytest

from quict_amd.core import (Circuit, GateType, H, CX, Rx, Ry, Rz, X, Measure, Reset, Perm, Unitary, CCX, CSwap,
                            iSwap, Rzx, QFT, FSim)
from quict_amd.core import qasm
from quict_amd.core.noise import DepolarizingError, DampingError, NoiseModel
from quict_amd import workloads
from oracle import statevector as osv
from oracle import density_matrix as odm

ALL_TYPES = [GateType.h, GateType.hy, GateType.s, GateType.sdg, GateType.x, GateType.y, GateType.z, GateType.sx,
             GateType.sy, GateType.sw, GateType.id, GateType.u1, GateType.u2, GateType.u3, GateType.rx,
             GateType.ry, GateType.rz, GateType.t, GateType.tdg, GateType.phase, GateType.gphase, GateType.cz,
             GateType.cx, GateType.cy, GateType.ch, GateType.crz, GateType.cu1, GateType.cu3, GateType.fsim,
             GateType.rxx, GateType.ryy, GateType.rzz, GateType.rzx, GateType.swap, GateType.iswap,
             GateType.iswapdg, GateType.sqiswap, GateType.cswap, GateType.ccx, GateType.ccz, GateType.ccrz]


def golden(golden_dir):
    c = qasm.load_file(os.path.join(golden_dir, "random_circuit_for_correction.qasm"))
    sv = np.load(os.path.join(golden_dir, "state_vector.npy"))
    return c, sv


def test_oracle_matches_reference_golden_vector(golden_dir):
    c, sv = golden(golden_dir)
    out = osv.run(c)
    assert np.max(np.abs(out - sv)) < 1e-12


def test_golden_vector_equals_brute_force_unitary(golden_dir):
    c, sv = golden(golden_dir)
    assert np.max(np.abs(c.matrix()[:, 0] - sv)) < 1e-12


@pytest.mark.parametrize("seed", range(6))
def test_oracle_vs_brute_force_all_gate_types(seed):
    n = 6
    c = Circuit(n)
    random.seed(seed)
    np.random.seed(seed)
    c.random_append(80, typelist=ALL_TYPES, random_params=True)
    psi0 = np.random.RandomState(seed).randn(1 << n) + 1j * np.random.RandomState(seed + 7).randn(1 << n)
    psi0 /= np.linalg.norm(psi0)
    out = osv.run(c, state_vector=psi0)
    ref = c.matrix() @ psi0
    assert np.max(np.abs(out - ref)) < 1e-12


def test_oracle_dense_multi_target_unitary_and_qft():
    n = 7
    rng = np.random.RandomState(3)
    q, _ = np.linalg.qr(rng.randn(8, 8) + 1j * rng.randn(8, 8))
    c = Circuit(n)
    H | c(0)
    Unitary(q) | c([5, 1, 3])
    QFT(4) | c([6, 2, 0, 4])
    Rzx(0.7) | c([2, 5])
    out = osv.run(c)
    assert np.max(np.abs(out - c.matrix()[:, 0])) < 1e-12


def test_measure_probability_closed_form():
    # Ry(theta)|0>: P0 = cos^2(theta/2)
    for theta in [0.1, 1.0, 2.5]:
        c = Circuit(1)
        Ry(theta) | c(0)
        v = osv.run(c)
        assert abs(osv.prob_zero(v, 1, 0) - np.cos(theta / 2) ** 2) < 1e-14


def test_measure_collapse_and_seeded_outcome():
    # Hand-worked: H|0>, P0 = 0.5; random.seed(5).random() = 0.6229... > 0.5 -> outcome 1 -> |1>
    c = Circuit(1)
    H | c(0)
    Measure | c(0)
    random.seed(5)
    u = random.random()
    random.seed(5)
    v = osv.run(c)
    assert c.qubits[0].measured == int(u > 0.5)
    expect = np.array([0, 1]) if u > 0.5 else np.array([1, 0])
    assert np.allclose(v, expect, atol=1e-15)


def test_reset_and_perm():
    c = Circuit(3)
    X | c(0)
    H | c(2)
    Reset | c(0)
    v = osv.run(c)
    ref = Circuit(3)
    H | ref(2)
    assert np.allclose(v, osv.run(ref), atol=1e-14)
    # Perm on 3 qubits [1, 2, 0]: mapping = pargs; source qubit i = output qubit mapping[i]
    c = Circuit(3)
    X | c(0)
    Perm(3, [1, 2, 0]) | c([0, 1, 2])
    v = osv.run(c)
    # |100> (qubit0=1); output qubit mapping[0]=1 carries source qubit 0 -> |010>
    assert np.argmax(np.abs(v)) == 0b010


def test_ghz_sampling_support_and_seeded_counts():
    n = 4
    c = Circuit(n)
    H | c(0)
    for i in range(n - 1):
        CX | c([i, i + 1])
    v = osv.run(c)
    random.seed(11)
    counts = osv.sample(v, n, 200)
    assert counts[0] + counts[-1] == 200
    # conditional-mass formulation gives the same outcomes from the same draws
    random.seed(11)
    u = np.array([[random.random() for _ in range(n)] for _ in range(200)])
    idx = osv.sample_outcomes(v, n, u)
    assert np.bincount(idx, minlength=1 << n).tolist() == counts


def test_sampling_chi_square():
    from scipy.stats import chisquare
    n = 5
    c = workloads.random_mix_circuit(n, 40, 9)
    v = osv.run(c)
    p = np.abs(v) ** 2
    random.seed(2)
    u = np.array([[random.random() for _ in range(n)] for _ in range(20000)])
    idx = osv.sample_outcomes(v, n, u)
    obs = np.bincount(idx, minlength=1 << n)
    mask = p * 20000 > 5
    exp = p[mask] * 20000
    stat, pval = chisquare(obs[mask], exp * obs[mask].sum() / exp.sum())
    assert pval > 1e-4


def test_dm_golden_reconstructed(golden_dir):
    c, sv = golden(golden_dir)
    rho = odm.run_rows_cols(c)
    assert np.max(np.abs(rho - np.outer(sv, sv.conj()))) < 1e-12


def test_dm_dense_vs_rows_cols_noisy():
    n = 4
    c = workloads.random_mix_circuit(n, 30, 5)
    nm = NoiseModel(basic_gates=[t.name for t in workloads.RANDOM_GATE_MIX])
    nm.add_noise_for_all_qubits(DepolarizingError(0.05), [t.name for t in workloads.RANDOM_GATE_MIX[:10]])
    nm.add_noise_for_all_qubits(DampingError(0.1, 0, 0).tensor(DampingError(0.1, 0, 0)),
                                [t.name for t in workloads.RANDOM_GATE_MIX[10:]])
    noisy = nm.transpile(c, accumulated_mode=True)
    a = odm.run_dense(noisy)
    b = odm.run_rows_cols(noisy)
    assert np.max(np.abs(a - b)) < 1e-12
    assert abs(np.trace(a) - 1) < 1e-12
    assert np.allclose(a, a.conj().T, atol=1e-14)
    assert np.min(np.linalg.eigvalsh(a)) > -1e-12


def test_channel_closed_forms():
    rng = np.random.RandomState(1)
    A = rng.randn(2, 2) + 1j * rng.randn(2, 2)
    rho = A @ A.conj().T
    rho /= np.trace(rho)
    dep = DepolarizingError(0.3)
    out = sum(k @ rho @ k.conj().T for k in dep.kraus)
    assert np.allclose(out, odm.depolarize_closed_form(rho, 0.3), atol=1e-14)
    damp = DampingError(0.2, 0, 0)
    out = sum(k @ rho @ k.conj().T for k in damp.kraus)
    assert np.allclose(out, odm.amp_damp_closed_form(rho, 0.2), atol=1e-14)
    # NoiseGate superoperator acting on vec(rho)
    from quict_amd.core import NoiseGate
    ng = NoiseGate(damp.kraus, 1)
    vec = ng.superoperator() @ rho.ravel()
    assert np.allclose(vec.reshape(2, 2), odm.amp_damp_closed_form(rho, 0.2), atol=1e-14)


def test_tensor_channel_factors_compose_to_the_channel():
    """A tensored channel after a two-qubit gate equals the gate, then the first
    factor on the first argument, then the second on the second (the form the
    density-matrix simulator applies).  Different damping strengths on the two
    factors make a swapped order fail."""
    from quict_amd.core import CRz
    rng = np.random.RandomState(7)
    A = rng.randn(4, 4) + 1j * rng.randn(4, 4)
    rho = A @ A.conj().T
    rho /= np.trace(rho)
    d1, d2 = DampingError(0.05, 0, 0), DampingError(0.4, 0, 0)
    t = d1.tensor(d2)
    assert t.factors == [d1, d2]
    assert d1.factors == [d1]
    U = CRz(0.7).matrix.astype(np.complex128)
    whole = sum(k @ rho @ k.conj().T for k in t.apply_to_gate(U))
    step = U @ rho @ U.conj().T
    k1 = [np.kron(k, np.identity(2)) for k in d1.kraus]
    k2 = [np.kron(np.identity(2), k) for k in d2.kraus]
    step = sum(k @ step @ k.conj().T for k in k1)
    step = sum(k @ step @ k.conj().T for k in k2)
    assert np.max(np.abs(whole - step)) < 1e-14
    # the per-qubit closed forms agree with the factor order too
    nm = NoiseModel(basic_gates=["crz"])
    nm.add_noise_for_all_qubits(t, ["crz"])
    c = Circuit(2)
    CRz(0.7) | c([0, 1])
    noisy = nm.transpile(c, accumulated_mode=True)
    ng = noisy.gates[0]
    assert ng.source[1] is t and ng.source[0].type == GateType.crz
