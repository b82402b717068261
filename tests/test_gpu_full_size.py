"""The benchmark workloads at their full sizes, through size-independent
properties (the oracle cannot run them):

* 30-qubit complex128 state, the seeded 2000-gate 60/40 circuit: fused passes
  equal one pass per gate (max-abs <= 1e-12), the norm stays 1, and the circuit
  followed by its inverse returns |0...0> (<= 1e-10);
* the same circuit in complex64: norm within 1e-4 and fused vs unfused within
  the single-precision error printed;
* 14-qubit noisy density matrix (300 gates, depolarizing and damping):
  trace 1, real non-negative diagonal, Hermitian on sampled entries, fused
  equal to unfused to 1e-12.
"""
import numpy as np
import pytest

from quict_amd import workloads
from quict_amd.core import Circuit

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def test_headline_30q_c128_properties():
    say("30q c128 headline: fused vs unfused, norm, inverse round trip")
    import torch
    from quict_amd.simulation import StateVectorSimulator
    n = 30
    fused = StateVectorSimulator(precision="double")
    va = fused.run(workloads.sv30_circuit(n=n)).tensor
    assert fused.last_passes < 100
    norm = torch.linalg.vector_norm(va).item()
    assert abs(norm - 1) < 1e-12
    unf = StateVectorSimulator(precision="double", fusion=False)
    vb = unf.run(workloads.sv30_circuit(n=n)).tensor
    diff = (va - vb).abs().max().item()
    print(f"  fused passes {fused.last_passes}, unfused {unf.last_passes}, max diff {diff:.2e}")
    assert diff < 1e-12
    del unf, vb
    torch.cuda.empty_cache()
    c = workloads.sv30_circuit(n=n)
    inv = Circuit(n)
    for g in reversed(c.gates):
        inv.append(g.inverse())
    back = fused.run(inv, use_previous=True).tensor
    err0 = abs(back[0].item() - 1)
    back[0] = 0
    rest = back.abs().max().item()
    print(f"  round trip: |a0 - 1| = {err0:.2e}, max |a_j|, j > 0: {rest:.2e}")
    assert err0 < 1e-10 and rest < 1e-10
    del fused, va, back
    torch.cuda.empty_cache()


def test_headline_30q_c64_properties():
    say("30q c64 headline: norm, fused vs unfused")
    import torch
    from quict_amd.simulation import StateVectorSimulator
    n = 30
    fused = StateVectorSimulator(precision="single")
    va = fused.run(workloads.sv30_circuit(n=n)).tensor
    norm = torch.linalg.vector_norm(va.to(torch.complex128)).item()
    unf = StateVectorSimulator(precision="single", fusion=False)
    vb = unf.run(workloads.sv30_circuit(n=n)).tensor
    diff = (va - vb).abs().max().item()
    print(f"  |norm - 1| = {abs(norm - 1):.2e}, fused vs unfused max diff {diff:.2e}")
    assert abs(norm - 1) < 1e-4
    assert diff < 1e-5
    del fused, unf, va, vb
    torch.cuda.empty_cache()


def test_dm14_noisy_properties():
    say("14q noisy density matrix: trace, diagonal, hermiticity, fused vs unfused")
    import torch
    from quict_amd.simulation import DensityMatrixSimulator
    nm = workloads.dm_noise_model()
    a = DensityMatrixSimulator(precision="double", accumulated_mode=True)
    ra = a.run(workloads.dm14_circuit(), noise_model=nm).tensor
    tr = a.trace()
    d = a.diagonal()
    assert abs(tr - 1) < 1e-12
    assert d.min() > -1e-15
    D = 1 << 14
    R = ra.view(D, D)
    rng = np.random.RandomState(3)
    for _ in range(64):
        i, j = (int(x) for x in rng.randint(0, D, 2))
        assert abs(R[i, j].item() - R[j, i].item().conjugate()) < 1e-14
    b = DensityMatrixSimulator(precision="double", accumulated_mode=True, fusion=False)
    rb = b.run(workloads.dm14_circuit(), noise_model=workloads.dm_noise_model()).tensor
    diff = (ra - rb).abs().max().item()
    print(f"  passes fused {a.last_passes} / unfused {b.last_passes}, max diff {diff:.2e}")
    assert diff < 1e-12
    del a, b, ra, rb, R
    torch.cuda.empty_cache()
