"""Adjoint gradients on the GPU against parameter-shift and finite differences.

Tolerance: |g_adjoint - g_shift| <= 1e-10 (complex128); finite differences
(h = 1e-6) are a coarse check (<= 1e-6)."""
import numpy as np
import pytest

from quict_amd.core import GateType
from quict_amd import workloads
from oracle import gradients as og

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def test_single_qubit_closed_form():
    say("closed form")
    from quict_amd.gradient.adjoint import ParameterizedCircuit, PauliSum, adjoint_gradient
    pc = ParameterizedCircuit(1).add(GateType.ry, [0], param=0)
    H = PauliSum(1, [(1.0, "Z")])
    for th in [0.3, 1.7, -2.2]:
        E, g = adjoint_gradient(pc, np.array([th]), H)
        assert abs(E - np.cos(th)) < 1e-14
        assert abs(g[0] + np.sin(th)) < 1e-14


@pytest.mark.parametrize("n,layers,seed", [(6, 3, 1), (10, 4, 2)])
def test_hea_adjoint_vs_oracle_shift(n, layers, seed):
    say("hea adjoint vs cpu shift", n, layers)
    from quict_amd.gradient.adjoint import ParameterizedCircuit, PauliSum, adjoint_gradient
    pc = ParameterizedCircuit.hea(n, layers)
    th = workloads.hea_params(n, layers, seed).ravel()
    terms = workloads.zz_x_hamiltonian(n)
    E, g = adjoint_gradient(pc, th, PauliSum(n, terms))
    assert abs(E - og.expectation(pc, th, terms)) < 1e-12
    ref = og.parameter_shift(pc, th, terms)
    assert np.max(np.abs(g - ref)) < 1e-10


def test_hea_14q_adjoint_vs_oracle_shift_subset():
    """14 qubits, 4 layers (112 parameters): the GPU adjoint gradient against the
    CPU oracle's parameter shift on 10 parameters spread over the layers."""
    say("hea adjoint 14q vs cpu shift subset")
    from quict_amd.gradient.adjoint import ParameterizedCircuit, PauliSum, adjoint_gradient
    n, layers = 14, 4
    pc = ParameterizedCircuit.hea(n, layers)
    th = workloads.hea_params(n, layers, 6).ravel()
    terms = workloads.zz_x_hamiltonian(n)
    E, g = adjoint_gradient(pc, th, PauliSum(n, terms))
    assert abs(E - og.expectation(pc, th, terms)) < 1e-12
    sub = [0, 13, 14, 27, 40, 55, 61, 80, 97, 111]
    ref = og.parameter_shift(pc, th, terms, params=sub)
    err = float(np.max(np.abs(g[sub] - ref[sub])))
    print(f"  max |adjoint - oracle shift| on 10 parameters: {err:.2e}")
    assert err < 1e-10


def test_mixed_generators_adjoint():
    say("mixed generators")
    from quict_amd.gradient.adjoint import ParameterizedCircuit, PauliSum, adjoint_gradient
    n = 5
    pc = ParameterizedCircuit(n)
    k = 0
    for q in range(n):
        pc.add(GateType.h, [q])
    for (gt, qs) in [(GateType.rx, [0]), (GateType.crz, [1, 3]), (GateType.rzz, [2, 4]), (GateType.rxx, [0, 2]),
                     (GateType.ryy, [3, 1]), (GateType.rzx, [4, 0]), (GateType.ry, [2]), (GateType.rz, [1])]:
        pc.add(gt, qs, param=k)
        k += 1
        pc.add(GateType.cx, [qs[0], (qs[0] + 1) % n])
    th = np.random.RandomState(5).uniform(-np.pi, np.pi, k)
    terms = workloads.zz_x_hamiltonian(n) + [(0.4, "XYZIX")]
    E, g = adjoint_gradient(pc, th, PauliSum(n, terms))
    ref = og.parameter_shift(pc, th, terms)
    assert np.max(np.abs(g - ref)) < 1e-10


def test_hea_20q_adjoint_vs_gpu_shift_and_fd():
    say("20q adjoint vs gpu shift subset")
    from quict_amd.gradient.adjoint import (ParameterizedCircuit, PauliSum, adjoint_gradient,
                                            parameter_shift_gradient, finite_difference_gradient)
    n, layers = 20, 12
    pc = ParameterizedCircuit.hea(n, layers)
    th = workloads.hea_params(n, layers, 3).ravel()
    H = PauliSum(n, workloads.zz_x_hamiltonian(n))
    E, g = adjoint_gradient(pc, th, H)
    assert g.shape == (480,)
    sub = [0, 7, 19, 20, 133, 250, 401, 479]
    ps = parameter_shift_gradient(pc, th, H, params=sub)
    fd = finite_difference_gradient(pc, th, H, params=sub[:3])
    err = np.max(np.abs(g[sub] - ps[sub]))
    print("max |adjoint - shift| on sampled params:", err)
    assert err < 1e-10
    assert np.max(np.abs(g[sub[:3]] - fd[sub[:3]])) < 1e-6


def test_autograd_function():
    say("autograd")
    import torch
    from quict_amd.gradient.adjoint import ParameterizedCircuit, PauliSum, AdjointExpectation
    n, layers = 6, 2
    pc = ParameterizedCircuit.hea(n, layers)
    H = PauliSum(n, workloads.zz_x_hamiltonian(n))
    th = torch.tensor(workloads.hea_params(n, layers, 4).ravel(), dtype=torch.float64, requires_grad=True)
    E = AdjointExpectation.apply(th, pc, H)
    (2 * E).backward()
    ref = og.parameter_shift(pc, th.detach().numpy(), workloads.zz_x_hamiltonian(n))
    assert np.max(np.abs(th.grad.numpy() - 2 * ref)) < 1e-10
