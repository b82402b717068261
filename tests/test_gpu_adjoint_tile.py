"""The adjoint sweep in LDS tile passes (k_adj_tile, default) against the
CPU oracle's parameter shift and against the three-bit register passes it
replaces (QS_ADJ_TILE=0): hardware-efficient ansatz and mixed generators
with controls and two-target rotations, tiles of 5 to 11 amplitude bits (more
passes, steps cut at tile boundaries), widths below the tile size, both
precisions.  Tolerance 1e-10 (complex128) against the shift rule."""
import numpy as np
import pytest

from quict_amd.core import GateType
from quict_amd import workloads
from oracle import gradients as og

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def mixed_circuit(n, seed):
    from quict_amd.gradient.adjoint import ParameterizedCircuit
    rng = np.random.RandomState(seed)
    pc = ParameterizedCircuit(n)
    for q in range(n):
        pc.add(GateType.h, [q])
    k = 0
    kinds = [GateType.rx, GateType.ry, GateType.rz, GateType.crz, GateType.rzz, GateType.rxx, GateType.ryy,
             GateType.rzx]
    for _ in range(6 * n):
        gt = kinds[rng.randint(len(kinds))]
        if gt in (GateType.rx, GateType.ry, GateType.rz):
            pc.add(gt, [int(rng.randint(n))], param=k)
        else:
            a, b = rng.choice(n, 2, replace=False)
            pc.add(gt, [int(a), int(b)], param=k)
        k += 1
        a, b = rng.choice(n, 2, replace=False)
        pc.add(GateType.cx, [int(a), int(b)])
        if rng.rand() < 0.3:
            a, b, c = rng.choice(n, 3, replace=False)
            pc.add(GateType.ccx, [int(a), int(b), int(c)])
    th = rng.uniform(-np.pi, np.pi, k)
    return pc, th


@pytest.mark.parametrize("tile_bits", ["5", "8", "11"])
def test_tile_adjoint_mixed_vs_oracle_shift(tile_bits, monkeypatch):
    say("adjoint tile vs oracle shift, mixed generators, tile bits", tile_bits)
    from quict_amd.gradient.adjoint import PauliSum, adjoint_gradient
    monkeypatch.setenv("QS_ADJ_TILE_BITS", tile_bits)
    n = 10
    pc, th = mixed_circuit(n, 7)
    terms = workloads.zz_x_hamiltonian(n) + [(0.3, "XYZIXZYIXZ")]
    E, g = adjoint_gradient(pc, th, PauliSum(n, terms))
    assert abs(E - og.expectation(pc, th, terms)) < 1e-12
    ref = og.parameter_shift(pc, th, terms)
    err = float(np.max(np.abs(g - ref)))
    print(f"  max |tile adjoint - oracle shift| over {len(th)} parameters: {err:.2e}")
    assert err < 1e-10


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_tile_adjoint_narrow_states(n):
    say("adjoint tile, narrow state", n)
    from quict_amd.gradient.adjoint import ParameterizedCircuit, PauliSum, adjoint_gradient
    pc = ParameterizedCircuit.hea(n, 3)
    th = workloads.hea_params(n, 3, 11).ravel()
    terms = [(1.0, "Z" * n), (0.5, "X" + "I" * (n - 1))]
    _, g = adjoint_gradient(pc, th, PauliSum(n, terms))
    assert np.max(np.abs(g - og.parameter_shift(pc, th, terms))) < 1e-10


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("single", 2e-4)])
def test_tile_adjoint_equals_register_passes_20q(precision, tol, monkeypatch):
    """The 20-qubit, 12-layer ansatz (480 parameters): tile passes and register
    passes give the same full gradient."""
    say("adjoint tile vs register passes 20q", precision)
    from quict_amd.gradient.adjoint import ParameterizedCircuit, PauliSum, adjoint_gradient
    n, layers = 20, 12
    pc = ParameterizedCircuit.hea(n, layers)
    th = workloads.hea_params(n, layers, 3).ravel()
    H = PauliSum(n, workloads.zz_x_hamiltonian(n))
    E1, g1 = adjoint_gradient(pc, th, H, precision=precision)
    monkeypatch.setenv("QS_ADJ_TILE", "0")
    E0, g0 = adjoint_gradient(pc, th, H, precision=precision)
    err = float(np.max(np.abs(g1 - g0)))
    print(f"  max |tile - register| over 480 parameters: {err:.2e}")
    assert abs(E1 - E0) < tol
    assert err < tol
