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


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("single", 2e-4)])
@pytest.mark.parametrize("tile_bits", ["5", "11"])
def test_forward_from_the_sweep_plan_equals_a_separate_forward(precision, tol, tile_bits, monkeypatch):
    """qs_adjoint_gradient_pauli with the forward records (psi from |0>, the
    sweep's tile passes in reverse with conjugate-transposed steps, gates
    outside the passes from the forward records) against the same call on a
    forward state made by qs_sv_apply_gates: energy, gradient and the forward
    state itself."""
    say("adjoint forward from the sweep plan", precision, "tile bits", tile_bits)
    import ctypes
    import torch
    from quict_amd import _native as nat
    from quict_amd.gradient import adjoint as A
    monkeypatch.setenv("QS_ADJ_TILE_BITS", tile_bits)
    n = 10
    pc, th = mixed_circuit(n, 12)
    H = A.PauliSum(n, workloads.zz_x_hamiltonian(n) + [(0.3, "XYZIXZYIXZ")])
    recs = H.records()
    eng = A._engine(n, precision, 0)
    fwd, inv, der, slot, keep = A._records(pc, th)

    def run(with_fwd):
        g = np.zeros(pc.nparams)
        E = ctypes.c_double(0)
        if not with_fwd:
            eng.forward(fwd)
        nat.check(eng.lib.qs_adjoint_gradient_pauli(
            eng.ctx, eng.ptr(eng.psi), eng.ptr(eng.lam), n, eng.dcode, ctypes.cast(recs, ctypes.c_void_p), len(recs),
            nat.gate_array_ptr(fwd) if with_fwd else None, nat.gate_array_ptr(inv), len(inv),
            nat.gate_array_ptr(der), slot.ctypes.data_as(ctypes.c_void_p), g.ctypes.data_as(ctypes.c_void_p),
            len(g), ctypes.byref(E), eng.stream()), "qs_adjoint_gradient_pauli")
        return E.value, g

    E0, g0 = run(False)
    E1, g1 = run(True)
    err = float(np.max(np.abs(g1 - g0)))
    print(f"  |dE| {abs(E1 - E0):.2e}, max |d grad| {err:.2e}")
    assert abs(E1 - E0) < tol and err < tol
    # the sweep returns psi to |0...0>: the forward state itself, checked through the engine
    torch.cuda.synchronize()
    psi = eng.psi.cpu().numpy()
    assert abs(abs(psi[0]) - 1.0) < tol
    if precision == "double":
        assert abs(E1 - og.expectation(pc, th, [(c, s) for c, s in H.terms])) < 1e-12


@pytest.mark.parametrize("tile", ["5", "11", "off"])
def test_three_target_gates_between_tile_passes(tile, monkeypatch):
    """Fixed three-target unitaries (no tile pass takes them: they run as
    per-gate passes through the context's scratch buffer, forward and
    backward) among the mixed generators: gradient against the oracle's shift."""
    say("adjoint with three-target unitaries, tile", tile)
    from quict_amd.gradient.adjoint import PauliSum, adjoint_gradient
    if tile == "off":
        monkeypatch.setenv("QS_ADJ_TILE", "0")
    else:
        monkeypatch.setenv("QS_ADJ_TILE_BITS", tile)
    n = 9
    pc, th = mixed_circuit(n, 21)
    rng = np.random.RandomState(5)
    ops = list(pc.ops)
    pc.ops = []
    for j, op in enumerate(ops):
        pc.ops.append(op)
        if j % 37 == 5:
            u, _ = np.linalg.qr(rng.randn(8, 8) + 1j * rng.randn(8, 8))
            pc.ops.append((GateType.unitary, [int(x) for x in rng.choice(n, 3, replace=False)], None, [u]))
    pc._template = None
    terms = workloads.zz_x_hamiltonian(n) + [(0.3, "XYZIXZYIX")]
    E, g = adjoint_gradient(pc, th, PauliSum(n, terms))
    assert abs(E - og.expectation(pc, th, terms)) < 1e-12
    ref = og.parameter_shift(pc, th, terms)
    err = float(np.max(np.abs(g - ref)))
    print(f"  max |adjoint - oracle shift| over {len(th)} parameters: {err:.2e}, |grad| max {np.max(np.abs(ref)):.2f}")
    assert err < 1e-10 and np.max(np.abs(ref)) > 1e-3
