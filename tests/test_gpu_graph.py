"""Captured circuits (qs_graph_capture / qs_graph_launch) on the GPU.

A captured circuit must give exactly what the eager path gives: the same
program runs through the same kernels, so every replay is compared bit for
bit with an eager run, with other GPU work and synchronisations between the
replays (a graph node that misbehaves on replays after the first shows up
there).  Also: capture without the |0...0> initialisation applies the sequence
to whatever the buffer holds; sequences and registers that are not captured
are refused with an error.
"""
import ctypes
import random

import numpy as np
import pytest

from quict_amd import _native as nat
from quict_amd import workloads
from quict_amd.core import Circuit, H, QFT

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("n", [16, 20])
def test_replays_equal_eager_runs(n, precision):
    say("graph replay", n, precision)
    import torch
    from quict_amd.simulation import StateVectorSimulator
    c = workloads.random_mix_circuit(n, 300, 21)
    eager = StateVectorSimulator(precision=precision).run(workloads.random_mix_circuit(n, 300, 21)).get()
    sim = StateVectorSimulator(precision=precision)
    cap = sim.capture(c)
    assert cap.passes > 0
    other = StateVectorSimulator(precision=precision)
    for k in range(3):
        out = cap.replay().get()
        assert np.array_equal(out, eager), f"replay {k}"
        # unrelated eager work between replays
        other.run(workloads.random_mix_circuit(n, 50, 100 + k))
        torch.ones(1 << 16, device="cuda:0").sum().item()
        torch.cuda.synchronize()
    # the simulator samples the replayed state exactly as after an eager run
    ref = StateVectorSimulator(precision=precision)
    ref.run(workloads.random_mix_circuit(n, 300, 21))
    random.seed(5)
    a = ref.sample(200)
    cap.replay()
    random.seed(5)
    b = sim.sample(200)
    assert a == b
    cap.close()


def test_replay_of_a_plan_with_qubit_moves(monkeypatch):
    """A circuit long enough for the planner to move qubits onto the low index
    bits between passes (20 qubits, 900 gates: fewer passes than without
    moves): replays equal the eager run bit for bit and the per-gate result
    to 1e-12."""
    say("graph replay with qubit moves 20q 900 gates")
    monkeypatch.delenv("QS_FUSION_REMAP", raising=False)  # the default fuse=2 rule plans both runs
    from quict_amd.simulation import StateVectorSimulator
    n = 20
    c = workloads.random_mix_circuit(n, 900, 33, workloads.RANDOM_GATE_PROBS_60_40)
    eager_sim = StateVectorSimulator(precision="double")
    eager = eager_sim.run(c).get()
    unfused = StateVectorSimulator(precision="double", fusion=False).run(c).get()
    assert np.max(np.abs(eager - unfused)) < 1e-12
    sim = StateVectorSimulator(precision="double")
    cap = sim.capture(c)
    assert cap.passes == eager_sim.last_passes
    for k in range(2):
        assert np.array_equal(cap.replay().get(), eager), f"replay {k}"
    cap.close()
    monkeypatch.setenv("QS_FUSION_REMAP", "0")
    plain = StateVectorSimulator(precision="double")
    plain.run(c)
    print(f"  passes with qubit moves {eager_sim.last_passes}, without {plain.last_passes}")
    assert eager_sim.last_passes < plain.last_passes


def test_capture_without_init_applies_to_the_buffer():
    say("graph no init")
    import torch
    lib, ctx, s = nat.load_library(), nat.context(0), nat.stream_ptr(0)
    from quict_amd.simulation.gate_compiler import compile_gates
    n = 15
    c = workloads.random_mix_circuit(n, 200, 8)
    recs, keep = compile_gates(c.gates, n)
    rng = np.random.RandomState(2)
    x = rng.randn(1 << n) + 1j * rng.randn(1 << n)
    x /= np.linalg.norm(x)
    v = torch.as_tensor(x, device="cuda:0")
    w = v.clone()
    g = ctypes.c_void_p(0)
    nat.check(lib.qs_graph_capture(ctx, ctypes.c_void_p(v.data_ptr()), n, nat.QS_C128, nat.gate_array_ptr(recs),
                                   len(recs), 0, ctypes.byref(g)))
    for _ in range(2):
        nat.check(lib.qs_graph_launch(g, s))
        nat.check(lib.qs_sv_apply_gates(ctx, ctypes.c_void_p(w.data_ptr()), n, nat.QS_C128,
                                        nat.gate_array_ptr(recs), len(recs), 1, None, s))
        torch.cuda.synchronize()
        assert torch.equal(v, w)
    lib.qs_graph_destroy(g)


def test_unsupported_captures_are_refused():
    say("graph refused")
    import torch
    lib, ctx = nat.load_library(), nat.context(0)
    from quict_amd.simulation.gate_compiler import compile_gates
    # register below 2^(tile+1) amplitudes
    small = Circuit(10)
    H | small(0)
    recs, keep = compile_gates(small.gates, 10)
    v = torch.zeros(1 << 10, dtype=torch.complex128, device="cuda:0")
    g = ctypes.c_void_p(0)
    rc = lib.qs_graph_capture(ctx, ctypes.c_void_p(v.data_ptr()), 10, nat.QS_C128, nat.gate_array_ptr(recs),
                              len(recs), 1, ctypes.byref(g))
    assert rc == nat.QS_ERR_UNSUPPORTED and not g.value
    # a five-target dense block runs as its own pass: not captured
    wide = Circuit(16)
    QFT(5) | wide([0, 3, 6, 9, 12])
    recs, keep = compile_gates(wide.gates, 16)
    v = torch.zeros(1 << 16, dtype=torch.complex128, device="cuda:0")
    rc = lib.qs_graph_capture(ctx, ctypes.c_void_p(v.data_ptr()), 16, nat.QS_C128, nat.gate_array_ptr(recs),
                              len(recs), 1, ctypes.byref(g))
    assert rc == nat.QS_ERR_UNSUPPORTED and not g.value
