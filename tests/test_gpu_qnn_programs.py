"""Batched QNN with per-circuit gate programs (qs_qnn_batch_programs): a batch of
mixed ansaetze and widths against the CPU oracle (expectations, and the loss
gradient from the oracle's parameter shift), the input gradient against finite
differences, and the hardware-efficient ansatz written as programs against the
hard-wired batch kernel.  complex128 at 1e-12 / 1e-10; the complex64 error
against the same oracle is printed and bounded."""
import numpy as np
import pytest

from quict_amd.core import GateType
from oracle import gradients as og

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def family(kind, w, nparams, rng):
    """Three ansatz families over a shared parameter vector."""
    from quict_amd.gradient.adjoint import ParameterizedCircuit
    pc = ParameterizedCircuit(w)
    k = int(rng.randint(nparams))
    nxt = lambda: int(rng.randint(nparams))
    if kind == 0:  # Ry Rz + CX ladder
        for _ in range(2):
            for q in range(w):
                pc.add(GateType.ry, [q], param=nxt())
                pc.add(GateType.rz, [q], param=nxt())
            for q in range(w - 1):
                pc.add(GateType.cx, [q, q + 1])
    elif kind == 1:  # Rx + fixed Rz + CZ ring
        for _ in range(2):
            for q in range(w):
                pc.add(GateType.rx, [q], param=nxt())
                pc.add(GateType.rz, [q], value=float(rng.uniform(-np.pi, np.pi)))
            for q in range(w):
                pc.add(GateType.cz, [q, (q + 1) % w])
    else:  # star of CX from qubit 0, mixed rotations
        for q in range(w):
            pc.add([GateType.rx, GateType.ry, GateType.rz][q % 3], [q], param=nxt())
        for q in range(1, w):
            pc.add(GateType.cx, [0, q])
        for q in range(w):
            pc.add(GateType.ry, [q], param=k)
    pc.nparams = nparams
    return pc


def batch(seed, ncirc=10, nparams=24, nfeat=5):
    rng = np.random.RandomState(seed)
    circuits, encode = [], []
    for c in range(ncirc):
        w = int(rng.randint(2, 9))
        circuits.append(family(c % 3, w, nparams, rng))
        encode.append([(q, int(rng.randint(nfeat))) for q in range(w) if rng.rand() < 0.7])
    x = rng.randn(ncirc, nfeat)
    t = 1.0 - 2.0 * (x[:, 0] > 0)
    theta = rng.uniform(-np.pi, np.pi, nparams)
    return circuits, encode, x, t, theta


def with_encoding(pc, enc, xrow, theta):
    """The circuit with its encodings as fixed-angle Ry gates and one parameter
    per trainable occurrence (the two-term shift rule holds per occurrence, not
    for a parameter shared by several gates): (circuit, expanded angles, the
    shared parameter of each occurrence)."""
    from quict_amd.gradient.adjoint import ParameterizedCircuit
    full = ParameterizedCircuit(pc.n)
    for q, j in enc:
        full.add(GateType.ry, [q], value=float(xrow[j]))
    owner = []
    for gtype, qubits, param, vals in pc.ops:
        if param is not None:
            full.add(gtype, qubits, param=len(owner))
            owner.append(param)
        else:
            full.add(gtype, qubits, value=vals)
    full.nparams = len(owner)
    return full, np.array([theta[p] for p in owner]), owner


@pytest.mark.parametrize("precision", ["double", "single"])
def test_mixed_ansatz_batch_vs_cpu_oracle(precision):
    say("qnn programs mixed batch vs oracle", precision)
    from quict_amd.gradient.qnn import QNNProgramBatch
    circuits, encode, x, t, theta = batch(3)
    n = len(circuits)
    for lds in (11, 4):  # LDS-resident, then wider circuits in the global slot
        qb = QNNProgramBatch(circuits, x, t, len(theta), encode=encode, precision=precision,
                             config={"lds_max_width": lds, "block_threads": 128, "blocks_per_cu": 1})
        loss, ev, g, gx = qb.forward_backward(theta)
        rev, rg = np.zeros(n), np.zeros(len(theta))
        for c, pc in enumerate(circuits):
            full, th_occ, owner = with_encoding(pc, encode[c], x[c], theta)
            z = [(1.0, "Z" + "I" * (pc.n - 1))]
            rev[c] = og.expectation(full, th_occ, z)
            ps = og.parameter_shift(full, th_occ, z)
            for m, p_ in enumerate(owner):
                rg[p_] += 2.0 * (rev[c] - t[c]) / n * ps[m]
        rl = float(np.mean((rev - t) ** 2))
        err_e, err_g = float(np.max(np.abs(ev - rev))), float(np.max(np.abs(g - rg)))
        print(f"  {precision} lds {lds}: max |E - oracle| {err_e:.2e}, max |grad - oracle shift| {err_g:.2e}")
        if precision == "double":
            assert err_e < 1e-12 and err_g < 1e-10 and abs(loss - rl) < 1e-12
        else:
            assert err_e < 1e-5 and err_g < 1e-5


def test_input_gradient_by_finite_differences():
    say("qnn programs grad x")
    from quict_amd.gradient.qnn import QNNProgramBatch
    circuits, encode, x, t, theta = batch(5, ncirc=6)
    loss, ev, g, gx = QNNProgramBatch(circuits, x, t, len(theta), encode=encode).forward_backward(theta)
    h = 1e-6
    for c, j in [(0, 0), (1, 3), (4, 2), (5, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[c, j] += h
        xm[c, j] -= h
        lp = QNNProgramBatch(circuits, xp, t, len(theta), encode=encode).forward_backward(theta)[0]
        lm = QNNProgramBatch(circuits, xm, t, len(theta), encode=encode).forward_backward(theta)[0]
        assert abs(gx[c, j] - (lp - lm) / (2 * h)) < 1e-7


def test_hea_programs_equal_the_hard_wired_batch():
    """The configuration ansatz written as programs (fixed-angle encodings) gives
    the hard-wired kernel's expectations, loss and parameter gradient."""
    say("qnn programs vs hard-wired HEA")
    from quict_amd.gradient.qnn import QNNBatch, QNNProgramBatch, synthetic_batch, circuit_for
    widths, x, t = synthetic_batch(48, 16, 4, 12, 8)
    layers, mw = 3, 12
    theta = np.random.RandomState(9).uniform(-np.pi, np.pi, layers * 2 * mw)
    ref = QNNBatch(widths, x, t, layers=layers, max_width=mw).forward_backward(theta)
    circuits = [circuit_for(int(w), x[c], layers, mw) for c, w in enumerate(widths)]
    out = QNNProgramBatch(circuits, x, t, len(theta)).forward_backward(theta)
    assert abs(out[0] - ref[0]) < 1e-12
    assert np.max(np.abs(out[1] - ref[1])) < 1e-12
    assert np.max(np.abs(out[2] - ref[2])) < 1e-10
