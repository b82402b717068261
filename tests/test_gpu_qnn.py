"""Batched QNN launch on the GPU against per-circuit references.

Reference per circuit: the same circuit built gate by gate (16 Ry encodings,
HEA layers) evaluated by the CPU oracle (small widths: <Z_0> and the
parameter-shift gradient) or by the single-circuit GPU adjoint path
(larger widths), then combined into the mean loss and its gradient."""
import numpy as np
import pytest

from quict_amd import workloads
from oracle import gradients as og

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def reference(widths, x, targets, theta, layers, max_width, use_gpu_adjoint=False):
    from quict_amd.gradient.qnn import circuit_for
    from quict_amd.gradient.adjoint import PauliSum, adjoint_gradient
    N = len(widths)
    ev = np.zeros(N)
    g = np.zeros(layers * 2 * max_width)
    for c in range(N):
        w = int(widths[c])
        pc = circuit_for(w, x[c], layers, max_width)
        zterm = [(1.0, "Z" + "I" * (w - 1))]
        if use_gpu_adjoint:
            e, gc = adjoint_gradient(pc, theta, PauliSum(w, zterm))
        else:
            e = og.expectation(pc, theta, zterm)
            gc = og.parameter_shift(pc, theta, zterm,
                                    params=[(l * 2 + k) * max_width + q for l in range(layers) for k in range(2)
                                            for q in range(w)])
        ev[c] = e
        g += 2 * (e - targets[c]) / N * gc
    loss = np.mean((ev - targets) ** 2)
    return loss, ev, g


@pytest.mark.parametrize("use_graph", [0, 1])
def test_small_batch_vs_cpu_shift(use_graph):
    say("qnn small batch graph", use_graph)
    from quict_amd.gradient.qnn import QNNBatch
    widths = np.array([3, 5, 4, 6, 2, 5], dtype=np.int32)
    rng = np.random.RandomState(1)
    x = rng.randn(6, 7)
    t = 1 - 2 * (x[:, 0] > 0)
    layers, mw = 2, 6
    theta = rng.uniform(-np.pi, np.pi, layers * 2 * mw)
    qb = QNNBatch(widths, x, t, layers=layers, max_width=mw,
                  config={"lds_max_width": 4, "block_threads": 128, "blocks_per_cu": 1, "use_graph": use_graph})
    loss, ev, gth, gx = qb.forward_backward(theta)
    rl, rev, rg = reference(widths, x, t, theta, layers, mw)
    assert np.max(np.abs(ev - rev)) < 1e-12
    assert abs(loss - rl) < 1e-12
    assert np.max(np.abs(gth - rg)) < 1e-10
    # second call reuses the captured graph with new inputs
    theta2 = theta * 0.5
    loss2, ev2, g2, _ = qb.forward_backward(theta2)
    rl2, rev2, rg2 = reference(widths, x, t, theta2, layers, mw)
    assert np.max(np.abs(ev2 - rev2)) < 1e-12 and np.max(np.abs(g2 - rg2)) < 1e-10


def test_grad_x_by_finite_differences():
    say("qnn grad x")
    from quict_amd.gradient.qnn import QNNBatch
    widths = np.array([4, 3], dtype=np.int32)
    rng = np.random.RandomState(2)
    x = rng.randn(2, 6)
    t = np.array([1.0, -1.0])
    theta = rng.uniform(-np.pi, np.pi, 2 * 2 * 4)
    qb = QNNBatch(widths, x, t, layers=2, max_width=4)
    loss, ev, gth, gx = qb.forward_backward(theta)
    h = 1e-6
    for c, j in [(0, 0), (0, 5), (1, 2)]:
        xp, xm = x.copy(), x.copy()
        xp[c, j] += h
        xm[c, j] -= h
        lp = QNNBatch(widths, xp, t, layers=2, max_width=4).forward_backward(theta)[0]
        lm = QNNBatch(widths, xm, t, layers=2, max_width=4).forward_backward(theta)[0]
        assert abs(gx[c, j] - (lp - lm) / (2 * h)) < 1e-7


@pytest.mark.parametrize("precision", ["double", "single"])
def test_config_batch_sample_vs_gpu_adjoint(precision):
    say("qnn config sample", precision)
    from quict_amd.gradient.qnn import QNNBatch, synthetic_batch
    widths, x, t = synthetic_batch(1024, 16, 8, 18, 4)
    layers, mw = 6, 18
    theta = np.random.RandomState(7).uniform(-np.pi, np.pi, layers * 2 * mw)
    qb = QNNBatch(widths, x, t, layers=layers, max_width=mw, precision=precision)
    loss, ev, gth, gx = qb.forward_backward(theta)
    # check a width-spanning subset of circuits against the single-circuit GPU adjoint path
    idx = [int(np.where(widths == w)[0][0]) for w in range(8, 19)]
    sub_w, sub_x, sub_t = widths[idx], x[idx], t[idx]
    sb = QNNBatch(sub_w, sub_x, sub_t, layers=layers, max_width=mw, precision=precision)
    sl, sev, sg, _ = sb.forward_backward(theta)
    rl, rev, rg = reference(sub_w, sub_x, sub_t, theta, layers, mw, use_gpu_adjoint=True)
    tol_e, tol_g = (1e-12, 1e-10) if precision == "double" else (1e-4, 1e-4)
    err_e = np.max(np.abs(sev - rev))
    err_g = np.max(np.abs(sg - rg))
    print(f"{precision}: max |E - ref| = {err_e:.2e}, max |grad - ref| = {err_g:.2e}")
    assert err_e < tol_e and err_g < tol_g
    assert np.max(np.abs(ev[idx] - sev)) < (1e-12 if precision == "double" else 1e-5)
    assert np.isfinite(loss) and np.all(np.isfinite(gth))


@pytest.mark.parametrize("precision", ["double", "single"])
def test_config_batch_widths_8_to_14_vs_cpu_oracle(precision):
    """Circuits of the 1024-circuit configuration batch at widths 8..14: each
    one's <Z_0> from the full batch against the CPU oracle, and its batched
    gradient (as a one-circuit batch, loss (E - t)^2) against
    2 (E - t) * the oracle's parameter shift on four parameters.  complex128 at
    1e-12 / 1e-10; the complex64 error against the same oracle is printed."""
    say("qnn config widths 8..14 vs cpu oracle", precision)
    from quict_amd.gradient.qnn import QNNBatch, synthetic_batch, circuit_for
    widths, x, t = synthetic_batch(1024, 16, 8, 18, 4)
    layers, mw = 6, 18
    theta = np.random.RandomState(7).uniform(-np.pi, np.pi, layers * 2 * mw)
    qb = QNNBatch(widths, x, t, layers=layers, max_width=mw, precision=precision)
    _, ev, _, _ = qb.forward_backward(theta)
    err_e = err_g = 0.0
    for w in range(8, 15):
        c = int(np.where(widths == w)[0][0])
        pc = circuit_for(w, x[c], layers, mw)
        z = [(1.0, "Z" + "I" * (w - 1))]
        e_ref = og.expectation(pc, theta, z)
        err_e = max(err_e, abs(ev[c] - e_ref))
        params = [(0 * 2 + 0) * mw + 0, (2 * 2 + 1) * mw + w // 2, (4 * 2 + 0) * mw + w - 1, (5 * 2 + 1) * mw + 1]
        ps = og.parameter_shift(pc, theta, z, params=params)
        one = QNNBatch(widths[c:c + 1], x[c:c + 1], t[c:c + 1], layers=layers, max_width=mw, precision=precision)
        _, ev1, g1, _ = one.forward_backward(theta)
        err_g = max(err_g, float(np.max(np.abs(g1[params] - 2 * (e_ref - t[c]) * ps[params]))))
    print(f"  {precision}: max |E - oracle| = {err_e:.2e}, max |grad - oracle shift| = {err_g:.2e}")
    if precision == "double":
        assert err_e < 1e-12 and err_g < 1e-10
    else:
        assert err_e < 1e-4 and err_g < 1e-4


def test_batch_gradient_is_bit_reproducible_across_launch_shapes():
    """Per-circuit gradient terms summed in circuit order: two runs, and runs
    with a different number of resident workgroups (same block size), give
    bit-identical gradients."""
    say("qnn determinism")
    from quict_amd.gradient.qnn import QNNBatch, synthetic_batch
    widths, x, t = synthetic_batch(256, 16, 8, 14, 9)
    theta = np.random.RandomState(3).uniform(-np.pi, np.pi, 4 * 2 * 14)
    outs = []
    for per_cu, graph in [(2, 1), (2, 1), (1, 0), (4, 1)]:
        qb = QNNBatch(widths, x, t, layers=4, max_width=14,
                      config={"lds_max_width": 11, "block_threads": 256, "blocks_per_cu": per_cu, "use_graph": graph})
        loss, ev, g, gx = qb.forward_backward(theta)
        outs.append((loss, ev.copy(), g.copy()))
    for loss, ev, g in outs[1:]:
        assert loss == outs[0][0]
        assert np.array_equal(ev, outs[0][1])
        assert np.array_equal(g, outs[0][2])


@pytest.mark.parametrize("fuse", [1, 2, 3])
def test_fusion_depth_vs_cpu_shift(fuse):
    """Rotations of 1, 2 or 3 qubits per pass (odd widths leave a shorter last
    group): the small batch against the CPU oracle's parameter shift, LDS-resident
    and global-slot states both."""
    say("qnn fusion depth", fuse)
    from quict_amd.gradient.qnn import QNNBatch
    widths = np.array([3, 5, 4, 7, 2, 6, 5], dtype=np.int32)
    rng = np.random.RandomState(12)
    x = rng.randn(7, 9)
    t = 1 - 2 * (x[:, 0] > 0)
    layers, mw = 3, 7
    theta = rng.uniform(-np.pi, np.pi, layers * 2 * mw)
    rl, rev, rg = reference(widths, x, t, theta, layers, mw)
    for lds in (4, 7):  # widths above `lds` live in the global slot
        qb = QNNBatch(widths, x, t, layers=layers, max_width=mw,
                      config={"lds_max_width": lds, "block_threads": 128, "blocks_per_cu": 1, "use_graph": 1,
                              "fuse_qubits": fuse})
        loss, ev, gth, gx = qb.forward_backward(theta)
        assert np.max(np.abs(ev - rev)) < 1e-12
        assert abs(loss - rl) < 1e-12
        assert np.max(np.abs(gth - rg)) < 1e-10


@pytest.mark.parametrize("precision,tol", [("double", 1e-11), ("single", 2e-4)])
def test_fusion_depths_agree_on_config_batch(precision, tol):
    """The 1024-circuit configuration batch (widths 8..18): fusion depths 2 and 3
    give the loss, <Z_0> and gradient of depth 1."""
    say("qnn fusion depths agree", precision)
    from quict_amd.gradient.qnn import QNNBatch, synthetic_batch
    widths, x, t = synthetic_batch(1024, 16, 8, 18, 4)
    layers, mw = 6, 18
    theta = np.random.RandomState(7).uniform(-np.pi, np.pi, layers * 2 * mw)
    outs = {}
    for fuse in (1, 2, 3):
        qb = QNNBatch(widths, x, t, layers=layers, max_width=mw, precision=precision,
                      config={"lds_max_width": 11, "block_threads": 256, "blocks_per_cu": 2, "use_graph": 1,
                              "fuse_qubits": fuse})
        outs[fuse] = qb.forward_backward(theta)
        print(f"  fuse {fuse}: {qb.last_ms:.1f} ms")
    for fuse in (2, 3):
        assert abs(outs[fuse][0] - outs[1][0]) < tol
        assert np.max(np.abs(outs[fuse][1] - outs[1][1])) < tol
        assert np.max(np.abs(outs[fuse][2] - outs[1][2])) < tol


def test_plan_reuse_when_a_wider_circuit_needs_the_global_slots():
    """Two batches of one shape (so the captured plan could be reused): the first
    fits LDS (no global state slots are allocated), the second has a circuit
    wider than the LDS limit; both against the CPU oracle."""
    say("qnn plan reuse, slots")
    from quict_amd.gradient.qnn import QNNBatch
    rng = np.random.RandomState(21)
    x = rng.randn(3, 5)
    t = np.array([1.0, -1.0, 1.0])
    layers, mw = 2, 8
    theta = rng.uniform(-np.pi, np.pi, layers * 2 * mw)
    cfg = {"lds_max_width": 6, "block_threads": 128, "blocks_per_cu": 1, "use_graph": 1, "fuse_qubits": 2}
    for widths in ([4, 5, 6], [4, 8, 6]):
        w = np.array(widths, dtype=np.int32)
        loss, ev, gth, gx = QNNBatch(w, x, t, layers=layers, max_width=mw, config=cfg).forward_backward(theta)
        rl, rev, rg = reference(w, x, t, theta, layers, mw)
        assert np.max(np.abs(ev - rev)) < 1e-12 and np.max(np.abs(gth - rg)) < 1e-10
