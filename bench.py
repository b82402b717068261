"""Benchmark of the circuit-simulation hot path on one MI355X.

Headline workload (one "step" = simulate the whole circuit from |0...0>):
  30-qubit complex128 state vector (16 GiB), seeded random circuit of 2000
  gates from {H,X,Y,Z,S,T,Rx,Ry,Rz,U3,CX,CZ,CRz,Swap} (60% one-qubit / 40%
  two-qubit, seed 2), gates fused into register/LDS-resident tile passes.

value = achieved HBM GB/s over the timed steps = passes * 2 * 2^n * 16 B / time.
roofline = the dominant kernel (the fused tile pass): algorithmic bytes per
launch (2 * 2^n * 16 B, one read and one write of the state) / its average
launch time, measured with HIP events on the launch stream.
cpu_baseline = the C restatement of the reference CPU simulator (one pass per
gate, gather / mat-vec / scatter) on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--qubits 30] [--no-secondary]
       python bench.py --only {dm14,unitary14,adjoint20,qnn_c128,qnn_c64}   (one secondary, for profiling)
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "30q HBM GB/s + gates/s; adjoint vs param-shift grads/s at ≤1e-10 error"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def profile_read(lib, ctx, reset=True):
    ms = (ctypes.c_double * 4)()
    cnt = (ctypes.c_int64 * 4)()
    from quict_amd import _native as nat
    nat.check(lib.qs_profile_read(ctx, ms, cnt, 1 if reset else 0), "qs_profile_read")
    return list(ms), list(cnt)


def run_sv(n, circuit, precision, fusion, steps, warmup, lib, ctx):
    import torch
    from quict_amd.simulation import StateVectorSimulator
    sim = StateVectorSimulator("GPU", precision=precision, fusion=fusion)
    for _ in range(warmup):
        sim.run(circuit)
    torch.cuda.synchronize()
    profile_read(lib, ctx, reset=True)
    lib.qs_profile_enable(ctx, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.run(circuit)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    lib.qs_profile_enable(ctx, 0)
    ms, cnt = profile_read(lib, ctx, reset=True)
    passes = sim.last_passes
    del sim
    torch.cuda.empty_cache()
    return dt, passes, ms, cnt


def run_stream_copy(n, lib, ctx, reps=5):
    """Plain streaming copy of one state (2^n complex128) with the library's copy kernel:
    the achievable-bandwidth reference for the tile passes."""
    import torch
    from quict_amd import _native as nat
    src = torch.zeros(1 << n, dtype=torch.complex128, device="cuda")
    dst = torch.empty_like(src)
    nbytes = (1 << n) * 16
    stream = nat.stream_ptr(0)
    out = {}
    for nt in (0, 1):
        for _ in range(2):
            nat.check(lib.qs_stream_copy(ctx, ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()),
                                         nbytes, nt, stream), "qs_stream_copy")
        torch.cuda.synchronize()
        profile_read(lib, ctx, reset=True)
        lib.qs_profile_enable(ctx, 1)
        for _ in range(reps):
            nat.check(lib.qs_stream_copy(ctx, ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()),
                                         nbytes, nt, stream), "qs_stream_copy")
        torch.cuda.synchronize()
        lib.qs_profile_enable(ctx, 0)
        ms, cnt = profile_read(lib, ctx, reset=True)
        avg = ms[1] / max(cnt[1], 1)
        out["nontemporal" if nt else "plain"] = {"avg_ms": avg, "gbs": 2 * nbytes / (avg * 1e-3) / 1e9}
    del src, dst
    torch.cuda.empty_cache()
    return out


def run_adjoint_20q(reps=20, cpu=False):
    """20-qubit 12-layer hardware-efficient ansatz, 480 parameters, H = sum ZZ + 0.5 sum X:
    adjoint gradient vs the parameter-shift gradient, both on this GPU."""
    import numpy as np
    from quict_amd import workloads
    from quict_amd.gradient.adjoint import (ParameterizedCircuit, PauliSum, adjoint_gradient,
                                            parameter_shift_gradient)
    n, layers = 20, 12
    pc = ParameterizedCircuit.hea(n, layers)
    th = workloads.hea_params(n, layers, 3).ravel()
    H = PauliSum(n, workloads.zz_x_hamiltonian(n))
    for _ in range(3):  # steady state: both pinned program slots grown, records templated
        adjoint_gradient(pc, th, H)
    t0 = time.perf_counter()
    for _ in range(reps):
        E, g = adjoint_gradient(pc, th, H)
    t_adj = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    ps = parameter_shift_gradient(pc, th, H)
    t_ps = time.perf_counter() - t0
    err = float(np.max(np.abs(g - ps)))
    out = {"qubits": n, "params": pc.nparams, "energy": E, "adjoint_ms": t_adj * 1e3,
           "adjoint_grads_per_s": pc.nparams / t_adj, "shift_ms": t_ps * 1e3,
           "shift_grads_per_s": pc.nparams / t_ps, "speedup_vs_shift": t_ps / t_adj,
           "max_abs_err_vs_shift": err}
    if cpu:
        # CPU parameter shift: two expectations per parameter, each the whole
        # circuit through the C restatement (all cores) plus <psi|H|psi>
        from quict_amd.simulation.gate_compiler import compile_gates
        from oracle import cpu_baseline, gradients as og
        threads = min(os.cpu_count() or 1, 16)  # the host-CPU share of one GPU
        recs, keep = compile_gates(pc.bind(th).gates, n)
        v = np.zeros(1 << n, dtype=np.complex128)
        v[0] = 1
        t0 = time.perf_counter()
        cpu_baseline.apply_gates(v, n, recs, threads)
        e_cpu = float(np.real(np.vdot(v, og.pauli_apply(v, n, workloads.zz_x_hamiltonian(n)))))
        t_e = time.perf_counter() - t0
        out["cpu_baseline"] = {"shift_grads_per_s": 1.0 / (2 * t_e), "expectation_s": t_e, "threads": threads,
                               "kind": "port", "energy_check": abs(e_cpu - E),
                               "sample": "one expectation (708 gates + H) on the C restatement; shift = 2 per parameter"}
    return out


def run_qnn_batch(precision, reps=3, sample_circuits=3, sample_params=6):
    """1024 circuits, widths 8..18, 16 Ry-encoded features, 6 layers, loss mean((<Z0>-t)^2):
    one autotuned persistent launch for forward + backward, against per-circuit parameter
    shift on this GPU (a sample of circuits and parameters, scaled to a rate)."""
    import numpy as np
    from quict_amd.gradient.qnn import QNNBatch, synthetic_batch, circuit_for
    from quict_amd.gradient.adjoint import PauliSum, parameter_shift_gradient
    widths, x, t = synthetic_batch(1024, 16, 8, 18, 4)
    layers, mw = 6, 18
    theta = np.random.RandomState(7).uniform(-np.pi, np.pi, layers * 2 * mw)
    qb = QNNBatch(widths, x, t, layers=layers, max_width=mw, precision=precision)
    t0 = time.perf_counter()
    cfg = qb.autotune(theta, reps=2, use_cache=False)
    t_tune = time.perf_counter() - t0
    qb.forward_backward(theta)
    t0 = time.perf_counter()
    for _ in range(reps):
        qb.forward_backward(theta)
    t_b = (time.perf_counter() - t0) / reps
    derivs = int(sum(2 * layers * int(w) for w in widths))  # ansatz derivatives over the batch
    # per-circuit parameter shift on the same GPU, and the batched gradient of one circuit
    err = 0.0
    t_ps = 0.0
    n_ps = 0
    for c in range(sample_circuits):
        w = int(widths[c])
        pc = circuit_for(w, x[c], layers, mw)
        params = [(l * 2 + k) * mw + q for l in range(layers) for k in range(2) for q in range(w)][:sample_params]
        t0 = time.perf_counter()
        ps = parameter_shift_gradient(pc, theta, PauliSum(w, [(1.0, "Z" + "I" * (w - 1))]), precision,
                                      params=params)
        t_ps += time.perf_counter() - t0
        n_ps += len(params)
        one = QNNBatch(widths[c:c + 1], x[c:c + 1], t[c:c + 1], layers=layers, max_width=mw, precision=precision)
        _, ev1, g1, _ = one.forward_backward(theta)
        scale = 2 * (ev1[0] - t[c])
        err = max(err, float(np.max(np.abs(g1[params] - scale * ps[params]))))
    shift_rate = n_ps / t_ps
    return {"circuits": len(widths), "config": cfg, "autotune_s": t_tune, "ms_per_batch": t_b * 1e3,
            "circuits_per_s": len(widths) / t_b, "grads_per_s": derivs / t_b,
            "shift_grads_per_s": shift_rate, "speedup_vs_shift": (derivs / t_b) / shift_rate,
            "max_abs_err_vs_shift": err,
            "note": f"shift rate from {sample_circuits} circuits x {sample_params} parameters, two GPU "
                    "expectations each; error compares the batched gradient of one circuit with "
                    "2(E - t) * shift derivative"}


def run_dm14(lib, ctx, reps=3, cpu=False):
    """14-qubit noisy density matrix (2^28 complex128 entries): 300-gate circuit with
    depolarizing / amplitude-damping channels, then diagonal and trace."""
    import torch
    from quict_amd import workloads
    from quict_amd.simulation import DensityMatrixSimulator
    c = workloads.dm14_circuit()
    nm = workloads.dm_noise_model()
    sim = DensityMatrixSimulator("GPU", precision="double", accumulated_mode=True)
    sim.run(c, noise_model=nm)
    torch.cuda.synchronize()
    profile_read(lib, ctx, reset=True)
    lib.qs_profile_enable(ctx, 1)
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.run(c, noise_model=nm)
        d = sim.diagonal()
        tr = sim.trace()
    dt = (time.perf_counter() - t0) / reps
    lib.qs_profile_enable(ctx, 0)
    ms, cnt = profile_read(lib, ctx, reset=True)
    pass_bytes = 2 * (1 << 28) * 16
    passes = sim.last_passes
    kern_ms = (ms[0] + ms[1]) / reps
    out = {"qubits": 14, "gates": 300, "passes": passes, "ms_per_run": dt * 1e3, "kernel_ms_per_run": kern_ms,
           "gbs_wall": passes * pass_bytes / dt / 1e9, "gbs_kernels": passes * pass_bytes / (kern_ms * 1e-3) / 1e9,
           "trace": [tr.real, tr.imag], "diag_sum": float(d.sum())}
    del sim
    torch.cuda.empty_cache()
    if cpu:
        out["cpu_baseline"] = cpu_dm14(c, nm)
    return out


def cpu_dm14(c, nm, sample_records=6, dense_sizes=(5, 6, 7, 8)):
    """CPU legs of the noisy density matrix: (a) the restated superoperator path
    (C restatement, one pass per record over the 2n-bit image) at 14 qubits, a
    sample of records on all cores scaled to the whole program; (b) the literal
    reference algorithm (dense U rho U^dagger per gate, Kraus sums) timed at
    5..8 qubits and extrapolated to 14 along its fitted exponent."""
    import numpy as np
    from quict_amd import workloads
    from quict_amd.simulation.density_matrix import program_records
    from oracle import cpu_baseline, density_matrix as odm
    n = 14
    recs, keep = program_records(nm.transpile(c, True).gates, n)
    rho = np.zeros(1 << (2 * n), dtype=np.complex128)
    rho[0] = 1
    threads = min(os.cpu_count() or 1, 16)  # the host-CPU share of one GPU
    k = min(sample_records, len(recs))
    t0 = time.perf_counter()
    cpu_baseline.apply_gates(rho, 2 * n, recs[:k], threads)
    t_rec = (time.perf_counter() - t0) / k
    del rho
    dense = {}
    for m in dense_sizes:
        cm = workloads.random_mix_circuit(m, 300, 5)
        noisy = nm.transpile(cm, True)
        t0 = time.perf_counter()
        odm.run_dense(noisy)
        dense[m] = time.perf_counter() - t0
    xs = np.array(list(dense.keys()), dtype=float)
    ys = np.log(np.array(list(dense.values())))
    slope, icpt = np.polyfit(xs, ys, 1)
    return {"superop_path_s_14q": t_rec * len(recs), "superop_path_threads": threads,
            "superop_path_sample": f"{k} of {len(recs)} records timed, scaled",
            "dense_reference_s": {str(m): v for m, v in dense.items()},
            "dense_reference_s_14q_extrapolated": float(np.exp(icpt + slope * n)),
            "dense_reference_growth_per_qubit": float(np.exp(slope)),
            "kind": "port", "note": "dense figure at 14 qubits is EXTRAPOLATED from 5..8 qubits (numpy, 1 process)"}


def run_unitary14(lib, ctx, reps=3, cpu=True):
    """14-qubit circuit unitary (2^28 complex128 entries) of the 300-gate mixed
    circuit: identity, fused gate passes on the column layout, U x, and the
    in-place transpose to the row-major matrix; each kernel timed with HIP
    events on the launch stream.  CPU leg: the restated reference algorithm
    (CircuitMatrix) at 10 qubits."""
    import numpy as np
    import torch
    from quict_amd import _native as nat
    from quict_amd import workloads
    from quict_amd.simulation.unitary import build_columns
    n = 14
    D = 1 << n
    c = workloads.dm14_circuit()
    stream = nat.stream_ptr(0)
    s = torch.cuda.current_stream()

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        fn()
        e1.record(s)
        e1.synchronize()
        return e0.elapsed_time(e1)

    W, passes = build_columns(c.gates, n, nat.QS_C128)
    x = torch.zeros(D, dtype=torch.complex128, device="cuda:0")
    x[0] = 1
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t_id, t_gates, t_mv, t_tr, t_all = [], [], [], [], []
    for _ in range(reps):
        t_id.append(timed(lambda: nat.check(lib.qs_mat_identity(ctx, ctypes.c_void_p(W.data_ptr()), n, nat.QS_C128,
                                                                stream))))
        profile_read(lib, ctx, reset=True)
        lib.qs_profile_enable(ctx, 1)
        t0 = time.perf_counter()
        W2, _ = build_columns(c.gates, n, nat.QS_C128)
        torch.cuda.synchronize()
        t_all.append((time.perf_counter() - t0) * 1e3)
        lib.qs_profile_enable(ctx, 0)
        ms, cnt = profile_read(lib, ctx, reset=True)
        t_gates.append(ms[0] + ms[1])
        t_mv.append(timed(lambda: nat.check(lib.qs_mat_vec(ctx, ctypes.c_void_p(W2.data_ptr()), n, nat.QS_C128, 1,
                                                           ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                                           stream))))
        t_tr.append(timed(lambda: nat.check(lib.qs_mat_transpose(ctx, ctypes.c_void_p(W2.data_ptr()), n, nat.QS_C128,
                                                                 stream))))
        del W2
    mat_bytes = D * D * 16
    med = lambda v: float(np.median(v))
    out = {"qubits": n, "gates": 300, "passes": passes, "matrix_GiB": mat_bytes / 2**30,
           "identity_ms": med(t_id), "identity_gbs": mat_bytes / (med(t_id) * 1e-3) / 1e9,
           "gate_passes_ms": med(t_gates), "gbs_per_pass": passes * 2 * mat_bytes / (med(t_gates) * 1e-3) / 1e9,
           "matvec_ms": med(t_mv), "matvec_gbs": mat_bytes / (med(t_mv) * 1e-3) / 1e9,
           "transpose_ms": med(t_tr), "transpose_gbs": 2 * mat_bytes / (med(t_tr) * 1e-3) / 1e9,
           "columns_wall_ms": med(t_all)}
    del W
    torch.cuda.empty_cache()
    if cpu:
        from oracle import unitary as ou
        c10 = workloads.random_mix_circuit(10, 300, 5)
        t0 = time.perf_counter()
        ou.circuit_matrix(c10.gates, 10)
        out["cpu_reference_algorithm_10q_ms"] = (time.perf_counter() - t0) * 1e3
        out["cpu_note"] = "restated CircuitMatrix (numpy) on a 300-gate 10-qubit circuit; 14 qubits is out of its reach"
    return out


def run_dense30(lib, ctx, n=30, reps=3):
    """Dense gates on 5 and 6 targets (a random Unitary and a QFT block) on a
    30-qubit complex128 state, one gate per pass: the matrix-core kernel
    (v_mfma_f64_16x16x4) against the per-amplitude kernel it replaces
    (QS_DENSE_MFMA=0), each timed with HIP events on the launch stream."""
    import numpy as np
    import torch
    from quict_amd.core import Circuit, QFT, Unitary
    from quict_amd.simulation.gate_compiler import compile_gates
    v = torch.zeros(1 << n, dtype=torch.complex128, device="cuda:0")
    v[0] = 1
    stream = nat_stream()
    out = {"qubits": n}
    rng = np.random.RandomState(3)
    for k, args in [(5, [3, 17, 9, 25, 0]), (6, [28, 4, 13, 21, 7, 1])]:
        for kind in ("unitary", "qft"):
            c = Circuit(n)
            if kind == "unitary":
                q, _ = np.linalg.qr(rng.randn(1 << k, 1 << k) + 1j * rng.randn(1 << k, 1 << k))
                Unitary(q) | c(args)
            else:
                QFT(k) | c(args)
            recs, keep = compile_gates(c.gates, n)
            for mode in ("mfma", "per_amplitude"):
                os.environ["QS_DENSE_MFMA"] = "1" if mode == "mfma" else "0"
                from quict_amd import _native as nat
                call = lambda: nat.check(lib.qs_sv_apply_gates(ctx, ctypes.c_void_p(v.data_ptr()), n, nat.QS_C128,
                                                                nat.gate_array_ptr(recs), len(recs), 0, None, stream),
                                         "qs_sv_apply_gates")
                call()
                torch.cuda.synchronize()
                profile_read(lib, ctx, reset=True)
                lib.qs_profile_enable(ctx, 1)
                for _ in range(reps):
                    call()
                torch.cuda.synchronize()
                lib.qs_profile_enable(ctx, 0)
                ms, cnt = profile_read(lib, ctx, reset=True)
                avg = ms[1] / max(cnt[1], 1)
                flops = 8.0 * (1 << k) * (1 << n)
                out[f"{kind}{k}_{mode}"] = {"avg_ms": avg, "touched_gbs": 2 * (1 << n) * 16 / (avg * 1e-3) / 1e9,
                                            "tflops": flops / (avg * 1e-3) / 1e12,
                                            "frac_fp64_peak": flops / (avg * 1e-3) / 78.6e12}
    os.environ.pop("QS_DENSE_MFMA", None)
    out["norm_after"] = float(torch.linalg.vector_norm(v).item())
    del v
    torch.cuda.empty_cache()
    return out


def nat_stream():
    from quict_amd import _native as nat
    return nat.stream_ptr(0)


def touched_bytes(circuit, n, amp_bytes):
    """Bytes the per-gate kernels read and write for the circuit: 2 * S per
    amplitude a gate changes (k_gate inserts control bits as fixed ones and
    phase / swap kernels visit only the amplitudes they change)."""
    from quict_amd import _native as nat
    from quict_amd.simulation.gate_compiler import compile_gates
    recs, keep = compile_gates(circuit.gates, n)
    total = 0
    for r in recs:
        c, t, kind = int(r["ncontrols"]), int(r["ntargets"]), int(r["kind"])
        frac = 2.0 ** -c
        if kind == nat.GATE_PHASE:
            frac = 2.0 ** -(c + t)
        elif kind == nat.GATE_SWAP:
            frac /= 2
        total += 2 * amp_bytes * frac * (1 << n)
    return total


def host_cpu():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def measured_traffic(n, ngates):
    """HBM bytes per tile-pass launch from the committed PMC summary (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated) when this run is the workload it was measured on."""
    path = os.path.join(ROOT, "profiles", "k_fused_sv30_c128_traffic.txt")
    if (n, ngates) != (30, 2000) or not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if "=" in line and not line.startswith("#"):
                k, v = line.strip().split("=", 1)
                vals[k] = int(v)
    return vals.get("traffic_bytes")


def cpu_sv20(c20, shot_sample=4):
    """CPU legs of the 20-qubit configuration: the 400 gates through the C
    restatement (one pass per gate) on 1 thread and on all cores, and the
    reference's per-shot sampling (copy the state, then measure and collapse
    qubit after qubit) on a few shots, scaled to 1024."""
    import numpy as np
    import random
    from quict_amd.simulation.gate_compiler import compile_gates
    from oracle import cpu_baseline
    n = 20
    recs, keep = compile_gates(c20.gates, n)
    out = {"kind": "port"}
    for threads in (1, min(os.cpu_count() or 1, 16)):
        v = np.zeros(1 << n, dtype=np.complex128)
        v[0] = 1
        t0 = time.perf_counter()
        cpu_baseline.apply_gates(v, n, recs, threads)
        dt = time.perf_counter() - t0
        out[f"gates_per_s_{threads}t"] = len(c20.gates) / dt
    random.seed(1001)
    u = np.array([random.random() for _ in range(shot_sample * n)], dtype=np.float64)
    t0 = time.perf_counter()
    cpu_baseline.sample(v, n, u, shot_sample)
    out["sample_1024_s"] = (time.perf_counter() - t0) / shot_sample * 1024
    out["sample"] = f"all 400 gates; {shot_sample} shots timed (1 thread) and scaled to 1024"
    return out


def cpu_baseline_sample(n_cpu=28, ngates=10, threads=1):
    import numpy as np
    from quict_amd import workloads
    from quict_amd.simulation.gate_compiler import compile_gates
    from oracle import cpu_baseline
    c = workloads.sv30_circuit(n=n_cpu)
    recs, keep = compile_gates(c.gates[:ngates], n_cpu)
    v = np.zeros(1 << n_cpu, dtype=np.complex128)
    v[0] = 1
    t0 = time.perf_counter()
    cpu_baseline.apply_gates(v, n_cpu, recs, threads)
    dt = time.perf_counter() - t0
    gbs = ngates * 2 * (1 << n_cpu) * 16 / dt / 1e9
    return gbs, ngates / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--gates", type=int, default=2000)
    ap.add_argument("--low-bits", type=int, default=0,
                    help="contiguous low index bits per fused tile (qs_set_fusion; 0 = built-in default 4)")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["dm14", "unitary14", "adjoint20", "qnn_c128", "qnn_c64", "dense30"],
                    help="run one secondary workload alone and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the simulator state lives on a single MI355X; "
                                   "sharding over several GPUs is designed but not built; run with --gpus 1"}))
        sys.exit(2)

    import torch
    from quict_amd import _native as nat, workloads

    lib = nat.load_library()
    ctx = nat.context(0)
    if args.low_bits:
        nat.check(lib.qs_set_fusion(ctx, 0, args.low_bits), "qs_set_fusion")
    if args.only:
        log(f"secondary workload {args.only} alone")
        only = {"dm14": lambda: run_dm14(lib, ctx, cpu=not args.no_cpu),
                "unitary14": lambda: run_unitary14(lib, ctx, cpu=not args.no_cpu),
                "adjoint20": lambda: run_adjoint_20q(cpu=not args.no_cpu),
                "qnn_c128": lambda: run_qnn_batch("double"),
                "qnn_c64": lambda: run_qnn_batch("single"),
                "dense30": lambda: run_dense30(lib, ctx)}[args.only]()
        print(json.dumps({"workload": args.only, **only}))
        return
    n = args.qubits
    log(f"building circuit n={n} gates={args.gates}")
    circuit = workloads.sv30_circuit(n=n, ngates=args.gates)
    state_bytes = (1 << n) * 16
    pass_bytes = 2 * state_bytes

    log("timed: complex128 fused")
    dt, passes, ms, cnt = run_sv(n, circuit, "double", True, args.steps, args.warmup, lib, ctx)
    step_ms = dt / args.steps * 1e3
    value = passes * pass_bytes / (dt / args.steps) / 1e9
    gates_per_s = args.gates / (dt / args.steps)
    fused_launches = cnt[0] // max(args.steps, 1)
    avg_fused_ms = ms[0] / max(cnt[0], 1)
    roof_achieved = pass_bytes / (avg_fused_ms * 1e-3) / 1e9 if cnt[0] else None
    log(f"c128 fused: {step_ms:.1f} ms/step, {passes} passes, {value:.0f} GB/s, avg tile-pass {avg_fused_ms:.3f} ms")

    secondary = {}
    log("streaming copy of one state (bandwidth reference; its profile calibrates the counters)")
    secondary["stream_copy"] = run_stream_copy(n, lib, ctx)
    if not args.no_secondary:
        log("secondary: complex64 fused")
        dt64, p64, ms64, cnt64 = run_sv(n, circuit, "single", True, 1, 1, lib, ctx)
        secondary["c64_fused"] = {
            "ms_per_step": dt64 * 1e3, "passes": p64, "gbs": p64 * 2 * (1 << n) * 8 / dt64 / 1e9,
            "gates_per_s": args.gates / dt64,
            "tile_pass_gbs": (2 * (1 << n) * 8) / (ms64[0] / max(cnt64[0], 1) * 1e-3) / 1e9 if cnt64[0] else None}
        log("secondary: complex128 unfused (one pass per gate)")
        dtu, pu, msu, cntu = run_sv(n, circuit, "double", False, 1, 0, lib, ctx)
        touched = touched_bytes(circuit, n, 16)
        secondary["c128_unfused"] = {
            "ms_per_step": dtu * 1e3, "passes": pu, "gates_per_s": args.gates / dtu,
            "touched_gbs": touched / dtu / 1e9,
            "touched_gbs_kernels": touched / (msu[1] * 1e-3) / 1e9 if cntu[1] else None,
            "touched_bytes_per_step": touched,
            "note": "one pass per gate; bytes are those each gate reads and writes (a control halves them, a "
                    "phase gate touches 1/2^(controls+targets) of the state, a swap 1/2)"}
        log("secondary: 20-qubit 400-gate circuit + 1024 shots")
        c20 = workloads.sv20_circuit()
        dt20, p20, _, _ = run_sv(20, c20, "double", True, 5, 2, lib, ctx)
        from quict_amd.simulation import StateVectorSimulator
        import random
        s20 = StateVectorSimulator("GPU", precision="double")
        s20.run(c20)
        random.seed(1001)
        t0 = time.perf_counter()
        s20.sample(1024)
        t_shots = time.perf_counter() - t0
        del s20
        s20 = StateVectorSimulator("GPU", precision="double")
        cap = s20.capture(c20)
        cap.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            cap.replay()
        torch.cuda.synchronize()
        t_rep = (time.perf_counter() - t0) / 20
        cap.close()
        del s20
        secondary["sv20_400"] = {"ms_per_step": dt20 / 5 * 1e3, "gates_per_s": 400 / (dt20 / 5), "passes": p20,
                                 "sample_1024_ms": t_shots * 1e3,
                                 "captured_replay_ms": t_rep * 1e3, "captured_gates_per_s": 400 / t_rep,
                                 "captured_note": "same circuit captured once (qs_graph_capture), each replay "
                                                  "re-runs it from |0...0> with no host planning"}
        if not args.no_cpu:
            log("secondary: 20-qubit circuit and shots on the CPU restatement")
            secondary["sv20_400"]["cpu_baseline"] = cpu_sv20(c20)
        log("secondary: adjoint gradient 20q x 12 layers vs parameter shift")
        secondary["adjoint_20q"] = run_adjoint_20q(cpu=not args.no_cpu)
        for prec in ("double", "single"):
            log(f"secondary: batched QNN {prec} (autotune + timed)")
            secondary[f"qnn_batch_{'c128' if prec == 'double' else 'c64'}"] = run_qnn_batch(prec)
        log("secondary: noisy density matrix 14q")
        secondary["dm14_noisy"] = run_dm14(lib, ctx, cpu=not args.no_cpu)
        log("secondary: circuit unitary 14q")
        secondary["unitary14"] = run_unitary14(lib, ctx, cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (C restatement of the reference CPU kernels, 1 thread)")
        gbs1, gps1, t1 = cpu_baseline_sample(28, 10, 1)
        cores = os.cpu_count() or 1
        gbsN, gpsN, tN = cpu_baseline_sample(28, 10, min(cores, 16))
        cpu = {"value": gbs1, "unit": "GB/s", "cores": 1, "kind": "port", "host": host_cpu(),
               "sample": "first 10 gates of the same 60/40 seeded generator at 28 qubits (4 GiB complex128), "
                         "one pass per gate as the reference's numba kernels; " + f"{t1:.1f} s",
               "gates_per_s": gps1,
               "all_cores": {"threads": min(cores, 16), "gbs": gbsN, "gates_per_s": gpsN,
                             "note": "threads capped at 16, the host-CPU share of one GPU on this machine"}}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"state-vector {n}q complex128, seeded random {args.gates}-gate circuit "
                               "(60% 1q / 40% 2q, seed 2), fused tile passes",
                   "qubits": n, "gates": args.gates, "seed": 2, "state_gib": state_bytes / 2 ** 30,
                   "tile_low_bits": args.low_bits or 4},
        "gates_per_s": gates_per_s,
        "passes_per_step": passes,
        "roofline": {"bound": "hbm", "achieved": roof_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (roof_achieved / HBM_PEAK_GBS) if roof_achieved else None,
                     "traffic": measured_traffic(n, args.gates),
                     "traffic_source": ("profiles/k_fused_sv30_c128_traffic.txt: rocprofv3 --pmc FETCH_SIZE and "
                                        "WRITE_SIZE runs of this workload, copy-calibrated; not measured in this run")
                     if measured_traffic(n, args.gates) else None,
                     "kernel": "k_fused<double,4,8>", "launches_per_step": fused_launches,
                     "avg_launch_ms": avg_fused_ms,
                     "bytes_per_launch": pass_bytes,
                     "frac_of_stream_copy": (roof_achieved / secondary["stream_copy"]["nontemporal"]["gbs"])
                     if roof_achieved and "stream_copy" in secondary else None},
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
