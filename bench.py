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
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the simulator state lives on a single MI355X; "
                                   "multi-GPU sharding is designed (DESIGN.md) but not built; run with --gpus 1"}))
        sys.exit(2)

    import torch
    from quict_amd import _native as nat, workloads

    lib = nat.load_library()
    ctx = nat.context(0)
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
    if not args.no_secondary:
        log("secondary: complex64 fused")
        dt64, p64, ms64, cnt64 = run_sv(n, circuit, "single", True, 1, 1, lib, ctx)
        secondary["c64_fused"] = {
            "ms_per_step": dt64 * 1e3, "passes": p64, "gbs": p64 * 2 * (1 << n) * 8 / dt64 / 1e9,
            "gates_per_s": args.gates / dt64,
            "tile_pass_gbs": (2 * (1 << n) * 8) / (ms64[0] / max(cnt64[0], 1) * 1e-3) / 1e9 if cnt64[0] else None}
        log("secondary: complex128 unfused (one pass per gate)")
        dtu, pu, msu, cntu = run_sv(n, circuit, "double", False, 1, 0, lib, ctx)
        secondary["c128_unfused"] = {
            "ms_per_step": dtu * 1e3, "passes": pu, "gbs": pu * pass_bytes / dtu / 1e9,
            "gates_per_s": args.gates / dtu,
            "gate_pass_gbs": pass_bytes / (msu[1] / max(cntu[1], 1) * 1e-3) / 1e9 if cntu[1] else None,
            "note": "gate passes touch only the amplitudes a gate changes (controls / phases touch 1/2 or 1/4)"}
        log("secondary: 20-qubit 400-gate circuit")
        c20 = workloads.sv20_circuit()
        dt20, p20, _, _ = run_sv(20, c20, "double", True, 5, 2, lib, ctx)
        secondary["sv20_400"] = {"ms_per_step": dt20 / 5 * 1e3, "gates_per_s": 400 / (dt20 / 5), "passes": p20}

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (C restatement of the reference CPU kernels, 1 thread)")
        gbs1, gps1, t1 = cpu_baseline_sample(28, 10, 1)
        cores = os.cpu_count() or 1
        gbsN, gpsN, tN = cpu_baseline_sample(28, 10, min(cores, 16))
        cpu = {"value": gbs1, "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": "first 10 gates of the same 60/40 seeded generator at 28 qubits (4 GiB complex128), "
                         "one pass per gate as the reference's numba kernels; " + f"{t1:.1f} s",
               "gates_per_s": gps1, "all_cores": {"threads": min(cores, 16), "gbs": gbsN, "gates_per_s": gpsN}}

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
                   "qubits": n, "gates": args.gates, "seed": 2, "state_gib": state_bytes / 2 ** 30},
        "gates_per_s": gates_per_s,
        "passes_per_step": passes,
        "roofline": {"bound": "hbm", "achieved": roof_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (roof_achieved / HBM_PEAK_GBS) if roof_achieved else None, "traffic": None,
                     "kernel": "k_fused<double,4,8>", "launches_per_step": fused_launches,
                     "avg_launch_ms": avg_fused_ms,
                     "bytes_per_launch": pass_bytes},
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
