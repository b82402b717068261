"""Writes the reference-generated fixtures under tests/golden/ (data only).

Runs only in a container that holds the reference source tree, with conda's
Python 3.9 (the interpreter that has the reference's `ply` dependency):

    /opt/conda/bin/python3.9 tests/golden/make_reference_fixtures.py [--ref /root/reference]

It imports the reference package from its source tree and records what its own
code produces for seeded inputs.  Nothing here is imported by the tests or the
library: the tests read only the JSON / NPZ files it writes.

Four adjustments, all outside the reference's algorithms:
  * numba 0.54 in that interpreter is ABI-broken, so a stand-in `numba`
    module (written to a temporary directory, never to the repository) maps
    `njit` to the identity decorator and `prange` to `range`: the reference's
    CPU kernels then run as the plain Python they are written in.  One
    consequence: `random.random()` inside the CPU measure kernel
    (QuICT/ops/linalg/cpu_calculator.py:329-363) is CPython's generator, which
    is the one the reference's GPU path draws from
    (QuICT/ops/gate_kernel/gate_function.py:1795); the recorded seeded samples
    follow that rule.
  * `np.complex` (removed in numpy 1.24) is aliased to `complex` before import.
  * that interpreter has no libuuid binding, so `uuid.uuid1()` (called for
    every Circuit / Qubit, QuICT/core/utils/id_generator.py) falls back to
    drawing its clock sequence with `random.getrandbits(14)`; CPython with
    libuuid (as on the GPU host's Python 3.10) draws nothing.  The script
    gives `uuid` a time-UUID source that draws nothing from `random`, so the
    seeded streams do not depend on which interpreter ran the reference (this
    matters where Circuits are built between draws: the trajectory-mode
    sampler transpiles a new Circuit per shot).
  * no bytecode is written (`sys.dont_write_bytecode`): the reference tree
    stays as it is.

What is recorded (each item is something no golden vector in the reference
pins):
  gate_lists.json       Circuit.random_append (QuICT/core/circuit/circuit.py:442-488,
                        gate_builder.py:107-131) gate lists for the seeded
                        workloads: type, cargs, targs, pargs.
  sv_cpu.npz            StateVectorSimulator(device="CPU").run amplitudes
                        (statevector_simulator.py:131-220) of seeded circuits,
                        incl. gate types absent from the golden QASM circuit
                        (rz, sx, sy, sw, phase, gphase, fsim, ccz, ccrz, ...),
                        and the documented quirk gates (rzx, iswap family).
  sv_sample.json        StateVectorSimulator.sample (statevector_simulator.py:595-626)
                        seeded counts.
  dm_cpu.npz            DensityMatrixSimulator(device="CPU") (density_matrix_simulator.py:88-240)
                        with NoiseModel.transpile (noise_model.py:152-185):
                        accumulated mode, trajectory mode, readout errors and
                        seeded samples.
  simulator_result.json Simulator(...).run(circuit, shots) result dictionaries
                        (simulator.py:101-151, utils/result.py:43-104).
"""
import argparse
import json
import os
import random
import sys
import tempfile

sys.dont_write_bytecode = True

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

MIX = ["h", "x", "y", "z", "s", "t", "rx", "ry", "rz", "u3", "cx", "cz", "crz", "swap"]
PROBS_60_40 = [0.06] * 10 + [0.10] * 4
# every gate type random_append accepts except the quirk gates (recorded separately)
ALL_NO_QUIRK = ["h", "hy", "s", "sdg", "x", "y", "z", "sx", "sy", "sw", "id", "u1", "u2", "u3", "rx", "ry", "rz",
                "t", "tdg", "phase", "gphase", "cz", "cx", "cy", "ch", "crz", "cu1", "cu3", "fsim", "rxx", "ryy",
                "rzz", "swap", "cswap", "ccx", "ccz", "ccrz"]
RZ_HEAVY = ["rz", "crz", "h", "rx", "ry", "u3", "cx"]
QUIRK = ["rzx", "iswap", "iswapdg", "sqiswap"]


def import_reference(ref_root):
    shim = tempfile.mkdtemp(prefix="numba_identity_")
    os.makedirs(os.path.join(shim, "numba"))
    with open(os.path.join(shim, "numba", "__init__.py"), "w") as f:
        f.write("def njit(*a, **k):\n"
                "    if len(a) == 1 and callable(a[0]) and not k:\n"
                "        return a[0]\n"
                "    return lambda f: f\n"
                "jit = njit\n"
                "prange = range\n")
    sys.path.insert(0, shim)
    sys.path.insert(0, ref_root)
    np.complex = complex
    import uuid
    if getattr(uuid, "_generate_time_safe", None) is None:
        uuid._generate_time_safe = lambda: (uuid.uuid4().bytes, None)
    import QuICT  # noqa: F401
    return QuICT


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    import_reference(args.ref)
    from QuICT.core import Circuit
    from QuICT.core.gate import GateType
    from QuICT.simulation.state_vector import StateVectorSimulator
    from QuICT.simulation.density_matrix import DensityMatrixSimulator
    from QuICT.simulation.simulator import Simulator
    from QuICT.core.noise import NoiseModel, DepolarizingError, DampingError, ReadoutError, BitflipError

    def types(names):
        return [GateType[x] for x in names]

    def gen(n, ngates, seed, names, probs=None):
        c = Circuit(n)  # seeds are set after the Circuit exists (its uuid may draw from `random`)
        random.seed(seed)
        np.random.seed(seed)
        c.random_append(ngates, typelist=types(names), random_params=True, probabilities=probs)
        return c

    def gate_list(c):
        return [[g.type.name, [int(x) for x in g.cargs], [int(x) for x in g.targs], [float(p) for p in g.pargs]]
                for g in c.gates]

    # ---- gate lists of the seeded workloads ---------------------------------
    specs = {
        "sv20": dict(n=20, ngates=400, seed=1, names=MIX, probs=None),
        "sv30": dict(n=30, ngates=2000, seed=2, names=MIX, probs=PROBS_60_40),
        "dm14": dict(n=14, ngates=300, seed=5, names=MIX, probs=None),
        "all_types_12": dict(n=12, ngates=300, seed=9, names=ALL_NO_QUIRK + QUIRK, probs=None),
    }
    lists = {}
    for name, s in specs.items():
        c = gen(s["n"], s["ngates"], s["seed"], s["names"], s["probs"])
        lists[name] = dict(n=s["n"], ngates=s["ngates"], seed=s["seed"], typelist=s["names"], probabilities=s["probs"],
                           gates=gate_list(c))
    with open(os.path.join(HERE, "gate_lists.json"), "w") as f:
        json.dump(lists, f, separators=(",", ":"))
    print("gate_lists.json:", {k: len(v["gates"]) for k, v in lists.items()})

    # ---- state-vector amplitudes ------------------------------------------------
    sv_cases = {
        "mix10": dict(n=10, ngates=200, seed=21, names=MIX, probs=None, psi0=None),
        "mix12_6040": dict(n=12, ngates=300, seed=22, names=MIX, probs=PROBS_60_40, psi0=None),
        "mix13_6040": dict(n=13, ngates=300, seed=28, names=MIX, probs=PROBS_60_40, psi0=None),
        "rz_heavy9": dict(n=9, ngates=150, seed=23, names=RZ_HEAVY, probs=None, psi0=None),
        "all8": dict(n=8, ngates=160, seed=24, names=ALL_NO_QUIRK, probs=None, psi0=None),
        "all8_psi0": dict(n=8, ngates=160, seed=25, names=ALL_NO_QUIRK, probs=None, psi0=26),
    }
    for q in QUIRK:
        sv_cases[f"quirk_{q}"] = dict(n=5, ngates=6, seed=30 + QUIRK.index(q), names=["h", "rx", q], probs=None,
                                      psi0=40 + QUIRK.index(q))
    npz = {}
    meta = {}
    for name, s in sv_cases.items():
        c = gen(s["n"], s["ngates"], s["seed"], s["names"], s["probs"])
        psi0 = None
        if s["psi0"] is not None:
            rng = np.random.RandomState(s["psi0"])
            psi0 = rng.randn(1 << s["n"]) + 1j * rng.randn(1 << s["n"])
            psi0 /= np.linalg.norm(psi0)
            npz[name + "__psi0"] = psi0
        sim = StateVectorSimulator(device="CPU", precision="double")
        out = sim.run(c, state_vector=None if psi0 is None else psi0.copy())
        npz[name] = np.asarray(out, dtype=np.complex128)
        meta[name] = dict(n=s["n"], gates=gate_list(c), has_psi0=psi0 is not None)
    np.savez(os.path.join(HERE, "sv_cpu.npz"), meta=json.dumps(meta), **npz)
    print("sv_cpu.npz:", sorted(npz))

    # ---- seeded sampling ----------------------------------------------------------
    samples = {}
    for name, n, ngates, seed, shots, sseed in [("mix10", 10, 200, 21, 1024, 1001), ("mix8", 8, 120, 27, 3000, 77)]:
        c = gen(n, ngates, seed, MIX)
        sim = StateVectorSimulator(device="CPU", precision="double")
        sim.run(c)
        random.seed(sseed)
        counts = sim.sample(shots)
        samples[name] = dict(n=n, ngates=ngates, seed=seed, shots=shots, sample_seed=sseed, gates=gate_list(c),
                             counts={str(i): int(v) for i, v in enumerate(counts) if v})
    with open(os.path.join(HERE, "sv_sample.json"), "w") as f:
        json.dump(samples, f, separators=(",", ":"))
    print("sv_sample.json:", {k: v["shots"] for k, v in samples.items()})

    # ---- density matrix with noise ---------------------------------------------------
    def config_noise(p_dep, gamma):
        nm = NoiseModel(basic_gates=list(MIX))
        nm.add_noise_for_all_qubits(DepolarizingError(p_dep), MIX[:10])
        d = DampingError(gamma, 0, 0)
        nm.add_noise_for_all_qubits(d.tensor(d), MIX[10:])
        return nm

    dm = {}
    dmeta = {}
    # accumulated mode: rho <- sum K rho K^dagger after every gate (config strengths and strong noise)
    for name, n, ngates, seed, p_dep, gamma in [("acc5_config", 5, 60, 5, 1e-3, 1e-2),
                                                 ("acc4_strong", 4, 40, 6, 0.2, 0.3)]:
        c = gen(n, ngates, seed, MIX)
        sim = DensityMatrixSimulator(device="CPU", precision="double", accumulated_mode=True)
        rho = sim.run(c, noise_model=config_noise(p_dep, gamma))
        dm[name] = np.asarray(rho, dtype=np.complex128)
        dmeta[name] = dict(n=n, gates=gate_list(c), p_dep=p_dep, gamma=gamma, accumulated=True)
    # trajectory mode: Pauli channels drawn per gate from numpy's global generator,
    # damping as Kraus; then seeded sampling with a readout error
    for name, n, ngates, seed, p_dep, gamma, shots in [("traj4", 4, 30, 7, 0.3, 0.25, 200)]:
        c = gen(n, ngates, seed, MIX)
        nm = config_noise(p_dep, gamma)
        nm.add_noise_for_all_qubits(BitflipError(0.2), ["x", "y"])
        nm.add_readout_error(ReadoutError(np.array([[0.8, 0.2], [0.1, 0.9]])), [0, 2])
        np.random.seed(100)
        random.seed(101)
        sim = DensityMatrixSimulator(device="CPU", precision="double", accumulated_mode=False)
        rho = sim.run(c, noise_model=nm)
        dm[name] = np.asarray(rho, dtype=np.complex128)
        counts = sim.sample(shots)
        dmeta[name] = dict(n=n, gates=gate_list(c), p_dep=p_dep, gamma=gamma, accumulated=False, np_seed=100,
                           py_seed=101, shots=shots, bitflip=0.2, bitflip_gates=["x", "y"],
                           readout=[[0.8, 0.2], [0.1, 0.9]], readout_qubits=[0, 2],
                           counts={str(i): int(v) for i, v in enumerate(counts) if v})
    # accumulated mode sampling with a readout error (rho restored per shot)
    for name, n, ngates, seed, shots in [("acc4_readout", 4, 30, 8, 300)]:
        c = gen(n, ngates, seed, MIX)
        nm = config_noise(0.1, 0.2)
        nm.add_readout_error(ReadoutError(np.array([[0.9, 0.1], [0.25, 0.75]])), 1)
        sim = DensityMatrixSimulator(device="CPU", precision="double", accumulated_mode=True)
        rho = sim.run(c, noise_model=nm)
        np.random.seed(200)
        random.seed(201)
        counts = sim.sample(shots)
        dm[name] = np.asarray(rho, dtype=np.complex128)
        dmeta[name] = dict(n=n, gates=gate_list(c), p_dep=0.1, gamma=0.2, accumulated=True, np_seed=200, py_seed=201,
                           shots=shots, readout=[[0.9, 0.1], [0.25, 0.75]], readout_qubits=[1],
                           counts={str(i): int(v) for i, v in enumerate(counts) if v})
    np.savez(os.path.join(HERE, "dm_cpu.npz"), meta=json.dumps(dmeta), **dm)
    print("dm_cpu.npz:", sorted(dm))

    # ---- Simulator result dictionaries -------------------------------------------------
    results = {}
    c = gen(6, 50, 31, MIX)
    random.seed(300)
    r = Simulator(device="CPU", backend="state_vector").run(c, shots=500)
    results["state_vector"] = dict(n=6, gates=gate_list(c), py_seed=300, shots_arg=500, result=strip(r))
    c = gen(4, 30, 32, MIX)
    np.random.seed(400)
    random.seed(401)
    r = Simulator(device="CPU", backend="density_matrix").run(c, shots=120, noise_model=config_noise(0.2, 0.3))
    results["density_matrix"] = dict(n=4, gates=gate_list(c), np_seed=400, py_seed=401, shots_arg=120,
                                     p_dep=0.2, gamma=0.3, result=strip(r))
    with open(os.path.join(HERE, "simulator_result.json"), "w") as f:
        json.dump(results, f, separators=(",", ":"))
    print("simulator_result.json:", list(results))


def strip(r):
    """The result dictionary with arrays as [re, im] lists."""
    def enc(a):
        if a is None:
            return None
        a = np.asarray(a)
        return {"shape": list(a.shape), "re": a.real.ravel().tolist(), "im": a.imag.ravel().tolist()}
    return {"keys": sorted(r.keys()), "shots": r["shots"], "backend_info": r["backend_info"],
            "data_keys": sorted(r["data"].keys()), "counts": r["data"]["counts"],
            "state_vector": enc(r["data"]["state_vector"]), "density_matrix": enc(r["data"]["density_matrix"])}


if __name__ == "__main__":
    main()
