"""Loaders for the reference-generated fixtures in tests/golden/ (written by
tests/golden/make_reference_fixtures.py from the reference's own code) and
builders that turn their gate lists and noise settings into this package's
objects."""
import json
import os

import numpy as np

from quict_amd import workloads
from quict_amd.core import Circuit, GateType, build_gate
from quict_amd.core.noise import BitflipError, DampingError, DepolarizingError, NoiseModel, ReadoutError

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MIX = [t.name for t in workloads.RANDOM_GATE_MIX]


def gate_lists() -> dict:
    with open(os.path.join(GOLDEN, "gate_lists.json")) as f:
        return json.load(f)


def _npz(name):
    z = np.load(os.path.join(GOLDEN, name))  # allow_pickle stays False: plain arrays and a JSON string
    meta = json.loads(str(z["meta"]))
    return meta, {k: z[k] for k in z.files if k != "meta"}


def sv_cases():
    return _npz("sv_cpu.npz")


def dm_cases():
    return _npz("dm_cpu.npz")


def sv_samples() -> dict:
    with open(os.path.join(GOLDEN, "sv_sample.json")) as f:
        return json.load(f)


def simulator_results() -> dict:
    with open(os.path.join(GOLDEN, "simulator_result.json")) as f:
        return json.load(f)


def gate_list_of(c: Circuit) -> list:
    return [[g.type.name, [int(x) for x in g.cargs], [int(x) for x in g.targs], [float(p) for p in g.pargs]]
            for g in c.gates]


def circuit_from(n: int, gates: list) -> Circuit:
    c = Circuit(n)
    for name, cargs, targs, pargs in gates:
        g = build_gate(GateType[name], list(cargs) + list(targs), list(pargs) if pargs else None)
        c.append(g)
    return c


def counts_list(counts: dict, n: int) -> list:
    out = [0] * (1 << n)
    for k, v in counts.items():
        out[int(k)] = v
    return out


def config_noise(p_dep: float, gamma: float) -> NoiseModel:
    """Depolarizing(p_dep) after every one-qubit gate of the mix, amplitude
    damping (gamma) tensored with itself after every two-qubit gate."""
    nm = NoiseModel(basic_gates=list(MIX))
    nm.add_noise_for_all_qubits(DepolarizingError(p_dep), MIX[:10])
    d = DampingError(gamma, 0, 0)
    nm.add_noise_for_all_qubits(d.tensor(d), MIX[10:])
    return nm


def noise_model_of(meta: dict) -> NoiseModel:
    nm = config_noise(meta["p_dep"], meta["gamma"])
    if "bitflip" in meta:
        nm.add_noise_for_all_qubits(BitflipError(meta["bitflip"]), meta["bitflip_gates"])
    if "readout" in meta:
        q = meta["readout_qubits"]
        nm.add_readout_error(ReadoutError(np.array(meta["readout"])), q if len(q) > 1 else q[0])
    return nm
