"""The HIP path against what the reference's own code produced
(tests/golden/*, written by tests/golden/make_reference_fixtures.py):
StateVectorSimulator amplitudes and seeded samples, DensityMatrixSimulator with
noise (accumulated and trajectory modes, readout errors, seeded samples) and
the Simulator result dictionaries.  complex128 at 1e-12; the complex64 error
against the same fixture is printed and bounded."""
import random

import numpy as np
import pytest

import reference_fixtures as rf

pytestmark = pytest.mark.gpu

SV_CASES = ["mix10", "mix12_6040", "mix13_6040", "rz_heavy9", "all8", "all8_psi0"]


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("fusion", [True, False], ids=["fused", "per-gate"])
@pytest.mark.parametrize("name", SV_CASES)
def test_statevector_matches_the_reference_cpu_simulator(name, fusion):
    say("reference sv fixture", name, "fused" if fusion else "per-gate")
    from quict_amd.simulation import StateVectorSimulator
    meta, arrays = rf.sv_cases()
    m = meta[name]
    psi0 = arrays.get(name + "__psi0")
    ref = arrays[name]
    out = StateVectorSimulator("GPU", precision="double", fusion=fusion).run(
        rf.circuit_from(m["n"], m["gates"]), state_vector=psi0).get()
    assert np.max(np.abs(out - ref)) < 1e-12
    out64 = StateVectorSimulator("GPU", precision="single", fusion=fusion).run(
        rf.circuit_from(m["n"], m["gates"]), state_vector=psi0).get()
    err = float(np.max(np.abs(out64 - ref)))
    print(f"  complex64 max-abs error vs the reference: {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("name", ["mix10", "mix8"])
def test_seeded_sampling_matches_the_reference(name):
    say("reference sample fixture", name)
    from quict_amd.simulation import StateVectorSimulator
    s = rf.sv_samples()[name]
    sim = StateVectorSimulator("GPU", precision="double")
    sim.run(rf.circuit_from(s["n"], s["gates"]))
    random.seed(s["sample_seed"])
    assert sim.sample(s["shots"]) == rf.counts_list(s["counts"], s["n"])


@pytest.mark.parametrize("name", ["acc5_config", "acc4_strong", "acc4_readout"])
def test_accumulated_noise_matches_the_reference(name):
    say("reference dm fixture", name)
    from quict_amd.simulation import DensityMatrixSimulator
    meta, arrays = rf.dm_cases()
    m = meta[name]
    sim = DensityMatrixSimulator("GPU", precision="double", accumulated_mode=True)
    rho = sim.run(rf.circuit_from(m["n"], m["gates"]), noise_model=rf.noise_model_of(m)).get()
    assert np.max(np.abs(rho - arrays[name])) < 1e-12
    if "counts" in m:
        np.random.seed(m["np_seed"])
        random.seed(m["py_seed"])
        assert sim.sample(m["shots"]) == rf.counts_list(m["counts"], m["n"])


def test_trajectory_noise_and_readout_sampling_match_the_reference():
    say("reference dm trajectory fixture")
    from quict_amd.simulation import DensityMatrixSimulator
    meta, arrays = rf.dm_cases()
    m = meta["traj4"]
    np.random.seed(m["np_seed"])
    random.seed(m["py_seed"])
    sim = DensityMatrixSimulator("GPU", precision="double", accumulated_mode=False)
    rho = sim.run(rf.circuit_from(m["n"], m["gates"]), noise_model=rf.noise_model_of(m)).get()
    assert np.max(np.abs(rho - arrays["traj4"])) < 1e-12
    assert sim.sample(m["shots"]) == rf.counts_list(m["counts"], m["n"])


def _check_result(res, ref, backend):
    assert sorted(res.keys()) == ref["keys"]
    assert sorted(res["data"].keys()) == ref["data_keys"]
    assert res["shots"] == ref["shots"]
    assert res["backend_info"]["backend"] == backend
    assert res["backend_info"]["options"] == ref["backend_info"]["options"]
    assert res["data"]["counts"] == ref["counts"]
    for key in ("state_vector", "density_matrix"):
        enc = ref[key]
        if enc is None:
            assert res["data"][key] is None
            continue
        a = np.asarray(enc["re"]) + 1j * np.asarray(enc["im"])
        assert list(np.shape(res["data"][key])) == enc["shape"]
        assert np.max(np.abs(np.asarray(res["data"][key]).ravel() - a)) < 1e-12


def test_simulator_state_vector_result_matches_the_reference():
    say("reference Simulator result, state_vector")
    from quict_amd.simulation import Simulator
    r = rf.simulator_results()["state_vector"]
    random.seed(r["py_seed"])
    res = Simulator(device="GPU", backend="state_vector").run(rf.circuit_from(r["n"], r["gates"]),
                                                             shots=r["shots_arg"])
    _check_result(res, r["result"], "state_vector")


def test_simulator_density_matrix_result_matches_the_reference():
    """Density-matrix backend with a noise model and no options: trajectory
    mode (per-shot transpile and rerun), as in the reference."""
    say("reference Simulator result, density_matrix")
    from quict_amd.simulation import Simulator
    r = rf.simulator_results()["density_matrix"]
    np.random.seed(r["np_seed"])
    random.seed(r["py_seed"])
    res = Simulator(device="GPU", backend="density_matrix").run(
        rf.circuit_from(r["n"], r["gates"]), shots=r["shots_arg"], noise_model=rf.config_noise(r["p_dep"], r["gamma"]))
    _check_result(res, r["result"], "density_matrix")

