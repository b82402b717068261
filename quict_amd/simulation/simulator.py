"""``Simulator`` façade and ``Result`` record.

Same constructor / ``run`` contract as QuICT/simulation/simulator.py:19-151
(device, backend, precision, circuit_record, amplitude_record, output_path,
per-backend options) and the result dictionary of
QuICT/simulation/utils/result.py:5-104:
``{"id", "shots", "backend_info": {"device", "backend", "options"},
   "data": {"counts", "state_vector", "density_matrix"}}``.
Backends: ``unitary``, ``state_vector`` and ``density_matrix``.
"""
import os

import numpy as np

from ..core import Circuit
from ..core import exception as qerr
from .state_vector import StateVectorSimulator
from .density_matrix import DensityMatrixSimulator
from .unitary import UnitarySimulator


class Result:
    def __init__(self, device, backend, precision, circuit_record, amplitude_record, options, output_path=None):
        self.device = device
        self.backend = backend
        self.precision = precision
        self._circuit_record = circuit_record
        self._amplitude_record = amplitude_record
        self.options = options
        self.id = None
        self.shots = 0
        self.counts = {}
        self.state_vector = None
        self.density_matrix = None
        self._dump = amplitude_record or circuit_record
        if self._dump:
            self._output_path = output_path or os.path.join(os.getcwd(), "output")
            os.makedirs(self._output_path, exist_ok=True)

    def __str__(self):
        return f"Device: {self.device}\nBackend: {self.backend}\nOptions: {self.options}\nResults: {self.counts}"

    def __dict__(self):
        return {
            "id": self.id,
            "shots": self.shots,
            "backend_info": {"device": self.device, "backend": self.backend, "options": self.options},
            "data": {"counts": self.counts, "state_vector": self.state_vector, "density_matrix": self.density_matrix},
        }

    def record_circuit(self, circuit: Circuit):
        self.id = circuit.name
        if self._circuit_record:
            with open(os.path.join(self._output_path, f"circuit_{self.id}.qasm"), "w") as f:
                f.write(circuit.qasm())

    def record_amplitude(self, amplitude):
        a = amplitude.get() if hasattr(amplitude, "get") else np.asarray(amplitude)
        if self._amplitude_record:
            np.savetxt(os.path.join(self._output_path, f"amp_{self.id}.txt"), a.reshape(-1) if a.ndim > 1 else a)
        if self.backend == "density_matrix":
            self.density_matrix = a.copy()
        else:
            self.state_vector = a.copy()

    def record_sample(self, result):
        if isinstance(result, dict):
            width = self._width
            self.shots = int(sum(result.values()))
            self.counts = {format(k, "b").zfill(width): v for k, v in sorted(result.items())}
        else:
            self.shots = int(sum(result))
            width = int(np.log2(len(result)))
            self.counts = {format(i, "b").zfill(width): int(c) for i, c in enumerate(result)}
        if self._dump:
            with open(os.path.join(self._output_path, f"result_{self.id}.log"), "w") as f:
                f.write(str(self.__dict__()))


class Simulator:
    __DEVICE = ["CPU", "GPU"]
    __BACKEND = ["unitary", "state_vector", "density_matrix"]
    __PRECISION = ["single", "double"]
    __OPTIONS = {"state_vector": ["gpu_device_id", "matrix_aggregation", "fusion"],
                 "density_matrix": ["accumulated_mode", "gpu_device_id", "fusion"],
                 "unitary": ["gpu_device_id", "fusion"]}

    def __init__(self, device: str = "GPU", backend: str = "state_vector", precision: str = "double",
                 circuit_record: bool = False, amplitude_record: bool = False, output_path: str = None, **options):
        if device != "GPU":
            raise qerr.ValueError("Simulator.device must be 'GPU' in this build, got", device)
        if backend not in self.__BACKEND:
            raise qerr.ValueError("Simulator.backend must be unitary / state_vector / density_matrix, got", backend)
        if precision not in self.__PRECISION:
            raise qerr.ValueError("Simulator.precision must be single / double, got", precision)
        for k in options:
            if k not in self.__OPTIONS[backend]:
                raise qerr.SimulatorOptionsUnmatchedError(f"option {k} does not apply to {device}/{backend}")
        self._device, self._backend, self._precision = device, backend, precision
        self._options = options
        if backend == "state_vector":
            self._simulator = StateVectorSimulator(device, precision=precision, **options)
        elif backend == "unitary":
            self._simulator = UnitarySimulator(device, precision=precision, **options)
        else:
            self._simulator = DensityMatrixSimulator(device, precision=precision, **options)
        self._result = Result(device, backend, precision, circuit_record, amplitude_record, options, output_path)

    def run(self, circuit: Circuit, shots: int = 1, state_vector=None, density_matrix=None, noise_model=None,
            use_previous: bool = False) -> dict:
        if not isinstance(circuit, Circuit) and self._backend != "unitary":
            raise qerr.SimulatorOptionsUnmatchedError(
                f"The unitary matrix input only allows in the unitary backend, not {self._backend}.")
        if (density_matrix is not None or noise_model is not None) and self._backend != "density_matrix":
            raise qerr.SimulatorOptionsUnmatchedError("density_matrix / noise_model need the density_matrix backend")
        if state_vector is not None and self._backend == "density_matrix":
            raise qerr.SimulatorOptionsUnmatchedError("state_vector input is not allowed in the density_matrix backend")
        if isinstance(circuit, Circuit):
            self._result.record_circuit(circuit)
            self._result._width = circuit.width()
        else:
            self._result._width = int(np.shape(circuit)[0]).bit_length() - 1
        if self._backend == "density_matrix":
            amp = self._simulator.run(circuit, density_matrix, noise_model, use_previous)
        else:
            amp = self._simulator.run(circuit, state_vector, use_previous)
        self._result.record_amplitude(amp)
        self._result.record_sample(self._simulator.sample(shots))
        return self._result.__dict__()
