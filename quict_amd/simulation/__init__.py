"""Simulators of the MI355X circuit-simulation path (QuICT.simulation API)."""
from .state_vector import StateVectorSimulator
from .density_matrix import DensityMatrixSimulator
from .unitary import UnitarySimulator, circuit_matrix
from .simulator import Simulator, Result
from .device_array import DeviceArray
