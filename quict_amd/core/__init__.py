"""Front-end objects of the simulation path (QuICT.core API mirror)."""
from .gate_type import GateType, MatrixType
from .gate import *  # noqa: F401,F403
from .gate import BasicGate, GATE_TYPE_TO_CLASS, build_gate, build_random_gate
from .composite_gate import CompositeGate
from .qubit import Qubit, Qureg
from .circuit import Circuit
from .operator import NoiseGate, Trigger, Operator
from . import noise
from .utils import matrix_product_to_circuit
