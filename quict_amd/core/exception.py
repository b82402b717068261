"""Exceptions with QuICT's numeric error codes.

The codes follow QuICT/tools/exception/quict_exception.py:1-35 and
QuICT/tools/exception/simulation/simulation_exception.py (2001-2007) so that
callers that switch on ``ecode`` behave the same.  Native (C-ABI) status codes
are translated into these classes by ``quict_amd._native``.
"""


class QuICTException(Exception):
    def __init__(self, error_code: int = 1000, msg: str = None):
        self.ecode = error_code
        self.strerror = str(msg) if msg is not None else f"QuICT error {error_code}"
        super().__init__(self.strerror)

    def __str__(self):
        return self.strerror

    def __repr__(self):
        return f"{self.__class__.__name__}({self.strerror})"


def _mk(name, code, base=QuICTException):
    def __init__(self, *args):
        msg = " ".join(str(a) for a in args) if args else None
        QuICTException.__init__(self, code, msg)
    return type(name, (base,), {"__init__": __init__})


# core (1001-1020)
TypeError = _mk("TypeError", 1001)                    # noqa: A001  (mirrors reference naming)
ValueError = _mk("ValueError", 1002)                  # noqa: A001
IndexExceedError = _mk("IndexExceedError", 1003)
CircuitAppendError = _mk("CircuitAppendError", 1004)
CircuitSpecialAppendError = _mk("CircuitSpecialAppendError", 1006)
QASMError = _mk("QASMError", 1009)
QubitMeasureError = _mk("QubitMeasureError", 1010)
KrausError = _mk("KrausError", 1012)
NoiseApplyError = _mk("NoiseApplyError", 1013)
PauliNoiseUnmatchedError = _mk("PauliNoiseUnmatchedError", 1014)
DampingNoiseMixedProbExceedError = _mk("DampingNoiseMixedProbExceedError", 1015)
GateQubitAssignedError = _mk("GateQubitAssignedError", 1016)
GateParametersAssignedError = _mk("GateParametersAssignedError", 1017)
GateAppendError = _mk("GateAppendError", 1018)
GateMatrixError = _mk("GateMatrixError", 1019)

# simulation (2001-2007)
SampleBeforeRunError = _mk("SampleBeforeRunError", 2001)
GateTypeNotImplementError = _mk("GateTypeNotImplementError", 2002)
SimulationMatrixError = _mk("SimulationMatrixError", 2003)
StateVectorUnmatchedError = _mk("StateVectorUnmatchedError", 2004)
GateAlgorithmNotImplementError = _mk("GateAlgorithmNotImplementError", 2005)
UnitaryMatrixUnmatchedError = _mk("UnitaryMatrixUnmatchedError", 2006)
SimulatorOptionsUnmatchedError = _mk("SimulatorOptionsUnmatchedError", 2007)

# native runtime failures (device OOM, launch failure, bad arguments caught in C)
NativeRuntimeError = _mk("NativeRuntimeError", 2100)
