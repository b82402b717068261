"""Batched QNN: many uneven-width parameterised circuits, forward + adjoint
backward in one persistent, graph-captured launch (qs_qnn_batch).

The circuit family (angle encoding, hardware-efficient ansatz, <Z_0> readout,
squared loss) is documented in quict_amd/csrc/qsim_qnn.hip.  ``QNNBatch``
mirrors a torch loss module: ``forward_backward(theta)`` returns the mean loss,
per-circuit <Z_0>, d loss / d theta and d loss / d x.

``autotune`` times every launch shape of a small search space on the real
batch (LDS width limit, workgroup size, workgroups per CU) and caches the
winner per (dtype, width histogram, layers, features) in a JSON file.
"""
import ctypes
import json
import os
import time
from typing import Dict, Optional, Tuple

import numpy as np

from ..core import GateType
from ..core import exception as qerr
from .. import _native as nat
from .adjoint import ParameterizedCircuit

CACHE_PATH = os.environ.get("QUICT_AMD_TUNE_CACHE",
                            os.path.join(os.path.expanduser("~"), ".cache", "quict_amd", "autotune.json"))

DEFAULT_CONFIG = {"lds_max_width": 11, "block_threads": 256, "blocks_per_cu": 2, "use_graph": 1, "fuse_qubits": 2}


def circuit_for(width: int, x_row: np.ndarray, layers: int, max_width: int) -> ParameterizedCircuit:
    """The same circuit as one batch entry, as a ParameterizedCircuit (parameter k =
    (l*2 + g)*max_width + q), for reference computations."""
    pc = ParameterizedCircuit(width)
    for j, xv in enumerate(x_row):
        pc.add(GateType.ry, [j % width], value=float(xv))
    for l in range(layers):
        for q in range(width):
            pc.add(GateType.ry, [q], param=(l * 2 + 0) * max_width + q)
            pc.add(GateType.rz, [q], param=(l * 2 + 1) * max_width + q)
        for q in range(width - 1):
            pc.add(GateType.cx, [q, q + 1])
    pc.nparams = layers * 2 * max_width
    return pc


def synthetic_batch(ncirc: int = 1024, nfeat: int = 16, wmin: int = 8, wmax: int = 18, seed: int = 4):
    """Widths U{wmin..wmax}, x ~ N(0, I), labels y = [x_0 > 0], targets 1 - 2y."""
    rng = np.random.RandomState(seed)
    widths = rng.randint(wmin, wmax + 1, size=ncirc).astype(np.int32)
    x = rng.randn(ncirc, nfeat)
    y = (x[:, 0] > 0).astype(np.float64)
    return widths, x, 1.0 - 2.0 * y


class QNNBatch:
    def __init__(self, widths, x, targets, layers: int = 6, max_width: Optional[int] = None,
                 precision: str = "double", device: int = 0, config: Optional[Dict] = None):
        self.widths = np.ascontiguousarray(widths, dtype=np.int32)
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.targets = np.ascontiguousarray(targets, dtype=np.float64)
        self.ncirc, self.nfeat = self.x.shape
        if len(self.widths) != self.ncirc or len(self.targets) != self.ncirc:
            raise qerr.ValueError("widths, x and targets must agree on the batch size")
        self.layers = layers
        self.max_width = int(max_width or self.widths.max())
        self.precision = precision
        self.dcode = nat.QS_C128 if precision == "double" else nat.QS_C64
        self.device = device
        self.lib = nat.load_library()
        self.ctx = nat.context(device)
        self.config = dict(config or DEFAULT_CONFIG)
        self.last_ms = None

    @property
    def nparams(self) -> int:
        return self.layers * 2 * self.max_width

    def _cfg(self, cfg: Dict) -> nat.QsQnnConfig:
        c = nat.QsQnnConfig()
        c.lds_max_width = cfg["lds_max_width"]
        c.block_threads = cfg["block_threads"]
        c.blocks_per_cu = cfg["blocks_per_cu"]
        c.use_graph = cfg.get("use_graph", 1)
        c.fuse_qubits = cfg.get("fuse_qubits", 1)
        return c

    def forward_backward(self, theta: np.ndarray, config: Optional[Dict] = None
                         ) -> Tuple[float, np.ndarray, np.ndarray, np.ndarray]:
        theta = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
        if theta.size != self.nparams:
            raise qerr.ValueError(f"theta needs {self.nparams} values")
        ev = np.zeros(self.ncirc)
        loss = ctypes.c_double(0)
        gth = np.zeros(self.nparams)
        gx = np.zeros((self.ncirc, self.nfeat))
        cfg = self._cfg(config or self.config)
        P = ctypes.c_void_p
        t0 = time.perf_counter()
        rc = self.lib.qs_qnn_batch(self.ctx, self.dcode, self.ncirc, self.widths.ctypes.data_as(P), self.nfeat,
                                   self.x.ctypes.data_as(P), self.targets.ctypes.data_as(P), self.layers,
                                   self.max_width, theta.ctypes.data_as(P), ev.ctypes.data_as(P),
                                   ctypes.byref(loss), gth.ctypes.data_as(P), gx.ctypes.data_as(P),
                                   ctypes.byref(cfg), nat.stream_ptr(self.device))
        self.last_ms = (time.perf_counter() - t0) * 1e3
        nat.check(rc, "qs_qnn_batch")
        return loss.value, ev, gth, gx

    # --- autotuning -----------------------------------------------------------
    def _key(self) -> str:
        hist = np.bincount(self.widths, minlength=self.max_width + 1).tolist()
        return json.dumps({"dtype": self.precision, "hist": hist, "layers": self.layers, "nfeat": self.nfeat})

    def search_space(self):
        """LDS width limit (tile size) x fusion depth (qubits per pass) x workgroup size x
        workgroups per CU."""
        lds_cap = 12 if self.precision == "double" else 13
        for lds in sorted({min(10, lds_cap), min(11, lds_cap), lds_cap}):
            for fuse in (1, 2, 3):
                for threads in (256, 512):
                    for per_cu in (1, 2):
                        amp = 16 if self.precision == "double" else 8
                        if per_cu * 2 * (1 << lds) * amp > 160 * 1024:
                            continue
                        yield {"lds_max_width": lds, "block_threads": threads, "blocks_per_cu": per_cu,
                               "use_graph": 1, "fuse_qubits": fuse}

    def autotune(self, theta: np.ndarray, reps: int = 2, use_cache: bool = True) -> Dict:
        key = self._key()
        cache = {}
        if use_cache and os.path.exists(CACHE_PATH):
            try:
                with open(CACHE_PATH) as f:
                    cache = json.load(f)
            except (OSError, ValueError):
                cache = {}
        if use_cache and key in cache:
            self.config = cache[key]["config"]
            return self.config
        best, best_ms = None, float("inf")
        trials = []
        for cfg in self.search_space():
            self.forward_backward(theta, cfg)  # capture / warm
            t0 = time.perf_counter()
            for _ in range(reps):
                self.forward_backward(theta, cfg)
            ms = (time.perf_counter() - t0) / reps * 1e3
            trials.append((cfg, ms))
            if ms < best_ms:
                best, best_ms = cfg, ms
        self.config = best
        if use_cache:
            cache[key] = {"config": best, "ms": best_ms}
            os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
            with open(CACHE_PATH, "w") as f:
                json.dump(cache, f)
        self.tuning_trials = trials
        return best


_PROGRAM_KINDS = {GateType.rx: nat.QS_QNN_RX, GateType.ry: nat.QS_QNN_RY, GateType.rz: nat.QS_QNN_RZ}


def program_ops(pc: ParameterizedCircuit, encode=()) -> list:
    """A circuit as qs_qnn_op records: Ry(x[feature]) encodings on the given
    (qubit, feature) pairs first, then the circuit's Rx / Ry / Rz (trainable
    parameter or fixed angle), CX and CZ gates."""
    ops = []
    for q, j in encode:
        ops.append((nat.QS_QNN_ENC_RY, int(q), 0, int(j), 0.0))
    for gtype, qubits, param, vals in pc.ops:
        if gtype in _PROGRAM_KINDS:
            if param is not None:
                ops.append((_PROGRAM_KINDS[gtype], qubits[0], 0, int(param), 0.0))
            else:
                ops.append((_PROGRAM_KINDS[gtype], qubits[0], 0, -1, float(vals[0])))
        elif gtype == GateType.cx:
            ops.append((nat.QS_QNN_CX, qubits[0], qubits[1], -1, 0.0))
        elif gtype == GateType.cz:
            ops.append((nat.QS_QNN_CZ, qubits[0], qubits[1], -1, 0.0))
        else:
            raise qerr.GateAlgorithmNotImplementError(
                f"{gtype.name}: QNN programs take Rx, Ry, Rz, CX and CZ gates")
    return ops


class QNNProgramBatch:
    """Circuits with their own gate programs (mixed ansaetze and widths), one
    persistent launch for forward and adjoint backward (qs_qnn_batch_programs).

    ``circuits``: ParameterizedCircuit per batch entry (parameters index the
    shared theta); ``encode``: per circuit, (qubit, feature) pairs encoded as
    Ry(x[c, feature]) before the circuit; loss = mean_c (<Z_0>_c - t_c)^2."""

    def __init__(self, circuits, x, targets, nparams: int, encode=None, precision: str = "double",
                 device: int = 0, config: Optional[Dict] = None):
        self.ncirc = len(circuits)
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.ncirc, -1)
        self.nfeat = self.x.shape[1]
        self.targets = np.ascontiguousarray(targets, dtype=np.float64)
        if len(self.targets) != self.ncirc:
            raise qerr.ValueError("circuits, x and targets must agree on the batch size")
        self.nparams = int(nparams)
        encode = encode or [()] * self.ncirc
        recs, off = [], [0]
        for pc, enc in zip(circuits, encode):
            recs.extend(program_ops(pc, enc))
            off.append(len(recs))
        self.widths = np.array([pc.n for pc in circuits], dtype=np.int32)
        self.offsets = np.array(off, dtype=np.int32)
        self.ops = (nat.QsQnnOp * max(len(recs), 1))()
        for i, r in enumerate(recs):
            self.ops[i] = nat.QsQnnOp(*r)
        self.precision = precision
        self.dcode = nat.QS_C128 if precision == "double" else nat.QS_C64
        self.device = device
        self.lib = nat.load_library()
        self.ctx = nat.context(device)
        self.config = dict(config or {"lds_max_width": 11, "block_threads": 256, "blocks_per_cu": 2})

    def forward_backward(self, theta: np.ndarray) -> Tuple[float, np.ndarray, np.ndarray, np.ndarray]:
        theta = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
        if theta.size != self.nparams:
            raise qerr.ValueError(f"theta needs {self.nparams} values")
        ev = np.zeros(self.ncirc)
        loss = ctypes.c_double(0)
        gth = np.zeros(max(self.nparams, 1))
        gx = np.zeros((self.ncirc, max(self.nfeat, 1)))
        c = nat.QsQnnConfig()
        c.lds_max_width = self.config["lds_max_width"]
        c.block_threads = self.config["block_threads"]
        c.blocks_per_cu = self.config["blocks_per_cu"]
        P = ctypes.c_void_p
        nat.check(self.lib.qs_qnn_batch_programs(
            self.ctx, self.dcode, self.ncirc, self.widths.ctypes.data_as(P), self.offsets.ctypes.data_as(P),
            ctypes.cast(self.ops, P), self.nparams, self.nfeat, self.x.ctypes.data_as(P),
            self.targets.ctypes.data_as(P), theta.ctypes.data_as(P), ev.ctypes.data_as(P), ctypes.byref(loss),
            gth.ctypes.data_as(P), gx.ctypes.data_as(P), ctypes.byref(c), nat.stream_ptr(self.device)),
            "qs_qnn_batch_programs")
        return loss.value, ev, gth[: self.nparams], gx[:, : self.nfeat]
