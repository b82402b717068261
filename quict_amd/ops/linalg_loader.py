"""``LinAlgLoader``: the reference's plugin table, bound to this build's C ABI.

QuICT's state-vector simulator reaches its GPU kernels through
``LinAlgLoader(device="GPU", enable_gate_kernel=True)``
(QuICT/ops/utils/algorithm_loader.py:27-60): an object whose attributes are
the 21 functions of ``gate_function.__outward_functions``
(QuICT/ops/gate_kernel/gate_function.py:6-28) plus the based linear algebra of
``BASED_LINALG`` (algorithm_loader.py:11-14; gpu_calculator.py).  The same
names, argument order and meaning are exported here, each one a single call of
``qs_sv_apply_gate`` (or of the measurement / permutation / mat-vec entry
points) on the caller's vector:

  * qubit arguments are the reversed indices the reference passes
    (``n - 1 - qubit`` = amplitude-index bit), ``vec_bit`` is n;
  * ``vec`` is a device tensor (torch, complex64 / complex128) updated in
    place, like the reference's CuPy array; ``matrix_dot_vector`` and
    ``VectorPermutation`` return a new vector as the reference does;
  * ``mat`` is the gate matrix (2x2 / 4x4 / 8x8, flat or square; a device or
    host tensor or a numpy array), read entry by entry exactly where the
    reference kernel reads it;
  * ``sync=True`` waits for the device, as ``cp.cuda.Device().synchronize()``.

A 2-target kernel's matrix index follows the reference kernel's own address
arithmetic: for ``normal_targs`` / ``diagonal_normal_targs`` the first index
is the matrix LSB; ``diagonal_targs``, ``swap_targs``, ``ctrl_normal_targs``
and ``normal_normal_targs`` sort the two indices and take the higher as the
MSB.  There is no CPU device: ``LinAlgLoader("CPU")`` raises.
"""
import ctypes
import math
import random
from typing import Sequence

import numpy as np

from .. import _native as nat
from ..core import exception as qerr

__outward_functions = [
    "diagonal_targ", "diagonal_targs", "normal_targ", "normal_targs", "control_targ", "diagonal_ctargs",
    "control_ctargs", "normal_ctargs", "ctrl_normal_targs", "normal_normal_targs", "diagonal_normal_targs",
    "swap_targ", "reverse_targ", "reverse_ctargs", "swap_targs", "reverse_more", "diagonal_more", "swap_tmore",
    "measured_prob_calculate", "apply_measuregate", "apply_resetgate",
]
GATE_KERNEL_FUNCTIONS = list(__outward_functions)
BASED_LINALG = ["dot", "tensor", "MatrixTensorI", "VectorPermutation", "matrix_dot_vector"]


class DeviceScalar(float):
    """A host float with CuPy's ``.get()`` (the reference returns a device
    scalar from ``measured_prob_calculate`` and its caller calls ``.get()``)."""

    def get(self) -> float:
        return float(self)


def _flat(mat) -> np.ndarray:
    if hasattr(mat, "detach"):
        mat = mat.detach().cpu().numpy()
    elif hasattr(mat, "get") and not isinstance(mat, np.ndarray):
        mat = mat.get()
    return np.asarray(mat, dtype=np.complex128).reshape(-1)


def _vec_info(vec, vec_bit):
    import torch
    if not isinstance(vec, torch.Tensor) or not vec.is_cuda:
        raise qerr.TypeError("vec must be a device tensor (torch, complex64/complex128), got", type(vec))
    if vec.numel() != (1 << vec_bit):
        raise qerr.ValueError(f"vec has {vec.numel()} amplitudes, vec_bit={vec_bit} needs {1 << vec_bit}")
    if not vec.is_contiguous():
        raise qerr.ValueError("vec must be contiguous")
    return nat.dtype_code(vec.dtype), vec.device.index or 0


def _gate(kind: int, controls: Sequence[int], targets: Sequence[int], payload) -> np.ndarray:
    rec = np.zeros(1, dtype=nat.QS_GATE_DTYPE)
    rec["kind"] = kind
    rec["ncontrols"] = len(controls)
    rec["ntargets"] = len(targets)
    bits = [int(b) for b in list(controls) + list(targets)]
    rec["bits"][0, :len(bits)] = bits
    p = np.asarray(payload, dtype=np.complex128).reshape(-1)
    m = np.zeros(32)
    m[0:2 * len(p):2] = p.real
    m[1:2 * len(p):2] = p.imag
    rec["m"][0] = m
    return rec


class LinAlgLoader:
    """Same constructor as QuICT/ops/utils/algorithm_loader.py:27-60; ``device``
    must be "GPU".  ``gpu_device_id`` selects the HIP device of the context
    (the reference uses CuPy's current device)."""

    def __init__(self, device: str = "GPU", enable_gate_kernel: bool = False,
                 enable_multigpu_gate_kernel: bool = False, gpu_device_id: int = 0):
        if device not in ("CPU", "GPU"):
            raise KeyError(f"Not supported the given device, please choice one of ['CPU', 'GPU']")
        if device != "GPU":
            raise qerr.ValueError("LinAlgLoader: this build has only the GPU (HIP) device")
        if enable_multigpu_gate_kernel:
            raise qerr.GateAlgorithmNotImplementError("multi-GPU gate kernels are not part of this build")
        self._lib = nat.load_library()
        self._device_id = gpu_device_id
        self._ctx = nat.context(gpu_device_id)
        for name in BASED_LINALG:
            self.__dict__[name] = getattr(self, "_" + name)
        if enable_gate_kernel:
            for name in GATE_KERNEL_FUNCTIONS:
                self.__dict__[name] = getattr(self, "_" + name)

    # ---- plumbing -----------------------------------------------------------------
    def _apply(self, rec: np.ndarray, vec, vec_bit: int, sync: bool):
        dcode, dev = _vec_info(vec, vec_bit)
        rc = self._lib.qs_sv_apply_gate(self._ctx, ctypes.c_void_p(vec.data_ptr()), vec_bit, dcode,
                                        nat.gate_array_ptr(rec), nat.stream_ptr(dev))
        nat.check(rc, "qs_sv_apply_gate")
        if sync:
            import torch
            torch.cuda.synchronize(dev)

    # ---- gate_function.py:1034-1590 ------------------------------------------------
    def _diagonal_targ(self, t_index, mat, vec, vec_bit, sync=False):
        """[[a,0],[0,b]] on bit t_index (gate_function.py:1034-1058)."""
        m = _flat(mat)
        self._apply(_gate(nat.GATE_DIAG, [], [t_index], [m[0], m[3]]), vec, vec_bit, sync)

    def _diagonal_targs(self, t_indexes, mat, vec, vec_bit, sync=False):
        """diag(m0, m5, m10, m15) on (high, low) (gate_function.py:1061-1092)."""
        m = _flat(mat)
        hi, lo = max(t_indexes), min(t_indexes)
        self._apply(_gate(nat.GATE_DIAG, [], [hi, lo], [m[0], m[5], m[10], m[15]]), vec, vec_bit, sync)

    def _normal_targ(self, t_index, mat, vec, vec_bit, sync=False):
        """2x2 on bit t_index (gate_function.py:1095-1119)."""
        self._apply(_gate(nat.GATE_DENSE, [], [t_index], _flat(mat)[:4]), vec, vec_bit, sync)

    def _normal_targs(self, t_indexes, mat, vec, vec_bit, sync=False):
        """4x4 whose index is bit(t_indexes[1]) * 2 + bit(t_indexes[0]) (gate_function.py:1122-1148)."""
        self._apply(_gate(nat.GATE_DENSE, [], [t_indexes[1], t_indexes[0]], _flat(mat)[:16]), vec, vec_bit, sync)

    def _control_targ(self, t_index, val, vec, vec_bit, sync=False):
        """amplitudes with bit t_index set times val (gate_function.py:1151-1175)."""
        self._apply(_gate(nat.GATE_PHASE, [], [t_index], [complex(val)]), vec, vec_bit, sync)

    def _diagonal_ctargs(self, c_index, t_index, mat, vec, vec_bit, sync=False):
        """control set: diag(m10, m15) on t_index (gate_function.py:1178-1204)."""
        m = _flat(mat)
        self._apply(_gate(nat.GATE_DIAG, [c_index], [t_index], [m[10], m[15]]), vec, vec_bit, sync)

    def _control_ctargs(self, c_index, t_index, value, vec, vec_bit, sync=False):
        """both bits set: times value (gate_function.py:1207-1233)."""
        self._apply(_gate(nat.GATE_PHASE, [c_index], [t_index], [complex(value)]), vec, vec_bit, sync)

    def _normal_ctargs(self, c_index, t_index, mat, vec, vec_bit, sync=False):
        """control set: [[m10, m11], [m14, m15]] on t_index (gate_function.py:1236-1262)."""
        m = _flat(mat)
        self._apply(_gate(nat.GATE_DENSE, [c_index], [t_index], [m[10], m[11], m[14], m[15]]), vec, vec_bit,
                    sync)

    def _masked_4x4(self, t_indexes, mat, keep, vec, vec_bit, sync):
        m = _flat(mat)[:16]
        full = np.zeros(16, dtype=np.complex128)
        full[keep] = m[keep]
        hi, lo = max(t_indexes), min(t_indexes)
        self._apply(_gate(nat.GATE_DENSE, [], [hi, lo], full), vec, vec_bit, sync)

    def _ctrl_normal_targs(self, t_indexes, mat, vec, vec_bit, sync=False):
        """diag corners m0, m15 and the middle block m5, m6, m9, m10 on (high, low)
        (gate_function.py:1265-1296, FSim)."""
        self._masked_4x4(t_indexes, mat, [0, 5, 6, 9, 10, 15], vec, vec_bit, sync)

    def _normal_normal_targs(self, t_indexes, mat, vec, vec_bit, sync=False):
        """outer block m0, m3, m12, m15 and middle block m5, m6, m9, m10 on
        (high, low) (gate_function.py:1299-1330, Rxx / Ryy)."""
        self._masked_4x4(t_indexes, mat, [0, 3, 12, 15, 5, 6, 9, 10], vec, vec_bit, sync)

    def _diagonal_normal_targs(self, t_indexes, mat, vec, vec_bit, sync=False):
        """two 2x2 blocks (m0, m1, m4, m5), (m10, m11, m14, m15); index
        bit(t_indexes[1]) * 2 + bit(t_indexes[0]) (gate_function.py:1333-1359, Rzx)."""
        m = _flat(mat)[:16]
        full = np.zeros(16, dtype=np.complex128)
        keep = [0, 1, 4, 5, 10, 11, 14, 15]
        full[keep] = m[keep]
        self._apply(_gate(nat.GATE_DENSE, [], [t_indexes[1], t_indexes[0]], full), vec, vec_bit, sync)

    def _swap_targ(self, t_index, vec, vec_bit, sync=False):
        """X on t_index (gate_function.py:1362-1386)."""
        self._apply(_gate(nat.GATE_XPERM, [], [t_index], [1, 1]), vec, vec_bit, sync)

    def _reverse_targ(self, t_index, mat, vec, vec_bit, sync=False):
        """[[0, m1], [m2, 0]] on t_index (gate_function.py:1389-1413)."""
        m = _flat(mat)
        self._apply(_gate(nat.GATE_XPERM, [], [t_index], [m[1], m[2]]), vec, vec_bit, sync)

    def _reverse_ctargs(self, c_index, t_index, mat, vec, vec_bit, sync=False):
        """control set: [[0, m11], [m14, 0]] on t_index (gate_function.py:1416-1442)."""
        m = _flat(mat)
        self._apply(_gate(nat.GATE_XPERM, [c_index], [t_index], [m[11], m[14]]), vec, vec_bit, sync)

    def _swap_targs(self, t_indexes, mat, vec, vec_bit, sync=False):
        """|01> <- m6 |10>, |10> <- m9 |01> on (high, low) (gate_function.py:1445-1476)."""
        m = _flat(mat)
        hi, lo = max(t_indexes), min(t_indexes)
        self._apply(_gate(nat.GATE_SWAP, [], [hi, lo], [m[6], m[9]]), vec, vec_bit, sync)

    def _reverse_more(self, c_indexes, t_index, vec, vec_bit, sync=False):
        """both controls set: X on t_index (gate_function.py:1479-1514, CCX)."""
        self._apply(_gate(nat.GATE_XPERM, list(c_indexes), [t_index], [1, 1]), vec, vec_bit, sync)

    def _diagonal_more(self, c_indexes, t_index, mat, vec, vec_bit, sync=False):
        """both controls set: diag(m54, m63) on t_index (gate_function.py:1517-1552, CCRz)."""
        m = _flat(mat)
        self._apply(_gate(nat.GATE_DIAG, list(c_indexes), [t_index], [m[54], m[63]]), vec, vec_bit, sync)

    def _swap_tmore(self, t_indexes, c_index, vec, vec_bit, sync=False):
        """control set: exchange |01> and |10> of (high, low) (gate_function.py:1555-1590, CSwap)."""
        hi, lo = max(t_indexes), min(t_indexes)
        self._apply(_gate(nat.GATE_SWAP, [c_index], [hi, lo], [1, 1]), vec, vec_bit, sync)

    # ---- measurement (gate_function.py:1751-1847) ------------------------------------------
    def _measured_prob_calculate(self, index, vec, vec_bit, all_measured=False, sync=False):
        """P(bit ``index`` = 0); with ``all_measured`` the squared norm
        (gate_function.py:1751-1776)."""
        dcode, dev = _vec_info(vec, vec_bit)
        out = ctypes.c_double(0)
        if all_measured:
            rc = self._lib.qs_sv_norm2(self._ctx, ctypes.c_void_p(vec.data_ptr()), vec_bit, dcode,
                                       ctypes.byref(out), nat.stream_ptr(dev))
            nat.check(rc, "qs_sv_norm2")
        else:
            rc = self._lib.qs_sv_prob_zero(self._ctx, ctypes.c_void_p(vec.data_ptr()), vec_bit, dcode, index,
                                           ctypes.byref(out), nat.stream_ptr(dev))
            nat.check(rc, "qs_sv_prob_zero")
        return DeviceScalar(out.value)

    def _apply_measuregate(self, index, vec, vec_bit, prob, sync=False):
        """u = random.random(); outcome = u > prob; the surviving half scaled by
        1/sqrt(P) (gate_function.py:1779-1815).  Returns the outcome (bool)."""
        dcode, dev = _vec_info(vec, vec_bit)
        prob = float(prob.get() if hasattr(prob, "get") else prob)
        one = random.random() > prob
        scale = 1.0 / math.sqrt(1 - prob) if one else 1.0 / math.sqrt(prob)
        if dcode == nat.QS_C64:  # the reference scales in the vector's own precision
            scale = float(np.float32(scale))
        rc = self._lib.qs_sv_collapse(self._ctx, ctypes.c_void_p(vec.data_ptr()), vec_bit, dcode, index, int(one),
                                      scale, nat.stream_ptr(dev))
        nat.check(rc, "qs_sv_collapse")
        if sync:
            import torch
            torch.cuda.synchronize(dev)
        return one

    def _apply_resetgate(self, index, vec, vec_bit, prob, sync=False):
        """Bit ``index`` back to 0: |0> branch renormalised, or the |1> branch
        moved down when P0 < 1e-12 (gate_function.py:1818-1847; the reference's
        double-precision |1> kernel writes the wrong index, see the oracle)."""
        dcode, dev = _vec_info(vec, vec_bit)
        prob = float(prob.get() if hasattr(prob, "get") else prob)
        rc = self._lib.qs_sv_reset(self._ctx, ctypes.c_void_p(vec.data_ptr()), vec_bit, dcode, index, prob,
                                   nat.stream_ptr(dev))
        nat.check(rc, "qs_sv_reset")
        if sync:
            import torch
            torch.cuda.synchronize(dev)

    # ---- based linear algebra (gpu_calculator.py) ----------------------------------------
    def _matrix_dot_vector(self, vec, vec_bit, mat, mat_args, sync=True):
        """(I (x) U (x) I) vec on the qubits ``mat_args`` (argument 0 = matrix
        MSB), returned as a new vector (gpu_calculator.py:430-478)."""
        import torch
        if not isinstance(vec, torch.Tensor):
            vec = torch.as_tensor(np.asarray(vec), device=f"cuda:{self._device_id}")
        k = len(mat_args)
        m = _flat(mat)
        if m.size != 1 << (2 * k):
            raise qerr.ValueError(f"matrix has {m.size} entries, {k} qubits need {1 << (2 * k)}")
        out = vec.clone()
        rec = _gate(nat.GATE_DENSE, [], [vec_bit - 1 - int(q) for q in mat_args], m[:16] if k <= 2 else [])
        big = None
        if k >= 3:
            big = np.ascontiguousarray(np.stack([m.real, m.imag], axis=1).reshape(-1))
            rec["big"] = big.ctypes.data
        self._apply(rec, out, vec_bit, sync)
        del big
        return out

    def _VectorPermutation(self, A, mapping, changeInput: bool = False, gpu_out: bool = False, sync: bool = True):
        """y[j] = A[x(j)] where bit (n-1-i) of x(j) is bit (n-1-mapping[i]) of j
        (gpu_calculator.py:190-250).  changeInput writes the result into A;
        gpu_out returns a host array."""
        import torch
        mapping = np.asarray(mapping, dtype=np.int64)
        n = int(mapping.shape[0])
        if not isinstance(A, torch.Tensor):
            A = torch.as_tensor(np.asarray(A), device=f"cuda:{self._device_id}")
        if A.numel() != 1 << n:
            raise IndexError("Indices do not match!")
        dcode, dev = _vec_info(A, n)
        perm = np.zeros(n, dtype=np.int32)
        for i in range(n):
            perm[n - 1 - int(mapping[i])] = n - 1 - i
        out = torch.empty_like(A)
        rc = self._lib.qs_sv_permute(self._ctx, ctypes.c_void_p(A.data_ptr()), ctypes.c_void_p(out.data_ptr()), n,
                                     dcode, perm.ctypes.data_as(ctypes.c_void_p), nat.stream_ptr(dev))
        nat.check(rc, "qs_sv_permute")
        if changeInput:
            A.copy_(out)
        if sync:
            torch.cuda.synchronize(dev)
        if gpu_out:
            return out.cpu().numpy()
        return out

    def _dot(self, A, B, gpu_out: bool = False, sync: bool = True):
        """Matrix product (gpu_calculator.py:13-38: cp.dot); a library GEMM."""
        import torch
        dev = f"cuda:{self._device_id}"
        a = A if isinstance(A, torch.Tensor) else torch.as_tensor(np.asarray(A), device=dev)
        b = B if isinstance(B, torch.Tensor) else torch.as_tensor(np.asarray(B), device=dev)
        out = torch.matmul(a, b.contiguous())
        if sync:
            torch.cuda.synchronize(self._device_id)
        return out.cpu().numpy() if gpu_out else out

    def _tensor(self, A, B, gpu_out: bool = False, sync: bool = True):
        """Kronecker product (gpu_calculator.py:41-71)."""
        import torch
        dev = f"cuda:{self._device_id}"
        a = A if isinstance(A, torch.Tensor) else torch.as_tensor(np.asarray(A), device=dev)
        b = B if isinstance(B, torch.Tensor) else torch.as_tensor(np.asarray(B), device=dev)
        ra, rb = a.shape[0], b.shape[0]
        ca = 1 if a.dim() == 1 else a.shape[1]
        cb = 1 if b.dim() == 1 else b.shape[1]
        out = torch.kron(a.reshape(ra, ca), b.reshape(rb, cb))  # always 2-D, as the reference's kernel
        if sync:
            torch.cuda.synchronize(self._device_id)
        return out.cpu().numpy() if gpu_out else out

    def _MatrixTensorI(self, A, n, m, gpu_out: bool = False, sync: bool = True):
        """I_n (x) A (x) I_m (gpu_calculator.py:74-110)."""
        import torch
        dev = f"cuda:{self._device_id}"
        a = A if isinstance(A, torch.Tensor) else torch.as_tensor(np.asarray(A), device=dev)
        out = torch.kron(torch.kron(torch.eye(n, dtype=a.dtype, device=a.device), a),
                         torch.eye(m, dtype=a.dtype, device=a.device))
        if sync:
            torch.cuda.synchronize(self._device_id)
        return out.cpu().numpy() if gpu_out else out
