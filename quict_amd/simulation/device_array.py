"""Device-resident result arrays.

The reference returns CuPy arrays from ``run`` (callers use ``.get()`` to copy
to the host).  These wrappers keep that calling convention over torch-ROCm
tensors: ``.get()`` -> numpy, ``.tensor`` -> the device tensor,
``np.asarray(x)`` works, and ``.data_ptr()`` exposes the device address.
"""
import numpy as np


class DeviceArray:
    def __init__(self, tensor, shape=None):
        self._t = tensor
        self._shape = tuple(shape) if shape is not None else tuple(tensor.shape)

    @property
    def tensor(self):
        return self._t

    @property
    def shape(self):
        return self._shape

    @property
    def size(self) -> int:
        return int(np.prod(self._shape)) if self._shape else 1

    @property
    def dtype(self):
        import torch
        return np.complex128 if self._t.dtype == torch.complex128 else np.complex64

    def get(self) -> np.ndarray:
        return self._t.detach().cpu().numpy().reshape(self._shape)

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a.astype(dtype) if dtype is not None else a

    def data_ptr(self) -> int:
        return self._t.data_ptr()

    def copy(self) -> "DeviceArray":
        return DeviceArray(self._t.clone(), self._shape)

    def __len__(self):
        return self._shape[0]

    def __repr__(self):
        return f"DeviceArray(shape={self._shape}, dtype={self.dtype.__name__}, device={self._t.device})"
