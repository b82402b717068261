"""Loader for the C restatement of the reference CPU simulator (oracle/csrc).

Measurement infrastructure only: bench.py times it as the CPU baseline and the
tests check it against the numpy oracle.  Built by ``build()`` (gcc -O3,
OpenMP) into oracle/libcpu_ref_sim.so.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "cpu_ref_sim.c")
LIB = os.path.join(HERE, "libcpu_ref_sim.so")
INC = os.path.join(os.path.dirname(HERE), "include")

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
        subprocess.check_call(["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-I", INC, SRC,
                               "-o", LIB, "-lm"])
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.cpu_apply_gates.restype = ctypes.c_int
        lib.cpu_apply_gates.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        lib.cpu_prob_zero.restype = ctypes.c_double
        lib.cpu_prob_zero.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        lib.cpu_sample.restype = ctypes.c_int
        lib.cpu_sample.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        _lib = lib
    return _lib


def apply_gates(vec: np.ndarray, n: int, recs: np.ndarray, threads: int = 1):
    assert vec.dtype == np.complex128 and vec.flags["C_CONTIGUOUS"]
    lib = load()
    lib.cpu_apply_gates(vec.ctypes.data, n, recs.ctypes.data if len(recs) else None, len(recs), threads)


def sample(vec: np.ndarray, n: int, uniforms: np.ndarray, shots: int) -> np.ndarray:
    lib = load()
    out = np.zeros(shots, dtype=np.uint64)
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    lib.cpu_sample(vec.ctypes.data, n, u.ctypes.data, shots, out.ctypes.data)
    return out
