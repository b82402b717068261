"""qs_sv_sample_noalloc (fixed per-device scratch, no device allocation)
against qs_sv_sample through the C ABI: identical outcomes on two calls in a
row, the second with a smaller register and fewer shots, both precisions; a
request beyond the fixed scratch is refused before anything runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _lib():
    from quict_amd import _native as nat
    lib = nat.load_library()
    fn = lib.qs_sv_sample_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                   ctypes.c_void_p, ctypes.c_void_p]
    return nat, lib


def _state(n, precision, seed):
    import torch
    rng = np.random.RandomState(seed)
    v = rng.randn(1 << n) + 1j * rng.randn(1 << n)
    v /= np.linalg.norm(v)
    dt = torch.complex128 if precision == "double" else torch.complex64
    return torch.as_tensor(v, device="cuda:0").to(dt)


def _sample(nat, lib, fn_name, vec, n, uniforms, shots):
    out = np.zeros(shots, dtype=np.uint64)
    rc = getattr(lib, fn_name)(nat.context(0), ctypes.c_void_p(vec.data_ptr()), n, nat.dtype_code(vec.dtype),
                               uniforms.ctypes.data_as(ctypes.c_void_p), shots,
                               out.ctypes.data_as(ctypes.c_void_p), nat.stream_ptr(0))
    nat.check(rc, fn_name)
    return out


@pytest.mark.parametrize("precision", ["double", "single"])
def test_noalloc_sampling_equals_the_allocating_entry_point(precision):
    nat, lib = _lib()
    for n, shots, seed in [(18, 3000, 1), (11, 400, 2)]:  # second call: smaller register, fewer shots
        say("qs_sv_sample_noalloc vs qs_sv_sample", precision, n, shots)
        vec = _state(n, precision, seed)
        u = np.random.RandomState(seed + 50).random_sample((shots, n))
        a = _sample(nat, lib, "qs_sv_sample", vec, n, u, shots)
        b = _sample(nat, lib, "qs_sv_sample_noalloc", vec, n, u, shots)
        assert np.array_equal(a, b)
        assert b.max() < (1 << n)


def test_noalloc_sampling_refuses_requests_beyond_its_scratch():
    say("qs_sv_sample_noalloc refuses 23 qubits")
    nat, lib = _lib()
    n = 23
    vec = _state(n, "single", 3)
    u = np.full((4, n), 0.5)
    out = np.zeros(4, dtype=np.uint64)
    rc = lib.qs_sv_sample_noalloc(nat.context(0), ctypes.c_void_p(vec.data_ptr()), n, nat.dtype_code(vec.dtype),
                                  u.ctypes.data_as(ctypes.c_void_p), 4, out.ctypes.data_as(ctypes.c_void_p),
                                  nat.stream_ptr(0))
    assert rc == nat.QS_ERR_ARG
    assert b"40 MiB" in lib.qs_last_error()
