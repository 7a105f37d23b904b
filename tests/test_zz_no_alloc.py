"""cgfdm_mlp_forward_noalloc (the learned-guess network evaluated with the
caller's device arrays and a static device scratch, no allocation) against
the allocating entry point cgfdm_mlp_forward, through the C ABI: two calls in
a row, the second with fewer rows, both matrix-core and scalar forms, bit for
bit (the same kernels); sizes beyond the scratch are rejected."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

dyn = pytest.importorskip("dynfd3dcuda_amd")
from dynfd3dcuda_amd import mlp  # noqa: E402

NET = os.path.join(os.path.dirname(dyn.__file__), "data", "newton_guess_4x64x64x1.f32")


def hip_runtime():
    """The HIP runtime the library runs on (mapped into this process once the
    library has made a HIP call), for the test's own device buffers."""
    for line in open("/proc/self/maps"):
        if "libamdhip64" in line:
            return ctypes.CDLL(line.split()[-1])
    pytest.fail("libamdhip64 not mapped")


class DevArray:
    """A device copy of a float32 array made with the test's own hipMalloc."""

    def __init__(self, hip, a):
        self.hip, self.n = hip, a.size
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(max(a.nbytes, 4))) == 0
        a = np.ascontiguousarray(a, np.float32)
        assert hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0  # H2D

    def get(self):
        out = np.empty(self.n, np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def features(n, seed):
    """Feature rows in the ranges the rate-state trials produce."""
    rng = np.random.default_rng(seed)
    f = np.empty((n, 4), np.float32)
    f[:, 0] = rng.uniform(-60, 10, n)
    f[:, 1] = rng.uniform(0, 40, n)
    f[:, 2] = rng.uniform(-9, -2, n)
    f[:, 3] = rng.uniform(-14, 1, n)
    return f


def test_mlp_forward_noalloc_matches_allocating_entry_point():
    lib = dyn.load_library()
    params = mlp.load(NET)
    mlp.forward(params, features(64, 0))  # a first HIP call maps the runtime
    hip = hip_runtime()
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    P = DevArray(hip, params)
    try:
        for mode in (0, 1):
            for n, seed in ((7001, 1), (1234, 2)):  # two calls, the second smaller
                f = features(n, seed)
                ref = mlp.forward(params, f, mode=mode)
                F, O = DevArray(hip, f), DevArray(hip, np.full(n, np.nan, np.float32))
                rc = lib.cgfdm_mlp_forward_noalloc(P.p, params.size, F.p, n, O.p, mode, None)
                assert rc == 0, lib.cgfdm_last_error()
                assert hip.hipDeviceSynchronize() == 0
                got = O.get()
                F.free()
                O.free()
                np.testing.assert_array_equal(got, ref, err_msg=f"mode {mode}, n {n}")
        # sizes beyond the scratch, and bad sizes, are rejected
        assert lib.cgfdm_mlp_forward_noalloc(P.p, params.size + 1, P.p, 1, P.p, 0, None) == dyn.ESIZE
        assert "exactly 4555" in lib.cgfdm_last_error().decode()
        assert lib.cgfdm_mlp_forward_noalloc(P.p, params.size, P.p, -1, P.p, 0, None) == dyn.ESIZE
        assert lib.cgfdm_mlp_forward_noalloc(P.p, params.size, P.p, 1 << 31, P.p, 0, None) == dyn.ESIZE
    finally:
        P.free()
