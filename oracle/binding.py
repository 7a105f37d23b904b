"""ctypes binding of the FP64 CPU oracle (oracle/cgfdm_oracle.cpp).

Test infrastructure only: the tests, the smoke check and the CPU-baseline
leg of bench.py use it as the checker; the product path never does.
"""
import ctypes
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    src = os.path.join(ORACLE_DIR, "cgfdm_oracle.cpp")
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    D = ctypes.c_double
    L = ctypes.c_long
    lib.orc_create.restype = P
    lib.orc_destroy.argtypes = [P]
    lib.orc_set_param.argtypes = [P, ctypes.c_char_p, D]
    lib.orc_set_geometry.argtypes = [P, P, P, P, L]
    lib.orc_set_coord3d.argtypes = [P, P, L]
    lib.orc_set_topography.argtypes = [P, P, L]
    lib.orc_set_media1d.argtypes = [P, ctypes.c_int, P, P, P, P]
    lib.orc_set_media3d.argtypes = [P, P, P, P, L]
    lib.orc_set_stress.argtypes = [P, ctypes.c_char_p, P, L]
    lib.orc_setup.argtypes = [P]
    lib.orc_step.argtypes = [P, ctypes.c_int]
    lib.orc_stage.argtypes = [P] + [ctypes.c_int] * 5
    lib.orc_fault_trial.argtypes = [P] + [ctypes.c_int] * 5
    lib.orc_fault_deriv.argtypes = [P] + [ctypes.c_int] * 3
    lib.orc_rup_sensor.argtypes = [P]
    lib.orc_state_rk.argtypes = [P, ctypes.c_int]
    lib.orc_set_mutation.argtypes = [ctypes.c_int]
    lib.orc_thermpress.argtypes = [P]
    lib.orc_smooth.argtypes = [P]
    lib.orc_wave_deriv.argtypes = [P] + [ctypes.c_int] * 3
    lib.orc_pml_deriv.argtypes = [P] + [ctypes.c_int] * 3
    lib.orc_wave_rk.argtypes = [P, ctypes.c_int]
    lib.orc_apply_cerjan.argtypes = [P]
    lib.orc_nr_search.argtypes = [P, D, D, D, D, D, ctypes.POINTER(D)]
    lib.orc_newton_stats.argtypes = [P, ctypes.POINTER(L), ctypes.POINTER(ctypes.c_int)]
    lib.orc_newton_stats.restype = L
    lib.orc_size.argtypes = [P, ctypes.c_char_p]
    lib.orc_size.restype = L
    lib.orc_get.argtypes = [P, ctypes.c_char_p, P, L]
    lib.orc_put.argtypes = [P, ctypes.c_char_p, P, L]
    for n in ["orc_set_param", "orc_set_geometry", "orc_set_coord3d", "orc_set_topography", "orc_set_media1d", "orc_set_media3d",
              "orc_set_stress", "orc_setup", "orc_step", "orc_stage", "orc_rup_sensor", "orc_smooth", "orc_wave_deriv",
              "orc_pml_deriv", "orc_wave_rk", "orc_apply_cerjan", "orc_nr_search", "orc_get", "orc_put",
              "orc_state_rk", "orc_set_mutation"]:
        getattr(lib, n).restype = ctypes.c_int
    _lib = lib
    return lib


def set_mutation(m):
    """Plant mistake number m in the oracle (0: none).  Pin tests use it to
    show that their check fails on a plausible error."""
    load().orc_set_mutation(int(m))


class Oracle:
    def __init__(self, params):
        self.lib = load()
        self.h = self.lib.orc_create()
        for k, v in params.items():
            if isinstance(v, (list, tuple)):
                for q, x in enumerate(v):
                    self.lib.orc_set_param(self.h, f"{k}{q}".encode(), float(x))
            elif isinstance(v, (int, float, bool)):
                self.lib.orc_set_param(self.h, k.encode(), float(v))

    def __del__(self):
        try:
            self.lib.orc_destroy(self.h)
        except Exception:
            pass

    def set_geometry(self, x, y, z):
        a = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (x, y, z)]
        self._g = a
        assert self.lib.orc_set_geometry(self.h, a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, a[0].size) == 0

    def set_coord3d(self, xyz):
        a = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
        assert self.lib.orc_set_coord3d(self.h, a.ctypes.data, a.size) == 0

    def set_topography(self, h):
        a = np.ascontiguousarray(h, dtype=np.float64).ravel()
        self._t = a
        assert self.lib.orc_set_topography(self.h, a.ctypes.data, a.size) == 0

    def set_media1d(self, depth, rho, vp, vs):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (depth, rho, vp, vs)]
        assert self.lib.orc_set_media1d(self.h, len(a[0]), *[x.ctypes.data for x in a]) == 0

    def set_media3d(self, vp, vs, rho):
        a = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (vp, vs, rho)]
        assert self.lib.orc_set_media3d(self.h, *[x.ctypes.data for x in a], a[0].size) == 0

    def set_stress(self, name, v):
        a = np.ascontiguousarray(v, dtype=np.float64).ravel()
        assert self.lib.orc_set_stress(self.h, name.encode(), a.ctypes.data, a.size) == 0

    def setup(self):
        rc = self.lib.orc_setup(self.h)
        assert rc == 0, f"oracle setup failed {rc}"

    def dims(self):
        d = self.get("dims")
        return dict(zip(["ni", "nj", "nk", "nx", "ny", "nz", "i0"], [int(x) for x in d]))

    def size(self, name):
        return self.lib.orc_size(self.h, name.encode())

    def get(self, name):
        n = self.size(name)
        assert n >= 0, f"unknown oracle field {name}"
        out = np.empty(n, dtype=np.float64)
        assert self.lib.orc_get(self.h, name.encode(), out.ctypes.data, n) == 0
        return out

    def put(self, name, v):
        a = np.ascontiguousarray(v, dtype=np.float64).ravel()
        assert self.lib.orc_put(self.h, name.encode(), a.ctypes.data, a.size) == 0

    def step(self, it):
        self.lib.orc_step(self.h, it)

    def stage(self, it, irk, fx, fy, fz):
        self.lib.orc_stage(self.h, it, irk, fx, fy, fz)

    def fault_trial(self, it, irk, fx, fy, fz):
        self.lib.orc_fault_trial(self.h, it, irk, fx, fy, fz)

    def fault_deriv(self, fx, fy, fz):
        self.lib.orc_fault_deriv(self.h, fx, fy, fz)

    def rup_sensor(self):
        self.lib.orc_rup_sensor(self.h)

    def state_rk(self, irk):
        self.lib.orc_state_rk(self.h, irk)

    def thermpress(self):
        self.lib.orc_thermpress(self.h)

    def smooth(self):
        self.lib.orc_smooth(self.h)

    def wave_deriv(self, fx, fy, fz):
        self.lib.orc_wave_deriv(self.h, fx, fy, fz)

    def pml_deriv(self, axis, face, flag):
        self.lib.orc_pml_deriv(self.h, axis, face, flag)

    def wave_rk(self, irk):
        self.lib.orc_wave_rk(self.h, irk)

    def apply_cerjan(self):
        self.lib.orc_apply_cerjan(self.h)

    def nr_search(self, x, psi, a, vhat, coef):
        out = ctypes.c_double()
        it = self.lib.orc_nr_search(self.h, x, psi, a, vhat, coef, ctypes.byref(out))
        return out.value, it

    def newton_stats(self):
        it = ctypes.c_long()
        mx = ctypes.c_int()
        n = self.lib.orc_newton_stats(self.h, ctypes.byref(it), ctypes.byref(mx))
        return n, it.value, mx.value
