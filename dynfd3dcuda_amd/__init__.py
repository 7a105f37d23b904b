"""MI355X-native CGFDM dynamic-rupture solver (drop-in for the time-step path
of wenqiang-geophys/dynfd3dcuda).

Python mirror of the C ABI in include/cgfdm.h, loaded with ctypes.  The
product path is the HIP library ``libcgfdm.so`` built in-tree; importing this
module never falls back to anything else: every call raises if the library is
missing or a call fails.
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcgfdm.so")
CLI_PATH = os.path.join(HERE, "cgfdm")

# error codes (include/cgfdm.h)
EPARAM, EFILE, EJSON, ENETCDF, ECFL, EHIP, ESTATE, ESIZE, ENONFINITE = range(1, 10)

# FD operator direction table (srcd/main.cu:86-94); negated after each stage
FD_FLAGS = [(-1, -1, -1), (1, 1, -1), (1, 1, 1), (-1, -1, 1),
            (-1, 1, -1), (1, -1, -1), (1, -1, 1), (-1, 1, 1)]

EXPORTED = [
    "cgfdm_last_error", "cgfdm_create", "cgfdm_create_from_string", "cgfdm_set_param",
    "cgfdm_get_param", "cgfdm_destroy", "cgfdm_set_fault_geometry", "cgfdm_set_coord3d",
    "cgfdm_set_topography", "cgfdm_set_media1d", "cgfdm_set_media3d", "cgfdm_set_init_stress", "cgfdm_setup", "cgfdm_step", "cgfdm_stage",
    "cgfdm_run", "cgfdm_synchronize", "cgfdm_wave_stage", "cgfdm_wave_rhs", "cgfdm_field_info",
    "cgfdm_get_field", "cgfdm_put_field", "cgfdm_get_dims", "cgfdm_open_outputs",
    "cgfdm_write_outputs", "cgfdm_close_outputs", "cgfdm_newton_stats", "cgfdm_profile",
    "cgfdm_profile_get", "cgfdm_profile_aux", "cgfdm_set_newton_mlp", "cgfdm_mlp_forward", "cgfdm_mlp_grad", "cgfdm_mlp_train",
    "cgfdm_stage_flags", "cgfdm_set_stream", "cgfdm_fault_trial", "cgfdm_fault_deriv", "cgfdm_fault_update",
    "cgfdm_rup_sensor", "cgfdm_fault_smooth", "cgfdm_thermpress", "cgfdm_profile_classes",
    "cgfdm_stream_probe", "cgfdm_check_field", "cgfdm_print_params", "cgfdm_mlp_forward_noalloc",
]

# kernel classes timed by cgfdm_profile / cgfdm_profile_get
PROFILE_KERNELS = ["stage_interior", "stage_band", "stage_surface", "pml_aux",
                   "fault_pre", "fault_deriv", "fault_post", "rup_sensor", "fault_rs"]


class CgfdmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cgfdm error {code}: {msg}")
        self.code = code


_lib = None


def stage_flags(it, irk):
    """FD flag triple of stage irk of step it (main.cu:440-442, 515)."""
    s = -1 if irk % 2 else 1
    f = FD_FLAGS[it % 8]
    return f[0] * s, f[1] * s, f[2] * s


def load_library(path=None):
    """Load libcgfdm.so; raises OSError when it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"libcgfdm.so not found at {p}: run python -m dynfd3dcuda_amd.build")
    lib = ctypes.CDLL(p)
    c_p = ctypes.c_void_p
    lib.cgfdm_last_error.restype = ctypes.c_char_p
    lib.cgfdm_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(c_p)]
    lib.cgfdm_create_from_string.argtypes = [ctypes.c_char_p, ctypes.POINTER(c_p)]
    lib.cgfdm_set_param.argtypes = [c_p, ctypes.c_char_p, ctypes.c_double]
    lib.cgfdm_get_param.argtypes = [c_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
    lib.cgfdm_destroy.argtypes = [c_p]
    lib.cgfdm_set_fault_geometry.argtypes = [c_p, c_p, c_p, c_p, ctypes.c_size_t]
    lib.cgfdm_set_coord3d.argtypes = [c_p, c_p, ctypes.c_size_t]
    lib.cgfdm_set_topography.argtypes = [c_p, c_p, ctypes.c_size_t]
    lib.cgfdm_set_media1d.argtypes = [c_p, ctypes.c_int, c_p, c_p, c_p, c_p]
    lib.cgfdm_set_media3d.argtypes = [c_p, c_p, c_p, c_p, ctypes.c_size_t]
    lib.cgfdm_set_init_stress.argtypes = [c_p, ctypes.c_char_p, c_p, ctypes.c_size_t]
    lib.cgfdm_setup.argtypes = [c_p]
    lib.cgfdm_step.argtypes = [c_p, ctypes.c_int]
    lib.cgfdm_stage.argtypes = [c_p, ctypes.c_int, ctypes.c_int]
    lib.cgfdm_run.argtypes = [c_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.cgfdm_synchronize.argtypes = [c_p]
    lib.cgfdm_wave_stage.argtypes = [c_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.cgfdm_wave_rhs.argtypes = [c_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.cgfdm_field_info.argtypes = [c_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int)]
    lib.cgfdm_get_field.argtypes = [c_p, ctypes.c_char_p, c_p, ctypes.c_size_t]
    lib.cgfdm_put_field.argtypes = [c_p, ctypes.c_char_p, c_p, ctypes.c_size_t]
    lib.cgfdm_get_dims.argtypes = [c_p, ctypes.POINTER(ctypes.c_int)]
    lib.cgfdm_open_outputs.argtypes = [c_p]
    lib.cgfdm_write_outputs.argtypes = [c_p, ctypes.c_int]
    lib.cgfdm_close_outputs.argtypes = [c_p]
    lib.cgfdm_newton_stats.argtypes = [c_p] + [ctypes.POINTER(ctypes.c_double)] * 3
    lib.cgfdm_profile.argtypes = [c_p, ctypes.c_int]
    lib.cgfdm_profile_get.argtypes = [c_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
    lib.cgfdm_profile_aux.argtypes = [c_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
    lib.cgfdm_set_newton_mlp.argtypes = [c_p, c_p, ctypes.c_long]
    lib.cgfdm_mlp_forward.argtypes = [c_p, c_p, ctypes.c_long, c_p, ctypes.c_int]
    lib.cgfdm_mlp_grad.argtypes = [c_p, c_p, c_p, ctypes.c_long, c_p, c_p]
    lib.cgfdm_mlp_train.argtypes = [c_p, c_p, c_p, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_uint, c_p]
    ip = ctypes.POINTER(ctypes.c_int)
    lib.cgfdm_stage_flags.argtypes = [ctypes.c_int, ctypes.c_int, ip, ip, ip]
    lib.cgfdm_set_stream.argtypes = [c_p, c_p]
    for name in ("cgfdm_fault_trial", "cgfdm_fault_deriv", "cgfdm_fault_update"):
        getattr(lib, name).argtypes = [c_p, ctypes.c_int, ctypes.c_int]
    lib.cgfdm_rup_sensor.argtypes = [c_p]
    lib.cgfdm_fault_smooth.argtypes = [c_p]
    lib.cgfdm_thermpress.argtypes = [c_p]
    lib.cgfdm_profile_classes.argtypes = [c_p, ctypes.c_char_p]
    lib.cgfdm_check_field.argtypes = [c_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_longlong),
                                      ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_double)]
    lib.cgfdm_print_params.argtypes = [c_p]
    lib.cgfdm_mlp_forward_noalloc.argtypes = [c_p, ctypes.c_long, c_p, ctypes.c_long, c_p, ctypes.c_int, c_p]
    lib.cgfdm_stream_probe.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double)]
    for name in EXPORTED[1:]:
        getattr(lib, name).restype = ctypes.c_int
    if path is None:
        _lib = lib
    return lib


_DT = {0: np.float32, 1: np.float64, 2: np.int32, 3: np.uint16}


def stream_probe(nread, nwrite, bytes_per_array=1 << 30, reps=10):
    """GB/s of a kernel that reads `nread` and writes `nwrite` arrays once
    each (non-temporal 16-byte accesses; forms 1:0, 1:1, 2:1, 3:1, 2:2)."""
    lib = load_library()
    v = ctypes.c_double()
    rc = lib.cgfdm_stream_probe(nread, nwrite, bytes_per_array, reps, ctypes.byref(v))
    if rc != 0:
        raise CgfdmError(rc, lib.cgfdm_last_error().decode())
    return v.value


class Solver:
    """One simulation on one GPU.  Mirrors the reference's params.json
    interface (set_params.cu:56-298) and its launcher set (main.cu:49-84)."""

    def __init__(self, params=None, path=None, **overrides):
        self.lib = load_library()
        h = ctypes.c_void_p()
        if path is not None:
            self._check(self.lib.cgfdm_create(os.fsencode(path), ctypes.byref(h)))
        else:
            text = json.dumps(params or {})
            self._check(self.lib.cgfdm_create_from_string(text.encode(), ctypes.byref(h)))
        self.h = h
        for k, v in overrides.items():
            self.set_param(k, v)
        self._setup = False

    def _check(self, rc):
        if rc != 0:
            raise CgfdmError(rc, self.lib.cgfdm_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.cgfdm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- parameters and inputs
    def set_param(self, key, value):
        self._check(self.lib.cgfdm_set_param(self.h, key.encode(), float(value)))

    def get_param(self, key):
        v = ctypes.c_double()
        self._check(self.lib.cgfdm_get_param(self.h, key.encode(), ctypes.byref(v)))
        return v.value

    def dims(self):
        d = (ctypes.c_int * 8)()
        self._check(self.lib.cgfdm_get_dims(self.h, d))
        return dict(zip(["ni", "nj", "nk", "nx", "ny", "nz", "i0", "NT"], list(d)))

    def set_fault_geometry(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float32).ravel() for a in (x, y, z))
        self._keep = (x, y, z)
        self._check(self.lib.cgfdm_set_fault_geometry(self.h, x.ctypes.data, y.ctypes.data, z.ctypes.data, x.size))

    def set_coord3d(self, xyz):
        a = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
        self._check(self.lib.cgfdm_set_coord3d(self.h, a.ctypes.data, a.size))

    def set_topography(self, h):
        a = np.ascontiguousarray(h, dtype=np.float32).ravel()
        self._keep_topo = a
        self._check(self.lib.cgfdm_set_topography(self.h, a.ctypes.data, a.size))

    def set_media1d(self, depth, rho, vp, vs):
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (depth, rho, vp, vs)]
        self._check(self.lib.cgfdm_set_media1d(self.h, len(arrs[0]), *[a.ctypes.data for a in arrs]))

    def set_media3d(self, vp, vs, rho):
        """Media3D on the interior grid, [NX][NZ][NY] (media.cu:140-204)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float32).ravel() for a in (vp, vs, rho)]
        self._check(self.lib.cgfdm_set_media3d(self.h, *[a.ctypes.data for a in arrs], arrs[0].size))

    def set_init_stress(self, name, data):
        a = np.ascontiguousarray(data, dtype=np.float32).ravel()
        self._check(self.lib.cgfdm_set_init_stress(self.h, name.encode(), a.ctypes.data, a.size))

    def setup(self):
        self._check(self.lib.cgfdm_setup(self.h))
        self._setup = True

    # ---- time stepping
    def step(self, it):
        self._check(self.lib.cgfdm_step(self.h, it))

    def stage(self, it, irk):
        self._check(self.lib.cgfdm_stage(self.h, it, irk))

    def run(self, it0, nt, write_outputs=False):
        self._check(self.lib.cgfdm_run(self.h, it0, nt, 1 if write_outputs else 0))

    def synchronize(self):
        self._check(self.lib.cgfdm_synchronize(self.h))

    def wave_stage(self, irk, fx, fy, fz):
        self._check(self.lib.cgfdm_wave_stage(self.h, irk, fx, fy, fz))

    def wave_rhs(self, fx, fy, fz):
        self._check(self.lib.cgfdm_wave_rhs(self.h, fx, fy, fz))

    # ---- one stage a phase at a time (the reference's launcher set,
    # main.cu:448-515): fault_trial, wave_stage, fault_deriv, fault_update
    def fault_trial(self, it, irk):
        self._check(self.lib.cgfdm_fault_trial(self.h, it, irk))

    def fault_deriv(self, it, irk):
        self._check(self.lib.cgfdm_fault_deriv(self.h, it, irk))

    def fault_update(self, it, irk):
        self._check(self.lib.cgfdm_fault_update(self.h, it, irk))

    def rup_sensor(self):
        self._check(self.lib.cgfdm_rup_sensor(self.h))

    def fault_smooth(self):
        self._check(self.lib.cgfdm_fault_smooth(self.h))

    def thermpress(self):
        self._check(self.lib.cgfdm_thermpress(self.h))

    def stage_by_phases(self, it, irk):
        """The phases of one stage in cgfdm_stage's order."""
        fx, fy, fz = self.stage_flags_lib(it, irk)
        if self.get_param("Rupture"):
            self.fault_trial(it, irk)
        self.wave_stage(irk, fx, fy, fz)
        if self.get_param("Rupture"):
            self.fault_deriv(it, irk)
            self.fault_update(it, irk)

    def stage_flags_lib(self, it, irk):
        f = [ctypes.c_int() for _ in range(3)]
        self._check(self.lib.cgfdm_stage_flags(it, irk, *[ctypes.byref(x) for x in f]))
        return tuple(x.value for x in f)

    def set_stream(self, stream):
        """Queue on `stream`, a hipStream_t (as int) created with the HIP
        runtime this library is linked against, or None for the solver's own
        stream."""
        self._check(self.lib.cgfdm_set_stream(self.h, ctypes.c_void_p(stream) if stream else None))

    # ---- fields
    def field_info(self, name):
        n = ctypes.c_size_t()
        t = ctypes.c_int()
        self._check(self.lib.cgfdm_field_info(self.h, name.encode(), ctypes.byref(n), ctypes.byref(t)))
        return n.value, _DT[t.value]

    def media_palette_entries(self):
        """Entries of the media palette the stage kernels read lambda, mu,
        rho from (MetricBF16=2, MetricPalette=1); 0 when they read the FP32
        media words."""
        n, t = ctypes.c_size_t(), ctypes.c_int()
        rc = self.lib.cgfdm_field_info(self.h, b"M_palette", ctypes.byref(n), ctypes.byref(t))
        return n.value // 4 if rc == 0 else 0

    def metric_skip_mask(self):
        """MetricBF16=2: bit p set when the 32-bit metric word p (coefficient
        pair p) is 0 at every point and the stage kernels skip loading it."""
        n, t = ctypes.c_size_t(), ctypes.c_int()
        if self.lib.cgfdm_field_info(self.h, b"M_skip", ctypes.byref(n), ctypes.byref(t)) != 0:
            return 0
        return int(self.get("M_skip")[0])

    def get(self, name):
        n, dt = self.field_info(name)
        out = np.empty(n, dtype=dt)
        self._check(self.lib.cgfdm_get_field(self.h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def put(self, name, data):
        n, dt = self.field_info(name)
        a = np.ascontiguousarray(data, dtype=dt).ravel()
        if a.size != n:
            raise ValueError(f"{name}: expected {n} values, got {a.size}")
        self._check(self.lib.cgfdm_put_field(self.h, name.encode(), a.ctypes.data, a.nbytes))

    # ---- outputs
    def open_outputs(self):
        self._check(self.lib.cgfdm_open_outputs(self.h))

    def write_outputs(self, it):
        self._check(self.lib.cgfdm_write_outputs(self.h, it))

    def close_outputs(self):
        self._check(self.lib.cgfdm_close_outputs(self.h))

    def check_field(self, name):
        """Device-side check of an FP32 field: {"nonfinite", "ghost_nonzero", "maxabs"}."""
        a, b, m = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
        self._check(self.lib.cgfdm_check_field(self.h, name.encode(), ctypes.byref(a), ctypes.byref(b), ctypes.byref(m)))
        return {"nonfinite": a.value, "ghost_nonzero": b.value, "maxabs": m.value}

    def print_params(self):
        self._check(self.lib.cgfdm_print_params(self.h))

    def set_newton_mlp(self, params):
        """Network for NewtonMode 2 (float32 blob of 4555 values)."""
        a = np.ascontiguousarray(params, dtype=np.float32).ravel()
        self._check(self.lib.cgfdm_set_newton_mlp(self.h, a.ctypes.data, a.size))

    def newton_stats(self):
        v = [ctypes.c_double() for _ in range(3)]
        self._check(self.lib.cgfdm_newton_stats(self.h, *[ctypes.byref(x) for x in v]))
        return {"solves": v[0].value, "iterations": v[1].value, "fallbacks": v[2].value}

    def profile(self, enable=True):
        """Start (clearing earlier records) or stop per-kernel GPU timing."""
        self._check(self.lib.cgfdm_profile(self.h, 1 if enable else 0))

    def profile_classes(self, names=None):
        """Time only these kernel classes (None: all)."""
        self._check(self.lib.cgfdm_profile_classes(self.h, ",".join(names).encode() if names else None))

    def profile_report(self):
        """{kernel class: (total ms, launches, grid points per launch)} since profile(True)."""
        out = {}
        for name in PROFILE_KERNELS:
            ms, n, pts = ctypes.c_double(), ctypes.c_long(), ctypes.c_long()
            self._check(self.lib.cgfdm_profile_get(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n),
                                                   ctypes.byref(pts)))
            out[name] = (ms.value, n.value, pts.value)
        return out

    def profile_aux(self):
        """{kernel class: (ADE-PML (point, face) pairs whose auxiliary value
        one launch reads for the stage, pairs whose auxiliary ODE step it
        takes)} for the classes with such work."""
        out = {}
        for name in PROFILE_KERNELS:
            rd, up = ctypes.c_long(), ctypes.c_long()
            self._check(self.lib.cgfdm_profile_aux(self.h, name.encode(), ctypes.byref(rd), ctypes.byref(up)))
            if rd.value or up.value:
                out[name] = (rd.value, up.value)
        return out
