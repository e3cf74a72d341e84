"""ctypes binding of libshark_gpu.so (include/shark_gpu.h).

``Step`` mirrors the reference's usage of `solve` (src/solver/solve.f90:10): build it from a
configuration, the grid arrays and the state after setup, then call ``step()`` once per time
step (or ``run(n, tend)`` for n steps without host round trips) and ``download()`` at output
times.  There is no CPU fallback: if the HIP library is missing or no GPU is visible the
constructor raises.
"""
import ctypes as C
import os

import numpy as np

from .cfg import ShARKCfg

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshark_gpu.so")
_P = C.POINTER(C.c_double)

# every entry point of include/shark_gpu.h
EXPORTS = [
    "shark_gpu_create", "shark_gpu_set_geometry", "shark_gpu_set_dvol", "shark_gpu_upload", "shark_gpu_step",
    "shark_gpu_run", "shark_gpu_download", "shark_gpu_get_time", "shark_gpu_set_time",
    "shark_gpu_kernel_time", "shark_gpu_set_timing", "shark_gpu_algorithmic_bytes",
    "shark_gpu_destroy", "shark_gpu_last_error", "shark_gpu_version",
    "shark_gpu_get_tune", "shark_gpu_set_tune", "shark_gpu_autotune",
    "shark_output_write", "shark_gpu_write_output", "shark_gpu_save_state", "shark_gpu_load_state",
    "shark_growth_create", "shark_growth_run", "shark_growth_download", "shark_growth_upload",
    "shark_growth_destroy", "shark_growth_kernel_time", "shark_growth_charge",
    "shark_growth_set_cell_dt", "shark_growth_set_wave_cells", "shark_growth_features", "shark_growth_set_surrogate",
    "shark_growth_surrogate_stats", "shark_gpu_attach_growth", "shark_charge_marchand",
    "shark_charge_marchand_noalloc", "shark_growth_zmean", "shark_growth_set_surrogate_residual",
    "shark_mlp_create", "shark_mlp_num_params", "shark_mlp_set_params", "shark_mlp_get_params",
    "shark_mlp_forward", "shark_mlp_forward_hilo", "shark_mlp_forward_layered", "shark_mlp_grads", "shark_mlp_grads_weighted",
    "shark_mlp_set_data_weighted", "shark_mlp_set_data", "shark_mlp_train",
    "shark_mlp_kernel_time", "shark_mlp_destroy", "shark_hbm_probe",
]


class SharkGPUError(RuntimeError):
    pass


class ShARKOutputInfo(C.Structure):
    """struct shark_output_info: what info.dat records besides the grid."""
    _fields_ = [("version", C.c_int32), ("gridspace", C.c_int32), ("mhd", C.c_int32), ("turb", C.c_int32),
                ("ndustpscal", C.c_int32), ("charging", C.c_int32), ("time", C.c_double),
                ("unit_t", C.c_double), ("unit_d", C.c_double), ("unit_l", C.c_double),
                ("unit_v", C.c_double), ("unit_p", C.c_double), ("unit_dcol", C.c_double),
                ("base_dir", C.c_char_p)]


# unit sets of the reference setups (src/setups/<SETUP>/units.f90); disk_polar in cgs from
# src/commons/phys_const.f90 (au, msun, grav)
def units_of(setup):
    if setup == "disk_polar":
        au, msun, grav = 1.5e13, 2e33, 6.67e-8
        unit_l = au
        unit_d = msun / au ** 3
        unit_dcol = msun / au ** 2
        unit_t = 1.0 / np.sqrt(grav * msun / unit_l ** 3)
        unit_v = unit_l / unit_t
        return dict(unit_t=unit_t, unit_d=unit_d, unit_l=unit_l, unit_v=unit_v, unit_p=unit_v ** 2,
                    unit_dcol=unit_dcol)
    if setup in ("KH", "disk"):
        from .setups import units_cdisk, units_kh
        u = units_kh() if setup == "KH" else units_cdisk()
        return {k: u[k] for k in ("unit_t", "unit_d", "unit_l", "unit_v", "unit_p", "unit_dcol")}
    return dict(unit_t=1.0, unit_d=1.0, unit_l=1.0, unit_v=1.0, unit_p=1.0, unit_dcol=1.0)


def output_info(time, units, gridspace=0, base_dir=None):
    info = ShARKOutputInfo()
    info.version = 1
    info.gridspace = int(gridspace)
    info.time = float(time)
    for k, v in units.items():
        setattr(info, k, float(v))
    info.base_dir = None if base_dir is None else os.fsencode(base_dir)
    return info


def output_write(info, iout, cfg, q, sdust, position, radii_c):
    """shark_output_write: the reference's output directory from host arrays (no GPU needed)."""
    lib = load()
    arrs = [np.ascontiguousarray(x, dtype=np.float64) if x is not None else None
            for x in (q, sdust, position, radii_c)]
    check(lib.shark_output_write(C.byref(info), int(iout), C.byref(cfg), *[_ptr(x) for x in arrs]))


class ShARKTune(C.Structure):
    """struct shark_tune: launch shapes of the step kernels."""
    _fields_ = [("strip_h", C.c_int32), ("post_tw", C.c_int32), ("xcd_order", C.c_int32),
                ("strip_gas", C.c_int32), ("reserved_i", C.c_int32 * 4),
                ("hydro_ms", C.c_double), ("post_ms", C.c_double), ("reserved", C.c_double * 6)]

    def as_dict(self):
        return dict(strip_h=self.strip_h, strip_gas=self.strip_gas, post_tw=self.post_tw,
                    xcd_order=self.xcd_order, hydro_ms=self.hydro_ms, post_ms=self.post_ms)


_LIB = None


def _ptr(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("arrays passed to libshark_gpu must be contiguous float64")
    return a.ctypes.data_as(_P)


def load():
    """Load libshark_gpu.so (built in-tree by __graft_entry__.build())."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise SharkGPUError(f"{LIB_PATH} is missing; build it with `python -c "
                            "'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(LIB_PATH)
    lib.shark_gpu_last_error.restype = C.c_char_p
    lib.shark_gpu_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    lib.shark_gpu_set_geometry.argtypes = [C.c_void_p] + [_P] * 5
    lib.shark_gpu_upload.argtypes = [C.c_void_p] + [_P] * 5
    lib.shark_gpu_step.argtypes = [C.c_void_p, _P]
    lib.shark_gpu_run.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.POINTER(C.c_int64), _P]
    lib.shark_gpu_download.argtypes = [C.c_void_p, _P, _P]
    lib.shark_gpu_get_time.argtypes = [C.c_void_p, _P, _P]
    lib.shark_gpu_set_time.argtypes = [C.c_void_p, C.c_double]
    lib.shark_gpu_kernel_time.argtypes = [C.c_void_p, C.c_char_p, _P, C.POINTER(C.c_int64)]
    lib.shark_gpu_set_timing.argtypes = [C.c_void_p, C.c_int]
    lib.shark_gpu_algorithmic_bytes.restype = C.c_int64
    lib.shark_gpu_algorithmic_bytes.argtypes = [C.c_void_p]
    lib.shark_gpu_destroy.argtypes = [C.c_void_p]
    lib.shark_output_write.argtypes = [C.POINTER(ShARKOutputInfo), C.c_int, C.c_void_p, _P, _P, _P, _P]
    lib.shark_gpu_write_output.argtypes = [C.c_void_p, C.POINTER(ShARKOutputInfo), C.c_int]
    lib.shark_gpu_save_state.argtypes = [C.c_void_p, C.c_char_p]
    lib.shark_gpu_load_state.argtypes = [C.c_void_p, C.c_char_p]
    lib.shark_gpu_get_tune.argtypes = [C.c_void_p, C.POINTER(ShARKTune)]
    lib.shark_gpu_set_tune.argtypes = [C.c_void_p, C.POINTER(ShARKTune)]
    lib.shark_gpu_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(ShARKTune)]
    _LIB = lib
    return lib


def hbm_probe(nbytes=1 << 30, reps=20):
    """shark_hbm_probe: GB/s of a plain write, copy (read + write bytes) and read kernel."""
    lib = load()
    lib.shark_hbm_probe.argtypes = [C.c_int64, C.c_int, _P]
    out = (C.c_double * 3)()
    rc = lib.shark_hbm_probe(int(nbytes), int(reps), out)
    if rc != 0:
        raise SharkGPUError(f"shark_hbm_probe failed ({rc})")
    return {"write_GBs": out[0], "copy_GBs": out[1], "read_GBs": out[2]}


def check(rc):
    if rc != 0:
        raise SharkGPUError(f"libshark_gpu error {rc}: {load().shark_gpu_last_error().decode()}")


class Step:
    """Device-resident SHARK step for one configuration (one HIP stream, one context)."""

    def __init__(self, cfg: ShARKCfg, geom: dict, state: dict):
        self.L = load()
        self.cfg = cfg
        h = C.c_void_p()
        check(self.L.shark_gpu_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.nvar = 5 + 4 * cfg.ndust
        nxm = cfg.nx + 4
        nym = cfg.ny + 4 if cfg.ny > 1 else 1
        self.ncells = nxm * nym
        g = {k: (None if v is None else np.ascontiguousarray(v, dtype=np.float64))
             for k, v in geom.items()}
        check(self.L.shark_gpu_set_geometry(self.h, _ptr(g["dx"]), _ptr(g["surf"]),
                                            _ptr(g["vol"]), _ptr(g["radii_c"]),
                                            _ptr(g.get("position"))))
        if g.get("dvol") is not None and cfg.gravity:
            self.L.shark_gpu_set_dvol.argtypes = [C.c_void_p, _P]
            check(self.L.shark_gpu_set_dvol(self.h, _ptr(g["dvol"])))
        self.upload(state)

    def upload(self, state):
        s = {k: (None if v is None else np.ascontiguousarray(v, dtype=np.float64))
             for k, v in state.items()}
        check(self.L.shark_gpu_upload(self.h, _ptr(s["u"]), _ptr(s["cs"]), _ptr(s.get("sdust")),
                                      _ptr(s.get("condinit")), _ptr(s.get("fargo_velocity"))))

    def attach_growth(self, growth, mode):
        """Dust growth inside every step (shark_gpu_attach_growth); growth=None detaches."""
        self.L.shark_gpu_attach_growth.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        self._growth = growth
        check(self.L.shark_gpu_attach_growth(self.h, None if growth is None else growth.h, int(mode)))

    def step(self):
        dt = C.c_double()
        check(self.L.shark_gpu_step(self.h, C.byref(dt)))
        return dt.value

    def run(self, nsteps, tend=float("inf")):
        n = C.c_int64()
        t = C.c_double()
        check(self.L.shark_gpu_run(self.h, int(nsteps), float(tend), C.byref(n), C.byref(t)))
        return n.value, t.value

    def download(self, prims=False):
        u = np.empty(self.nvar * self.ncells)
        q = np.empty(self.nvar * self.ncells) if prims else None
        check(self.L.shark_gpu_download(self.h, _ptr(u), _ptr(q)))
        u = u.reshape(self.nvar, self.ncells)
        return (u, q.reshape(self.nvar, self.ncells)) if prims else u

    @property
    def time(self):
        t = C.c_double()
        d = C.c_double()
        check(self.L.shark_gpu_get_time(self.h, C.byref(t), C.byref(d)))
        return t.value

    def set_timing(self, on=True):
        check(self.L.shark_gpu_set_timing(self.h, 1 if on else 0))

    def kernel_time(self, name):
        ms = C.c_double()
        n = C.c_int64()
        check(self.L.shark_gpu_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def write_output(self, iout, units, gridspace=0, base_dir=None):
        """The reference's output_NNNNN/ directory of the current state (shark_gpu_write_output)."""
        info = output_info(self.time, units, gridspace, base_dir)
        check(self.L.shark_gpu_write_output(self.h, C.byref(info), int(iout)))

    def save_state(self, path):
        check(self.L.shark_gpu_save_state(self.h, os.fsencode(path)))

    def load_state(self, path):
        check(self.L.shark_gpu_load_state(self.h, os.fsencode(path)))

    def get_tune(self):
        t = ShARKTune()
        check(self.L.shark_gpu_get_tune(self.h, C.byref(t)))
        return t.as_dict()

    def set_tune(self, strip_h=None, post_tw=None, xcd_order=None, strip_gas=None):
        t = ShARKTune()
        check(self.L.shark_gpu_get_tune(self.h, C.byref(t)))
        if strip_h is not None:
            t.strip_h = int(strip_h)
        if strip_gas is not None:
            t.strip_gas = int(strip_gas)
        if post_tw is not None:
            t.post_tw = int(post_tw)
        if xcd_order is not None:
            t.xcd_order = int(xcd_order)
        check(self.L.shark_gpu_set_tune(self.h, C.byref(t)))

    def autotune(self, reps=5):
        """Time the candidate launch shapes on this state and keep the fastest (the state is
        restored afterwards); returns the chosen shape and its kernel times."""
        t = ShARKTune()
        check(self.L.shark_gpu_autotune(self.h, int(reps), C.byref(t)))
        return t.as_dict()

    def algorithmic_bytes(self):
        return self.L.shark_gpu_algorithmic_bytes(C.byref(self.cfg))

    def close(self):
        if getattr(self, "h", None):
            self.L.shark_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
