"""ctypes wrapper of oracle/liboracle.so — the CPU restatement used as the checker.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the shark_amd package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_P = C.POINTER(C.c_double)


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


def load():
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    lib = C.CDLL(path)
    lib.oracle_create.restype = C.c_void_p
    lib.oracle_create.argtypes = [C.c_void_p]
    lib.oracle_destroy.argtypes = [C.c_void_p]
    lib.oracle_ncells.argtypes = [C.c_void_p]
    lib.oracle_nvar.argtypes = [C.c_void_p]
    lib.oracle_set_geometry.argtypes = [C.c_void_p] + [_P] * 5
    lib.oracle_upload.argtypes = [C.c_void_p] + [_P] * 5
    lib.oracle_step.restype = C.c_double
    lib.oracle_step.argtypes = [C.c_void_p]
    lib.oracle_run.restype = C.c_int64
    lib.oracle_run.argtypes = [C.c_void_p, C.c_int64, C.c_double]
    lib.oracle_download.argtypes = [C.c_void_p, _P, _P]
    lib.oracle_primitives.argtypes = [C.c_void_p, _P]
    lib.oracle_get_aux.argtypes = [C.c_void_p, _P, _P, _P, _P]
    lib.oracle_time.restype = C.c_double
    lib.oracle_time.argtypes = [C.c_void_p]
    lib.oracle_dt.restype = C.c_double
    lib.oracle_dt.argtypes = [C.c_void_p]
    lib.oracle_slope_limit.restype = C.c_double
    lib.oracle_slope_limit.argtypes = [C.c_int, C.c_double, C.c_double]
    lib.oracle_gas_flux.argtypes = [C.c_int, C.c_double, _P, _P, _P, C.c_double, C.c_double,
                                    C.c_int, C.c_int]
    lib.oracle_dust_flux.argtypes = [C.c_int, C.c_double, _P, _P, _P, C.c_double, C.c_double]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


class Oracle:
    """CPU step on a problem dict (see shark_amd.setups)."""

    def __init__(self, cfg, geom, state):
        self.L = lib()
        self.cfg = cfg
        self.h = self.L.oracle_create(C.byref(cfg))
        self.nc = self.L.oracle_ncells(self.h)
        self.nv = self.L.oracle_nvar(self.h)
        self._keep = []
        g = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in geom.items()}
        self.L.oracle_set_geometry(self.h, _ptr(g["dx"]), _ptr(g["surf"]), _ptr(g["vol"]),
                                   _ptr(g["radii_c"]), _ptr(g.get("position")))
        s = {k: (None if v is None else np.ascontiguousarray(v, dtype=np.float64))
             for k, v in state.items()}
        self.L.oracle_upload(self.h, _ptr(s["u"]), _ptr(s["cs"]), _ptr(s.get("sdust")),
                             _ptr(s.get("condinit")), _ptr(s.get("fargo_velocity")))

    def step(self):
        return self.L.oracle_step(self.h)

    def run(self, nsteps, tend=1e300):
        return self.L.oracle_run(self.h, nsteps, tend)

    def u(self):
        out = np.empty(self.nv * self.nc)
        self.L.oracle_download(self.h, _ptr(out), None)
        return out.reshape(self.nv, self.nc)

    def q(self):
        out = np.empty(self.nv * self.nc)
        self.L.oracle_primitives(self.h, _ptr(out))
        return out.reshape(self.nv, self.nc)

    @property
    def time(self):
        return self.L.oracle_time(self.h)

    @property
    def dt(self):
        return self.L.oracle_dt(self.h)

    def close(self):
        if self.h:
            self.L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
