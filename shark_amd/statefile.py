"""State files read by the Fortran driver (shark_amd/fortran/shark_gpu_driver.f90, `state` mode)
and the Fortran driver's result files.

State file (little-endian stream): int32 magic 0x53484B31, int32 version 1; the shark_cfg
struct; int64 ncells, int32 ndim, int32 nvar; int32 has_sdust, has_condinit, has_fargo, pad;
FP64 dx, surf, vol, radii_c, position, u_prim, cs (column-major as the reference holds them);
the optional sdust, uprim_condinit, fargo_velocity; int64 nsteps, FP64 tend.
Result file: int64 ncells, nvar; FP64 initial u_prim, final u_prim; int64 steps; FP64 time.
"""
import ctypes as C
import struct

import numpy as np

MAGIC = 0x53484B31


def _w(f, a):
    f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())


def write_state(path, cfg, geom, state, nsteps, tend=float("inf")):
    nc = geom["vol"].size
    ndim = geom["dx"].shape[0]
    nvar = state["u"].shape[0]
    has = [state.get(k) is not None for k in ("sdust", "condinit", "fargo_velocity")]
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", MAGIC, 1))
        f.write(bytes(cfg))
        f.write(struct.pack("<qii", nc, ndim, nvar))
        f.write(struct.pack("<iiii", *[int(h) for h in has], 0))
        for k in ("dx", "surf", "vol", "radii_c", "position"):
            _w(f, geom[k])
        _w(f, state["u"])
        _w(f, state["cs"])
        for k, h in zip(("sdust", "condinit", "fargo_velocity"), has):
            if h:
                _w(f, state[k])
        f.write(struct.pack("<qd", int(nsteps), float(tend)))


def read_result(path):
    raw = open(path, "rb").read()
    nc, nv = struct.unpack_from("<qq", raw, 0)
    o = 16
    n = nc * nv
    u0 = np.frombuffer(raw, np.float64, n, o).reshape(nv, nc)
    o += 8 * n
    u = np.frombuffer(raw, np.float64, n, o).reshape(nv, nc)
    o += 8 * n
    steps, time = struct.unpack_from("<qd", raw, o)
    return u0.copy(), u.copy(), steps, time
