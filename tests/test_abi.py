"""The C-ABI library loads without a GPU and exports every entry point include/*.h declares."""
import ctypes
import glob
import os
import re

import pytest

from shark_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms |= set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w]*\s*\**\s*(shark_\w+)\s*\(", text, re.M))
    return syms


def test_header_declares_the_step_boundary():
    syms = declared_symbols()
    for s in ["shark_gpu_create", "shark_gpu_set_geometry", "shark_gpu_upload", "shark_gpu_step",
              "shark_gpu_run", "shark_gpu_download", "shark_gpu_destroy", "shark_gpu_last_error"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gpu.LIB_PATH)
    missing = [s for s in sorted(declared_symbols()) if not hasattr(lib, s)]
    assert not missing, missing


def test_algorithmic_bytes_formula():
    from shark_amd import cfg as C
    lib = gpu.load()
    c = C.make_cfg(nx=1024, ny=2048, ndust=16, fargo=1, geom=2)
    assert lib.shark_gpu_algorithmic_bytes(ctypes.byref(c)) == 2216 * 1024 * 2048
    assert C.algorithmic_bytes(c) == 2216 * 1024 * 2048
    c = C.make_cfg(nx=1024, ny=1024, ndust=8)
    assert lib.shark_gpu_algorithmic_bytes(ctypes.byref(c)) == 600 * 1024 * 1024


def test_create_rejects_out_of_scope_configs_without_gpu():
    """Validation runs before any device call, so error paths are testable on the CPU."""
    from shark_amd import cfg as C
    lib = gpu.load()
    h = ctypes.c_void_p()
    for kw, frag in [(dict(nx=16, ny=16, ndust=1, iso_cs=0, geom=C.GEOM_POLAR), b"cartesian"),
                     (dict(nx=64, ndust=1, iso_cs=2), b"iso_cs=2"),
                     (dict(nx=16, ny=16, ndust=1, geom=C.GEOM_POLAR, force_kind=C.FORCE_DISK, mplanet=1e-3),
                      b"rplanet > 0"),
                     (dict(nx=16, ny=16, ndust=1, force_kind=C.FORCE_DISK, mplanet=1e-3, rplanet=1.0),
                      b"2D polar grid"),
                     (dict(nx=16, ny=16, ndust=1, force_kick=1, force_kind=C.FORCE_SI, self_gravity=1),
                      b"self_gravity"),
                     # explicit drag: built for the shearing box (constant stopping times) only
                     (dict(nx=16, ny=16, ndust=1, geom=C.GEOM_POLAR, force_kick=1, drag=0,
                           force_kind=C.FORCE_DISK, tstop_kind=C.TSTOP_EPSTEIN_DISK), b"shearing box"),
                     (dict(nx=64, ndust=65), b"ndust=65"),
                     (dict(nx=64, ndust=1, slope_type=7), b"slope limiter"),
                     (dict(nx=64, ndust=1, viscosity=1), b"viscosity"),
                     (dict(nx=64, ndust=1, fargo=1), b"FARGO"),
                     # FARGO shifts rings of constant radius: polar grids only (the cartesian
                     # disk's hooks would otherwise run in a kernel that leaves them out)
                     (dict(nx=16, ny=16, ndust=1, fargo=1, force_kick=1, force_kind=C.FORCE_CDISK,
                           inloop_kind=C.INLOOP_CDISK), b"FARGO"),
                     (dict(nx=16, ny=16, ndust=1, fargo=1), b"polar")]:
        c = C.make_cfg(**kw)
        assert lib.shark_gpu_create(ctypes.byref(c), ctypes.byref(h)) != 0
        assert frag in lib.shark_gpu_last_error(), lib.shark_gpu_last_error()


def test_growth_create_rejects_out_of_scope_configs_without_gpu():
    """Growth parameters are validated before any device call."""
    from shark_amd import growth as G
    p = G.disk_params(nbins=16)
    p.env = G.ENV_COLLAPSE
    with pytest.raises(gpu.SharkGPUError, match="disk environment"):
        G.Growth(p, 10)
    p = G.disk_params(nbins=16, decomposed=0)
    with pytest.raises(gpu.SharkGPUError, match="decomposed"):
        G.Growth(p, 10)
    p = G.disk_params(nbins=65)
    with pytest.raises(gpu.SharkGPUError, match="nbins"):
        G.Growth(p, 10)


def test_noalloc_marchand_rejects_oversized_calls_without_gpu():
    """The size limit of shark_charge_marchand_noalloc's static scratch is checked before any
    device call."""
    from shark_amd import growth as G
    lib = gpu.load()
    lib.shark_charge_marchand_noalloc.argtypes = ([ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
                                                  + [ctypes.c_void_p] * 13)
    p = G.marchand_params()
    dummy = ctypes.c_void_p(8)   # never dereferenced: the size check comes first
    rc = lib.shark_charge_marchand_noalloc(ctypes.byref(p), (1 << 20) // 16 + 1, 16, *([dummy] * 12), None)
    assert rc != 0
    assert b"exceeds the static scratch" in lib.shark_gpu_last_error()
