"""shark_charge_marchand_noalloc (Marchand charging on caller-owned device arrays, scratch in a
static device array, no allocation) against shark_charge_marchand, which allocates its own device
memory, on the same inputs: two calls in a row, the second on fewer cells, every output bit for bit
(the same kernels run in both).  The test's own device buffers come from hipMalloc of the HIP
runtime the library is linked to."""
import ctypes as C

import numpy as np
import pytest

from shark_amd import gpu
from shark_amd import growth as G
from test_marchand_oracle import fixture

pytestmark = pytest.mark.gpu

_P = C.POINTER(C.c_double)


def inputs(nc, seed):
    p, z = fixture()
    rng = np.random.default_rng(seed)
    nb = z["rhod"].shape[0]
    rho = z["u_rho"][0] * 10.0 ** rng.uniform(0, 6, nc)
    rhod = rho[None, :] * 0.01 * rng.dirichlet(np.ones(nb), nc).T
    T = 10.0 ** rng.uniform(1, 2.5, nc)
    cs = z["cs"][0] * np.sqrt(T / z["T"][0])
    sd = np.repeat(z["sdust"][:, :1], nc, axis=1)
    md = np.repeat(z["mdust"][:, :1], nc, axis=1)
    return p, rho, rhod, T, cs, sd, md


def hip():
    gpu.load()   # the library's HIP runtime is then the one in the process
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class DevBuf:
    """A device copy of a float64 host array (hipMalloc / hipMemcpy)."""

    def __init__(self, h, a):
        self.h, self.a = h, np.ascontiguousarray(a, dtype=np.float64)
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), max(self.a.nbytes, 8)) == 0
        assert h.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, 1) == 0   # host to device

    def get(self):
        out = np.empty_like(self.a)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        self.h.hipFree(self.p)


def noalloc(p, rho, rhod, T, cs, sd, md):
    h = hip()
    L = gpu.load()
    L.shark_charge_marchand_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_int] + [C.c_void_p] * 12 + [C.c_void_p]
    nb, nc = rhod.shape
    bufs = [DevBuf(h, a) for a in (rho, rhod, T, cs, sd, md, np.zeros(nc), np.zeros((nb, nc)), np.zeros(nc),
                                   np.zeros(nc), np.zeros((6, nc)), np.zeros((nb, nc)))]
    try:
        rc = L.shark_charge_marchand_noalloc(C.byref(p), nc, nb, *[b.p for b in bufs], None)
        gpu.check(rc)
        psi, zd, ni, ne, cond, gd = [b.get() for b in bufs[6:]]   # blocking copies on the default stream
    finally:
        for b in bufs:
            b.free()
    out = dict(zd=zd, ni=ni, ne=ne, psi=psi)
    if p.resistivities:
        out.update(eta_o=cond[0], eta_h=cond[1], eta_a=cond[2], sigma_o=cond[3], sigma_p=cond[4],
                   sigma_h=cond[5], gamma_d=gd)
    return out


@pytest.mark.parametrize("resist", [1, 0])
def test_noalloc_marchand_equals_allocating_entry_point(resist):
    for call, nc in ((1, 8192), (2, 3000)):
        p, rho, rhod, T, cs, sd, md = inputs(nc, seed=10 + call)
        p.resistivities = resist
        print(f"marchand no-alloc call {call}: {nc} cells x {rhod.shape[0]} bins, resistivities {resist}",
              flush=True)
        want = G.charge_marchand(p, rho, rhod, T, cs, sd, md)
        got = noalloc(p, rho, rhod, T, cs, sd, md)
        assert set(got) == set(want)
        for k in want:
            assert np.array_equal(got[k], want[k]), (call, k)


def test_noalloc_marchand_rejects_more_pairs_than_its_scratch():
    L = gpu.load()
    L.shark_charge_marchand_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_int] + [C.c_void_p] * 12 + [C.c_void_p]
    p = G.marchand_params()
    dummy = C.c_void_p(8)   # never dereferenced: the size check comes first
    rc = L.shark_charge_marchand_noalloc(C.byref(p), (1 << 20) // 16 + 1, 16, *([dummy] * 12), None)
    assert rc != 0
    assert b"exceeds the static scratch" in L.shark_gpu_last_error()
