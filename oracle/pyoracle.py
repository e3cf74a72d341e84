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
    lib.oracle_set_dvol.argtypes = [C.c_void_p, _P]
    lib.oracle_barotrop.restype = C.c_double
    lib.oracle_barotrop.argtypes = [C.c_void_p, C.c_double]
    lib.oracle_cs_eos.restype = C.c_double
    lib.oracle_cs_eos.argtypes = [C.c_void_p, C.c_double]
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
    lib.oracle_fargo_remap.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _P, _P, _P,
                                       _P, _P]
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


def fargo_remap(u, nx, ny, slope, dt, dx2, radii_c, fargo_v):
    """The oracle's FARGO remap alone (the code its step runs, fargo_scheme.f90:17-77) on a 2D
    state u of shape (nvar, (ny+4)*(nx+4)); dx2, radii_c and fargo_v are per cell (ncells,).
    The y-ghost rows of u are taken as periodic copies; cells outside the active block are
    returned unchanged."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    nvar = u.shape[0]
    assert u.size == nvar * (nx + 4) * (ny + 4)
    out = np.empty_like(u)
    c = [np.ascontiguousarray(x, dtype=np.float64) for x in (dx2, radii_c, fargo_v)]
    for x in c:
        assert x.size == (nx + 4) * (ny + 4)
    lib().oracle_fargo_remap(nx, ny, nvar, slope, float(dt), _ptr(c[0]), _ptr(c[1]), _ptr(c[2]),
                             _ptr(u), _ptr(out))
    return out


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
        if g.get("dvol") is not None:
            self.L.oracle_set_dvol(self.h, _ptr(g["dvol"]))
        s = {k: (None if v is None else np.ascontiguousarray(v, dtype=np.float64))
             for k, v in state.items()}
        self.L.oracle_upload(self.h, _ptr(s["u"]), _ptr(s["cs"]), _ptr(s.get("sdust")),
                             _ptr(s.get("condinit")), _ptr(s.get("fargo_velocity")))

    def set_growth(self, params):
        """Dust growth (disk environment, FP64) after the drag of every step."""
        self.L.oracle_set_growth.argtypes = [C.c_void_p, C.c_void_p]
        self._gp = params
        if self.L.oracle_set_growth(self.h, None if params is None else C.byref(params)) != 0:
            raise ValueError("growth bins must equal the number of dust species")

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


# ----------------------------------------------------------------------------- growth
class GrowthTables(C.Structure):
    _fields_ = [("nb", C.c_int), ("zeta", C.c_double), ("eta", C.c_double),
                ("massmin", C.c_double), ("m_mono", C.c_double), ("ebr_mono", C.c_double),
                ("aplus", C.c_double * 64), ("aminus", C.c_double * 64),
                ("mminus", C.c_double * 64), ("sdust", C.c_double * 64),
                ("mdust", C.c_double * 64), ("redis", C.c_double * 4096)]


def _glib():
    L = lib()
    if not getattr(L, "_growth_bound", False):
        L.growth_oracle_tables.argtypes = [C.c_void_p, C.c_void_p]
        L.growth_oracle_collapse.restype = C.c_int64
        L.growth_oracle_collapse.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                             C.c_int64] + [_P] * 7 + [C.c_void_p, C.c_double, _P,
                                                                      C.POINTER(C.c_int64)]
        L.growth_oracle_disk.restype = C.c_int64
        L.growth_oracle_disk.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, _P, _P, _P,
                                         C.c_double, _P, C.POINTER(C.c_int64), _P]
        L._growth_bound = True
    return L


def growth_tables(params):
    t = GrowthTables()
    _glib().growth_oracle_tables(C.byref(params), C.byref(t))
    return t


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def growth_collapse(params, tables, T0, unit_nH, q_rho, u_rho, sdust, mdust, tstop, vd, eta_a,
                    active, dt, rhod):
    """Reference collapse branch of dust_growth on ncells cells (arrays (nbins, ncells))."""
    nc = q_rho.size
    r = _c(rhod).copy()
    n = np.zeros(nc, dtype=np.int64)
    act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
    args = [_c(q_rho), _c(u_rho), _c(sdust), _c(mdust), _c(tstop), _c(vd), _c(eta_a)]
    _glib().growth_oracle_collapse(C.byref(params), C.byref(tables), T0, unit_nH, nc,
                                   *[_ptr(a) for a in args],
                                   None if act is None else act.ctypes.data_as(C.c_void_p),
                                   dt, _ptr(r), n.ctypes.data_as(C.POINTER(C.c_int64)))
    return r, n


def growth_disk(params, tables, sigma, cs, radius, dt, rhod, first_rates=False):
    """Disk-environment growth over dt for every cell; returns (rhod, substeps[, rates])."""
    nc = sigma.size
    r = _c(rhod).copy()
    n = np.zeros(nc, dtype=np.int64)
    d1 = np.zeros_like(r) if first_rates else None
    _glib().growth_oracle_disk(C.byref(params), C.byref(tables), nc, _ptr(_c(sigma)),
                               _ptr(_c(cs)), _ptr(_c(radius)), dt, _ptr(r),
                               n.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(d1))
    return (r, n, d1) if first_rates else (r, n)


def growth_disk_env(params, tables, sigma, cs, r):
    """Per-cell inputs of the disk environment: T, t_L, t_eta, Re and the stopping times."""
    L = _glib()
    nb = tables.nb
    out = [C.c_double() for _ in range(4)]
    ts = np.zeros(nb)
    L.growth_oracle_disk_env.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                         C.c_double] + [C.POINTER(C.c_double)] * 4 + [_P]
    L.growth_oracle_disk_env(C.byref(params), C.byref(tables), sigma, cs, r,
                             *[C.byref(o) for o in out], _ptr(ts))
    return dict(T=out[0].value, t_L=out[1].value, t_eta=out[2].value, Re=out[3].value, tstop=ts)


def charge_collapse(q_rho, rhod, mdust, sdust, T0, zeta, mu_ions, mu_gas, electrons=True,
                    maxiter=100000):
    """analytical_charge over collapse cells; arrays (nbins, ncells). Returns zd, ni, ne."""
    L = _glib()
    nb, nc = rhod.shape
    zd = np.zeros((nb, nc))
    ni = np.zeros(nc)
    ne = np.zeros(nc)
    L.growth_oracle_charge_collapse.argtypes = [C.c_int64, C.c_int] + [_P] * 4 + [
        C.c_double] * 4 + [C.c_int, C.c_int] + [_P] * 3
    L.growth_oracle_charge_collapse(nc, nb, _ptr(_c(q_rho)), _ptr(_c(rhod)), _ptr(_c(mdust)),
                                    _ptr(_c(sdust)), T0, zeta, mu_ions, mu_gas, int(electrons),
                                    maxiter, _ptr(zd), _ptr(ni), _ptr(ne))
    return zd, ni, ne


def charge_root(zeta, rho, m_d, rho_d, s_d, T, mu_ions=25.0, mu_gas=2.31, electrons=True,
                maxiter=100000, steptol=0.0):
    L = _glib()
    L.growth_oracle_charge_root.restype = C.c_double
    L.growth_oracle_charge_root.argtypes = [C.c_double] * 8 + [C.c_int, C.c_int, C.c_double,
                                                               C.POINTER(C.c_int)]
    it = C.c_int()
    z = L.growth_oracle_charge_root(zeta, rho, m_d, rho_d, s_d, T, mu_ions, mu_gas, int(electrons),
                                    maxiter, steptol, C.byref(it))
    return z, it.value


def charge_disk(params, tables, sigma, cs, r, rhod, zeta=5e-17, mu_ions=25.0, electrons=True,
                maxiter=100, per_bin=False):
    """Disk-environment grain charges: (zmean, xion[, zd])."""
    L = _glib()
    nb, nc = rhod.shape
    zd = np.zeros((nb, nc))
    zm = np.zeros(nc)
    xi = np.zeros(nc)
    L.growth_oracle_charge_disk.argtypes = [C.c_void_p, C.c_void_p, C.c_int64] + [_P] * 4 + [
        C.c_double, C.c_double, C.c_int, C.c_int] + [_P] * 3
    L.growth_oracle_charge_disk(C.byref(params), C.byref(tables), nc, _ptr(_c(sigma)), _ptr(_c(cs)),
                                _ptr(_c(r)), _ptr(_c(rhod)), zeta, mu_ions, int(electrons), maxiter,
                                _ptr(zd), _ptr(zm), _ptr(xi))
    return (zm, xi, zd) if per_bin else (zm, xi)


def charge_marchand(p, rho, rhod, T, cs, sdust, mdust, psi_old=None):
    """Marchand grain charging (ionisation.f90:4-214) on the CPU; p: shark_amd.growth.MarchandParams.
    Arrays (nbins, ncells) / (ncells); returns the same dict as shark_amd.growth.charge_marchand."""
    L = _glib()
    nb, nc = rhod.shape
    psi = np.zeros(nc) if psi_old is None else np.array(psi_old, dtype=np.float64)
    zd = np.zeros((nb, nc))
    ni = np.zeros(nc)
    ne = np.zeros(nc)
    cond = np.zeros((6, nc))
    gd = np.zeros((nb, nc))
    L.marchand_oracle_charge.argtypes = [C.c_void_p, C.c_int64, C.c_int] + [_P] * 12
    L.marchand_oracle_charge(C.byref(p), nc, nb, _ptr(_c(rho)), _ptr(_c(rhod)), _ptr(_c(T)), _ptr(_c(cs)),
                             _ptr(_c(sdust)), _ptr(_c(mdust)), _ptr(psi), _ptr(zd), _ptr(ni), _ptr(ne),
                             _ptr(cond), _ptr(gd))
    out = dict(zd=zd, ni=ni, ne=ne, psi=psi)
    if p.resistivities:
        out.update(eta_o=cond[0], eta_h=cond[1], eta_a=cond[2], sigma_o=cond[3], sigma_p=cond[4],
                   sigma_h=cond[5], gamma_d=gd)
    return out
