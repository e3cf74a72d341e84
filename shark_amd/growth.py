"""Dust growth: the parameter struct of include/shark_growth.h and the GPU binding.

``GrowthParams`` mirrors ``shark_growth_params``; ``Growth`` drives the growth kernels of
libshark_gpu.so: mode FP64 (pair loop in the reference's order, bit-exact with the FP64
oracle) or the MFMA contraction in FP32, BF16 hi+lo split, or plain BF16.
The reference routine is dust_growth (src/dust/dust_growth.f90:3-256).
"""
import ctypes as C

import numpy as np

from . import gpu

SHARK_GROWTH_VERSION = 1
MAX_BINS = 64
ENV_COLLAPSE, ENV_DISK = 0, 1
FP64, MFMA_FP32, MFMA_BF16, MFMA_BF16_PLAIN, SURROGATE = 0, 1, 2, 3, 4
MODE_NAMES = {FP64: "fp64", MFMA_FP32: "mfma_fp32", MFMA_BF16: "mfma_bf16x3", MFMA_BF16_PLAIN: "mfma_bf16",
              SURROGATE: "surrogate"}

AU = 1.5e13
MSUN = 2e33
GRAV = 6.67e-8


class GrowthParams(C.Structure):
    _fields_ = [
        ("version", C.c_int32), ("nbins", C.c_int32), ("env", C.c_int32),
        ("kernel_type", C.c_int32), ("fragmentation", C.c_int32), ("frag_mode", C.c_int32),
        ("turb", C.c_int32), ("brownian", C.c_int32), ("drift", C.c_int32),
        ("ambipolar", C.c_int32), ("decomposed", C.c_int32), ("max_substeps", C.c_int32),
        ("smin", C.c_double), ("smax", C.c_double), ("rhograin", C.c_double),
        ("ice_mantle", C.c_double), ("slope_mono", C.c_double), ("size_mono", C.c_double),
        ("abr", C.c_double), ("gamma_grains", C.c_double), ("estar_grains", C.c_double),
        ("turb_coef", C.c_double), ("vfrag", C.c_double), ("cfl_growth", C.c_double),
        ("eps_threshold", C.c_double), ("eps_threshold_frag", C.c_double),
        ("dust_ratio_min", C.c_double), ("sticking", C.c_double), ("clustered", C.c_double),
        ("delta_vambi", C.c_double), ("unit_l", C.c_double), ("unit_d", C.c_double),
        ("unit_t", C.c_double), ("unit_m", C.c_double), ("unit_v", C.c_double),
        ("mu_gas", C.c_double), ("gamma", C.c_double), ("hoverr", C.c_double),
        ("mstar", C.c_double), ("reserved", C.c_double * 8),
    ]


class MarchandParams(C.Structure):
    """struct shark_marchand_params (include/shark_growth.h)."""
    _fields_ = [
        ("zeta", C.c_double), ("mu_ions", C.c_double), ("stickeff_el", C.c_double),
        ("epsilon_ionis", C.c_double), ("nitermax_ionis", C.c_int32), ("resistivities", C.c_int32),
        ("B_0_lee", C.c_double), ("rhograin", C.c_double), ("gamma", C.c_double),
        ("unit_l", C.c_double), ("unit_d", C.c_double), ("unit_m", C.c_double), ("unit_v", C.c_double),
        ("unit_nh", C.c_double),
    ]


def marchand_params(**kw):
    """Defaults of &IONISATION_PARAMS (src/chemistry/ionisation_commons.f90:71-77), overridden by kw."""
    p = MarchandParams()
    p.zeta, p.mu_ions, p.stickeff_el, p.epsilon_ionis = 5e-17, 25.0, 0.5, 1e-6
    p.nitermax_ionis, p.resistivities = 1000, 1
    p.B_0_lee, p.rhograin, p.gamma = 3e-5, 2.3, 1.66667
    p.unit_l = p.unit_d = p.unit_m = p.unit_v = p.unit_nh = 1.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def charge_marchand(p, rho, rhod, T, cs, sdust, mdust, psi_old=None):
    """shark_charge_marchand on the GPU: arrays (nbins, ncells) / (ncells).  Returns a dict with
    zd, ni, ne, psi (the new psi_old) and, with p.resistivities, eta_o, eta_h, eta_a, sigma_o,
    sigma_p, sigma_h, gamma_d."""
    L = gpu.load()
    L.shark_charge_marchand.argtypes = [C.c_void_p, C.c_int64, C.c_int] + [_P] * 12
    nb, nc = rhod.shape
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    rho, rhod, T, cs, sdust, mdust = map(f, (rho, rhod, T, cs, sdust, mdust))
    psi = np.zeros(nc) if psi_old is None else f(psi_old).copy()
    zd = np.zeros((nb, nc))
    ni = np.zeros(nc)
    ne = np.zeros(nc)
    cond = np.zeros((6, nc))
    gd = np.zeros((nb, nc))
    gpu.check(L.shark_charge_marchand(C.byref(p), nc, nb, *[gpu._ptr(a) for a in (
        rho, rhod, T, cs, sdust, mdust, psi, zd, ni, ne, cond, gd)]))
    out = dict(zd=zd, ni=ni, ne=ne, psi=psi)
    if p.resistivities:
        out.update(eta_o=cond[0], eta_h=cond[1], eta_a=cond[2], sigma_o=cond[3], sigma_p=cond[4],
                   sigma_h=cond[5], gamma_d=gd)
    return out


def make_params(**kw):
    """Defaults of &DUST_PARAMS (src/dust/dust_commons.f90:85-143), overridden by kw."""
    p = GrowthParams()
    p.version = SHARK_GROWTH_VERSION
    p.nbins = 16
    p.env = ENV_DISK
    p.kernel_type = 0
    p.fragmentation = 0
    p.frag_mode = 0
    p.turb = p.brownian = 1
    p.drift = p.ambipolar = 0
    p.decomposed = 1
    p.max_substeps = 1000000
    p.smin, p.smax, p.rhograin, p.ice_mantle = 1e-7, 1e-2, 2.7, 8.7e-7
    p.slope_mono, p.size_mono, p.abr = 3.5, 1e-5, 2.8e3
    p.gamma_grains, p.estar_grains = 25.0, 2.8e11
    p.turb_coef = 1.5
    p.vfrag = 100.0
    p.cfl_growth, p.eps_threshold, p.eps_threshold_frag = 0.1, 1e-10, 1e-10
    p.dust_ratio_min, p.sticking, p.clustered, p.delta_vambi = 1e-11, 1.0, 1.0, 1.0
    p.unit_l = p.unit_d = p.unit_t = p.unit_m = p.unit_v = 1.0
    p.mu_gas, p.gamma = 2.31, 1.66667
    p.hoverr, p.mstar = 0.05, 1.0
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def disk_params(nbins=64, smin=1e-4, smax=10.0, alpha=1e-3, vfrag=100.0, hoverr=0.05, **kw):
    """The growth configuration of the 512x512 disk workload: bins log-spaced in size from
    smin to smax (cm), Brownian + turbulent relative velocities with turbulent speed
    sqrt(alpha)*cs, velocity-threshold fragmentation at vfrag (cm/s), disk_polar units
    (src/setups/disk_polar/units.f90)."""
    unit_l = AU
    unit_d = MSUN / AU ** 3
    unit_t = 1.0 / np.sqrt(GRAV * MSUN / unit_l ** 3)
    unit_m = MSUN
    unit_v = unit_l / unit_t
    base = dict(nbins=nbins, env=ENV_DISK, fragmentation=1, frag_mode=1, turb=1, brownian=1,
                drift=0, ambipolar=0, decomposed=1, smin=smin, smax=smax, rhograin=2.7,
                ice_mantle=0.0, turb_coef=float(np.sqrt(alpha)), vfrag=vfrag, cfl_growth=0.1,
                unit_l=unit_l, unit_d=unit_d, unit_t=unit_t, unit_m=unit_m, unit_v=unit_v,
                hoverr=hoverr, mstar=1.0)
    base.update(kw)
    return make_params(**base)


def disk_state(params, nr=512, nphi=512, r_in=0.4, r_out=2.5, sigma_r0=1700.0, eps_tot=0.01,
               noise=1e-3, seed=20261202, bins=None):
    """Growth inputs of a polar disk with nr x nphi cells (the disk setup's radial grid and
    Sigma ~ r^-1 normalised to sigma_r0 g/cm^2 at r = 1, locally isothermal cs = (H/R) v_K,
    seeded Gaussian density noise) and an MRN size distribution (mass per bin ~ s^(4-3.5),
    src/dust/dust_init.f90:164-197) of total dust-to-gas eps_tot.  `bins` are the bin centre
    sizes in code units (from the growth tables).  Returns (rhod (nbins, ncells), sigma, cs, r)
    with cells ordered radius-fastest as the step's cell index."""
    nb = params.nbins
    dr = (r_out - r_in) / nr
    rc = r_in + (np.arange(nr) + 0.5) * dr
    r = np.tile(rc, nphi)
    rng = np.random.Generator(np.random.Philox(seed))
    sig0 = sigma_r0 / (params.unit_m / params.unit_l ** 2)
    sigma = sig0 / r * (1.0 + noise * rng.standard_normal(r.size))
    cs = params.hoverr * np.sqrt(params.mstar / r)
    s = np.asarray(bins, dtype=np.float64)
    w = s ** (4.0 - 3.5) / s ** 0.0
    w = w / w.sum()
    rhod = np.ascontiguousarray((eps_tot * sigma)[None, :] * w[:, None])
    return rhod, sigma, cs, r


_P = C.POINTER(C.c_double)


def _lib():
    lib = gpu.load()
    if not getattr(lib, "_growth_bound", False):
        lib.shark_growth_create.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p)]
        lib.shark_growth_upload.argtypes = [C.c_void_p, _P, _P, _P, _P]
        lib.shark_growth_run.argtypes = [C.c_void_p, C.c_int, C.c_double]
        lib.shark_growth_download.argtypes = [C.c_void_p, _P, C.POINTER(C.c_int64)]
        lib.shark_growth_kernel_time.argtypes = [C.c_void_p, _P]
        lib.shark_growth_charge.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _P, _P]
        lib.shark_growth_set_cell_dt.argtypes = [C.c_void_p, _P]
        lib.shark_growth_set_wave_cells.argtypes = [C.c_void_p, C.c_int64]
        lib.shark_growth_features.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_float)]
        lib.shark_growth_set_surrogate.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_float),
                                                   C.c_double, C.c_double, C.c_double, C.c_double]
        lib.shark_growth_surrogate_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64), _P]
        lib.shark_growth_destroy.argtypes = [C.c_void_p]
        lib._growth_bound = True
    return lib


class Growth:
    """Device growth state: rhod (ncells, nbins) plus per-cell gas density, cs and radius."""

    def __init__(self, params, ncells):
        self.L = _lib()
        self.p = params
        self.ncells = int(ncells)
        h = C.c_void_p()
        gpu.check(self.L.shark_growth_create(C.byref(params), self.ncells, C.byref(h)))
        self.h = h

    def upload(self, rhod, rho_gas, cs, radius):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (rhod, rho_gas, cs, radius)]
        gpu.check(self.L.shark_growth_upload(self.h, *[x.ctypes.data_as(_P) for x in a]))

    def run(self, mode, dt):
        gpu.check(self.L.shark_growth_run(self.h, int(mode), float(dt)))

    def download(self):
        r = np.empty(self.p.nbins * self.ncells)
        n = np.empty(self.ncells, dtype=np.int64)
        gpu.check(self.L.shark_growth_download(self.h, r.ctypes.data_as(_P),
                                               n.ctypes.data_as(C.POINTER(C.c_int64))))
        return r.reshape(self.p.nbins, self.ncells), n

    def charge(self, zeta=5e-17, mu_ions=25.0, electrons=True):
        """Mean grain charge and ion fraction per cell of the current state."""
        zm = np.empty(self.ncells)
        xi = np.empty(self.ncells)
        gpu.check(self.L.shark_growth_charge(self.h, float(zeta), float(mu_ions), int(electrons),
                                             zm.ctypes.data_as(_P), xi.ctypes.data_as(_P)))
        return zm, xi

    def set_cell_dt(self, dt_cells):
        """Per-cell growth intervals for the following runs (None: run's dt everywhere)."""
        if dt_cells is None:
            gpu.check(self.L.shark_growth_set_cell_dt(self.h, None))
            return
        d = np.ascontiguousarray(dt_cells, dtype=np.float64)
        assert d.size == self.ncells
        self._dtc = d
        gpu.check(self.L.shark_growth_set_cell_dt(self.h, d.ctypes.data_as(_P)))

    def set_wave_cells(self, max_cells):
        """Chunks of at most max_cells cells run mode 0 one wave per cell, larger ones one lane per
        cell (shark_growth_set_wave_cells; 0: always one lane per cell)."""
        gpu.check(self.L.shark_growth_set_wave_cells(self.h, int(max_cells)))

    def features(self, dt):
        f = np.empty((self.ncells, self.p.nbins + 5), dtype=np.float32)
        gpu.check(self.L.shark_growth_features(self.h, float(dt), f.ctypes.data_as(C.POINTER(C.c_float))))
        return f

    def set_surrogate(self, net, norm, alpha=0.0, zeta=5e-17, mu_ions=25.0, resid_max=3e-3):
        """Attach a shark_amd.mlp.MLP (kept referenced here) with its normalisation vector.
        resid_max: cells whose raw predicted mass residual exceeds it run the plain step (trained
        weights carry their own calibrated value, surrogate.resid_max(meta))."""
        self._net = net
        self._norm = np.ascontiguousarray(norm, dtype=np.float32)
        gpu.check(self.L.shark_growth_set_surrogate(
            self.h, None if net is None else net.h, self._norm.ctypes.data_as(C.POINTER(C.c_float)),
            float(alpha), float(zeta), float(mu_ions), float(resid_max)))

    def set_surrogate_residual(self, net2, out_norm2):
        """Attach a residual network (its outputs, de-standardised with out_norm2 = out_mean2,
        out_std2, are added to the first network's); None detaches."""
        self.L.shark_growth_set_surrogate_residual.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]
        self._net2 = net2
        self._norm2 = None if net2 is None else np.ascontiguousarray(out_norm2, dtype=np.float32)
        gpu.check(self.L.shark_growth_set_surrogate_residual(
            self.h, None if net2 is None else net2.h,
            None if net2 is None else self._norm2.ctypes.data_as(C.POINTER(C.c_float))))

    def surrogate_stats(self):
        n = C.c_int64()
        r = C.c_double()
        gpu.check(self.L.shark_growth_surrogate_stats(self.h, C.byref(n), C.byref(r)))
        return n.value, r.value

    def zmean(self):
        """Per-cell <Z> of the last learned step (network head; charge solve where it fell back)
        or of the last charge()."""
        self.L.shark_growth_zmean.argtypes = [C.c_void_p, _P]
        z = np.empty(self.ncells)
        gpu.check(self.L.shark_growth_zmean(self.h, z.ctypes.data_as(_P)))
        return z

    def kernel_ms(self):
        ms = C.c_double()
        gpu.check(self.L.shark_growth_kernel_time(self.h, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            self.L.shark_growth_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
