"""GPU dust growth vs the FP64 oracle (oracle/growth_oracle.c, pinned to the reference's
dust_growth in tests/test_growth_oracle.py).

Mode FP64 must agree bit for bit (densities and sub-step counts).  The MFMA modes are
reported per field against the FP64 oracle: mass-weighted size-distribution L1 error
sum_k |rho_k - rho_k,ref| / sum_k rho_k,ref per cell, with tolerances written here; their
per-cell mass rescale must conserve total dust mass to FP64 round-off (1e-13)."""
import numpy as np
import pytest

import pyoracle
from shark_amd import growth as G

pytestmark = pytest.mark.gpu


def cells(nb=64, nr=64, nphi=8, **kw):
    p = G.disk_params(nbins=nb, max_substeps=100000, **kw)
    t = pyoracle.growth_tables(p)
    rhod, sigma, cs, r = G.disk_state(p, nr=nr, nphi=nphi, bins=np.array(t.sdust[:nb]))
    return p, t, rhod, sigma, cs, r


def run_gpu(p, rhod, sigma, cs, r, mode, dt, wave_cells=None):
    g = G.Growth(p, sigma.size)
    try:
        if wave_cells is not None:
            g.set_wave_cells(wave_cells)
        g.upload(rhod, sigma, cs, r)
        g.run(mode, dt)
        return g.download()
    finally:
        g.close()


# mode 0's two launch shapes: one lane per cell (k_gfp64) and one wave per cell (k_gfp64w)
SHAPES = [pytest.param(0, id="lane-per-cell"), pytest.param(1 << 40, id="wave-per-cell")]


def l1_mass(a, ref):
    return np.abs(a - ref).sum(axis=0) / ref.sum(axis=0)


@pytest.mark.parametrize("nb,kw,dt", [
    (64, {}, 2e-8),                       # velocity-threshold fragmentation (stiff)
    (64, dict(frag_mode=0), 1e4),         # energy-threshold fragmentation
    (33, dict(fragmentation=0), 1e4),
    (16, dict(kernel_type=1, fragmentation=0), 1e-38),
    (16, dict(kernel_type=2, fragmentation=0), 1e-8),
    (24, dict(fragmentation=0, decomposed=0, eps_threshold=1e-3), 1e4),
])
@pytest.mark.parametrize("wave_cells", SHAPES)
def test_fp64_bit_exact(nb, kw, dt, wave_cells):
    p, t, rhod, sigma, cs, r = cells(nb=nb, **kw)
    ref, nref = pyoracle.growth_disk(p, t, sigma, cs, r, dt, rhod)
    out, n = run_gpu(p, rhod, sigma, cs, r, G.FP64, dt, wave_cells)
    print(f"nb={nb} {kw} dt={dt} wave_cells={wave_cells} substeps {nref.min()}..{nref.max()}", flush=True)
    assert nref.max() >= 2
    assert np.array_equal(n, nref)
    assert np.array_equal(out, ref), np.abs(out / ref - 1).max()


@pytest.mark.parametrize("wave_cells", SHAPES)
def test_fp64_chunked_relaunch_matches(wave_cells):
    """Many sub-steps per cell (the first launch of the one-lane kernel runs one sub-step, the
    relaunches resume every unfinished cell)."""
    p, t, rhod, sigma, cs, r = cells(nb=32, nr=16, nphi=4)
    dt = 4e-7
    ref, nref = pyoracle.growth_disk(p, t, sigma, cs, r, dt, rhod)
    assert nref.max() > 64
    out, n = run_gpu(p, rhod, sigma, cs, r, G.FP64, dt, wave_cells)
    assert np.array_equal(n, nref)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("mode,tol", [(G.MFMA_FP32, 1e-7), (G.MFMA_BF16, 2e-6), (G.MFMA_BF16_PLAIN, 2e-3)])
def test_mfma_error_vs_fp64_oracle(mode, tol):
    p, t, rhod, sigma, cs, r = cells(nb=64, nr=64, nphi=8)
    dt = 2e-8
    ref, _ = pyoracle.growth_disk(p, t, sigma, cs, r, dt, rhod)
    out, n = run_gpu(p, rhod, sigma, cs, r, mode, dt)
    err = l1_mass(out, ref)
    change = l1_mass(ref, rhod)
    mass = np.abs(out.sum(axis=0) / rhod.sum(axis=0) - 1)
    print(f"{G.MODE_NAMES[mode]}: L1 err mean {err.mean():.3e} max {err.max():.3e} "
          f"(change {change.mean():.3e}); mass residual max {mass.max():.3e}; "
          f"substeps {n.min()}..{n.max()}")
    assert change.mean() > 1e-3          # the step does something measurable
    assert err.max() <= tol
    assert mass.max() <= 1e-13
    assert (out > 0).all()


def test_charge_matches_oracle():
    """Grain charging on the GPU vs the oracle (same Newton iteration; exp may differ in the last
    bit between the device and host libraries, so 1e-12 relative)."""
    p, t, rhod, sigma, cs, r = cells(nb=64, nr=32, nphi=4)
    zm_ref, xi_ref = pyoracle.charge_disk(p, t, sigma, cs, r, rhod)
    g = G.Growth(p, sigma.size)
    try:
        g.upload(rhod, sigma, cs, r)
        zm, xi = g.charge()
    finally:
        g.close()
    assert (zm_ref < 0).all()
    np.testing.assert_allclose(zm, zm_ref, rtol=1e-12)
    np.testing.assert_allclose(xi, xi_ref, rtol=1e-12)


def test_feature_ion_fraction_matches_charge_solve():
    """The learned step's ion-fraction feature (regrouped Newton, warm-started per bin) agrees with
    the reference-order charge solve far below the feature's FP32 rounding, on the initial MRN
    state and on a state aged by the plain coagulation."""
    p, t, rhod, sigma, cs, r = cells(nb=64, nr=64, nphi=8)
    g = G.Growth(p, sigma.size)
    try:
        for aged in (False, True):
            g.upload(rhod, sigma, cs, r)
            if aged:
                g.run(G.MFMA_FP32, 1e-3)
                s2, _ = g.download()
                g.upload(s2, sigma, cs, r)
            _, xi = g.charge()
            f = g.features(1e-3)
            lx = f[:, p.nbins + 3].astype(np.float64)
            ref = np.log10(xi)
            assert np.all(np.isfinite(lx))
            assert np.abs(lx - ref).max() <= 4 * np.spacing(np.abs(ref).astype(np.float32)).max(), aged
    finally:
        g.close()
