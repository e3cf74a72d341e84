"""The GPU's FARGO remap (the remap phase of the post kernel) on the 1024 x 2048, 16-species disk.

Two contexts start from the same state, one with FARGO and one without, with drag, forces and
the in-loop hook switched off: their post passes then differ by the remap alone, so one step of
the FARGO context must equal the oracle's remap (fargo_scheme.f90:17-77, pinned by
tests/test_fargo_oracle.py) applied to one step of the other, bit for bit.  Integer shifts must
give the cyclic permutation of each ring in the reference's direction, and the ring sums are
conserved.
"""
import copy
import math

import numpy as np
import pytest

import pyoracle
from refcases import FARGO_REF_CASES, load_ref, run_fargo_fixture

pytestmark = pytest.mark.gpu


def remap_only(cfg, fargo):
    c = copy.copy(cfg)
    c.fargo = 1 if fargo else 0
    c.drag = 0
    c.force_kick = 0
    c.force_kind = 0
    c.inloop_kind = 0
    return c


def active(u, cfg):
    return u.reshape(u.shape[0], cfg.ny + 4, cfg.nx + 4)[:, 2:cfg.ny + 2, 2:cfg.nx + 2]


def one_step(cfg, geom, state):
    from shark_amd import gpu
    g = gpu.Step(cfg, geom, state)
    dt = g.step()
    u = g.download()
    g.close()
    return u, dt


def integer_shift_velocities(cfg, geom, dt, seed):
    """Per-column shift velocities with |vk dt| = m dy exactly (m a power of two below ny, so that
    m dy is exact; a random sign), so nshft = m and eps = 0; columns where no double near
    m dy / dt gives that product for any m keep vk = 0 (the identity).  Returns per-cell velocities and per-column m*sign."""
    nxm = cfg.nx + 4
    rng = np.random.default_rng(seed)
    dy = geom["dx"][1][:nxm] * geom["radii_c"][:nxm]
    vk = np.zeros(nxm)
    shift = np.zeros(nxm, dtype=np.int64)
    for ix in range(2, cfg.nx + 2):
        sgn = 1.0 if rng.random() < 0.5 else -1.0
        for m in rng.permutation([1 << e for e in range(11)]):
            m = int(m)
            target = m * dy[ix]
            v = target / dt
            for _ in range(64):
                sh = abs(v * dt)
                if sh == target and int(sh / dy[ix]) == m and math.fmod(sh, dy[ix]) == 0.0:
                    break
                v = float(np.nextafter(v, math.inf if sh < target else -math.inf))
            else:
                continue
            vk[ix], shift[ix] = sgn * v, int(sgn) * m
            break
    return np.tile(vk, cfg.ny + 4), shift


def pair(cfg0, geom, state, vk):
    st = dict(state)
    st["fargo_velocity"] = vk
    u0, dt0 = one_step(remap_only(cfg0, False), geom, st)
    u1, dt1 = one_step(remap_only(cfg0, True), geom, st)
    assert dt0 == dt1
    return u0, u1, dt0


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_gpu_remap_equals_oracle_remap_full_size(sign):
    from shark_amd import setups
    cfg, geom, state = setups.disk()
    print(f"disk 1024x2048 nd=16, remap only, vk sign {sign}", flush=True)
    vk = sign * state["fargo_velocity"]
    u0, u1, dt = pair(cfg, geom, state, vk)
    want = pyoracle.fargo_remap(u0, cfg.nx, cfg.ny, cfg.slope_type, dt, geom["dx"][1],
                                geom["radii_c"], vk)
    a1, aw, a0 = active(u1, cfg), active(want, cfg), active(u0, cfg)
    bad = np.argwhere(a1 != aw)
    assert bad.size == 0, f"{len(bad)} values differ from the oracle's remap, first {bad[:5]}"
    # ring sums (over phi, per radius and variable) are conserved
    s0, s1 = a0.sum(axis=1), a1.sum(axis=1)
    assert np.all(np.abs(s1 - s0) <= 1e-14 * np.abs(a0).sum(axis=1))
    assert not np.array_equal(a1, a0)


@pytest.mark.parametrize("slope", [2, 3])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_gpu_remap_equals_oracle_remap_other_limiters(slope, sign):
    from shark_amd import setups
    cfg, geom, state = setups.disk(nr=256, nphi=512, nd=16, slope_type=slope)
    print(f"disk 256x512 nd=16, remap only, slope {slope}, vk sign {sign}", flush=True)
    vk = sign * state["fargo_velocity"]
    u0, u1, dt = pair(cfg, geom, state, vk)
    want = pyoracle.fargo_remap(u0, cfg.nx, cfg.ny, cfg.slope_type, dt, geom["dx"][1],
                                geom["radii_c"], vk)
    assert np.array_equal(active(u1, cfg), active(want, cfg))


def test_gpu_integer_shift_is_cyclic_permutation_full_size():
    from shark_amd import setups
    cfg, geom, state = setups.disk()
    print("disk 1024x2048 nd=16, integer shifts per ring", flush=True)
    # the first step's dt does not depend on the shift velocity
    _, dt = one_step(remap_only(cfg, False), geom, state)
    vk, shift = integer_shift_velocities(cfg, geom, dt, seed=7)
    assert np.count_nonzero(shift) > 0.75 * cfg.nx
    u0, u1, dt1 = pair(cfg, geom, state, vk)
    assert dt1 == dt
    a0, a1 = active(u0, cfg), active(u1, cfg)
    for i in range(cfg.nx):
        # vk > 0: cell iy takes cell iy - m (fargo_scheme.f90:49); vk < 0: cell iy + m (:62)
        want = np.roll(a0[:, :, i], int(shift[i + 2]), axis=1)
        assert np.array_equal(a1[:, :, i], want), (i, int(shift[i + 2]))


@pytest.mark.parametrize("name", FARGO_REF_CASES)
def test_gpu_fargo_matches_reference(name):
    """The flang-built reference's fargo_scheme (with `one` = 1 and double-precision implicit
    reals): integer shifts set per step for six full steps, and the setup's own velocity for one
    step on the cells where the reference's remap is defined; every dt and value bit for bit."""
    from shark_amd import gpu

    def make(c, g, s):
        return gpu.Step(c, g, s)

    print(f"FARGO reference fixture {name}", flush=True)
    dts, u, final, mask = run_fargo_fixture(name, make, lambda x: x.download())
    np.testing.assert_array_equal(dts, load_ref(name)[4])
    cfg = load_ref(name)[0]
    nxm = cfg.nx + 4
    act = np.zeros((cfg.ny + 4, nxm), bool)
    act[2:cfg.ny + 2, 2:cfg.nx + 2] = True
    m = mask & act.reshape(-1)
    bad = np.argwhere(u[:, m] != final[:, m])
    assert bad.size == 0, f"{len(bad)} values differ, first {bad[:5]}"
