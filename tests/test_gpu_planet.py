"""The polar disk's planet and self-gravity (disk_polar/setup.f90:300-378: the softened potential
of a planet on a circular orbit, or the disk's own potential summed directly over its cells, the
reference's ghost values of apply_boundaries_phi, the centred gradient in every fluid's force) on
the GPU.  The oracle is bit-exact with the flang-built reference on the fixtures
(tests/test_oracle_ref.py, ref_disk_2d_planet*, ref_disk_2d_selfgrav*); here the GPU is bit-exact
with the oracle on larger disks, with and without FARGO and dust."""
import numpy as np
import pytest

from refcases import active_mask
from shark_amd import setups

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fargo,nd", [(False, 4), (True, 4), (False, 0)])
def test_gpu_planet_matches_oracle(fargo, nd):
    import pyoracle
    from shark_amd import gpu
    cfg, g, st = setups.disk(nr=128, nphi=192, nd=nd, fargo=fargo)
    cfg.mplanet, cfg.rplanet, cfg.a_smooth_pl = 3e-3, 1.3, 0.05
    print(f"polar disk 128x192 nd={nd} fargo={fargo} with a planet: 30 steps, GPU vs oracle", flush=True)
    o = pyoracle.Oracle(cfg, g, st)
    dts_o = [o.step() for _ in range(30)]
    s = gpu.Step(cfg, g, st)
    try:
        dts = [s.step() for _ in range(30)]
        u = s.download()
    finally:
        s.close()
    m = active_mask(cfg)
    print(f"  dt bit-exact {dts == dts_o}; state bit-exact {np.array_equal(u[:, m], o.u()[:, m])}", flush=True)
    np.testing.assert_array_equal(dts, dts_o)
    np.testing.assert_array_equal(u[:, m], o.u()[:, m])


@pytest.mark.parametrize("fargo,nd", [(False, 2), (True, 2), (False, 0)])
def test_gpu_self_gravity_matches_oracle(fargo, nd):
    import pyoracle
    from shark_amd import gpu
    cfg, g, st = setups.disk(nr=48, nphi=64, nd=nd, fargo=fargo)
    cfg.self_gravity, cfg.a_smooth_pl = 1, 0.1
    print(f"polar disk 48x64 nd={nd} fargo={fargo} with self-gravity: 20 steps, GPU vs oracle", flush=True)
    o = pyoracle.Oracle(cfg, g, st)
    dts_o = [o.step() for _ in range(20)]
    s = gpu.Step(cfg, g, st)
    try:
        dts = [s.step() for _ in range(20)]
        u = s.download()
    finally:
        s.close()
    m = active_mask(cfg)
    print(f"  dt bit-exact {dts == dts_o}; state bit-exact {np.array_equal(u[:, m], o.u()[:, m])}", flush=True)
    np.testing.assert_array_equal(dts, dts_o)
    np.testing.assert_array_equal(u[:, m], o.u()[:, m])


def test_gpu_self_gravity_graph_replay_matches_steps():
    """The step's hipGraph (shark_gpu_run) replays the self-gravity kernels and the dt taken after
    them: 15 steps by run() equal 15 single steps bit for bit."""
    from shark_amd import gpu
    cfg, g, st = setups.disk(nr=48, nphi=64, nd=2, fargo=True)
    cfg.self_gravity, cfg.a_smooth_pl = 1, 0.1
    out = []
    for by_run in (False, True):
        s = gpu.Step(cfg, g, st)
        try:
            if by_run:
                n, t = s.run(15)
                assert n == 15
            else:
                for _ in range(15):
                    s.step()
                t = s.time
            out.append((s.download(), t))
        finally:
            s.close()
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])
