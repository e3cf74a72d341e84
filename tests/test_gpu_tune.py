"""Launch shapes and the autotuner (shark_gpu_set_tune / shark_gpu_autotune).

Every shape must give the same state bit for bit: the hydro strips re-read their halo rows and
compute each face once per owning strip, and the post pass keeps the drag sums in species order
in one thread whatever the segment width.  The autotuner must leave the run exactly where it was.
"""
import numpy as np
import pytest

import pyoracle
from refcases import active_mask
from shark_amd import gpu, setups

pytestmark = pytest.mark.gpu


def _run(case, nsteps, **shape):
    g = gpu.Step(*case)
    try:
        if shape:
            g.set_tune(**shape)
            got = g.get_tune()
            for k, v in shape.items():
                assert got[k] == v
        dts = [g.step() for _ in range(nsteps)]
        return np.array(dts), g.download()
    finally:
        g.close()


@pytest.mark.parametrize("which", ["disk16", "si8", "disk64"])
def test_every_shape_is_bit_identical(which):
    if which == "disk16":
        case, tws, n = setups.disk(nr=96, nphi=128, nd=16), (16, 32), 6
    elif which == "si8":
        case, tws, n = setups.si_box(n=64, nd=8), (16, 32), 6
    else:
        case, tws, n = setups.disk(nr=64, nphi=96, nd=64), (8, 16), 3
    print(f"shapes {which}", flush=True)
    cfg, geom, state = case
    o = pyoracle.Oracle(cfg, geom, state)
    dto = np.array([o.step() for _ in range(n)])
    m = active_mask(cfg)
    uo = o.u()
    for strip, sg in ((4, 8), (13, 13), (32, 4), (200, 200)):
        for tw in tws:
            for xcd in (0, 1):
                dts, u = _run(case, n, strip_h=strip, strip_gas=sg, post_tw=tw, xcd_order=xcd)
                assert np.array_equal(dts, dto), (strip, sg, tw, xcd)
                assert np.array_equal(u[:, m], uo[:, m]), (strip, sg, tw, xcd)


def test_autotune_restores_the_run():
    case = setups.disk(nr=96, nphi=128, nd=16)
    dts_ref, u_ref = _run(case, 5)
    g = gpu.Step(*case)
    try:
        t = g.autotune(reps=3)
        print("chosen", t, flush=True)
        assert t["strip_h"] >= 1 and t["strip_gas"] >= 1 and t["post_tw"] in (16, 32) and t["xcd_order"] in (0, 1)
        assert t["hydro_ms"] > 0 and t["post_ms"] > 0
        assert g.get_tune()["strip_h"] == t["strip_h"]
        assert g.time == 0.0
        dts = np.array([g.step() for _ in range(5)])
        u = g.download()
    finally:
        g.close()
    assert np.array_equal(dts, dts_ref)
    assert np.array_equal(u, u_ref)


def test_autotune_one_dimensional():
    case = setups.dustywave(nx=512)
    dts_ref, u_ref = _run(case, 20)
    g = gpu.Step(*case)
    try:
        t = g.autotune(reps=2)
        assert t["strip_h"] == 32 and t["strip_gas"] == 32   # a 1D run has one strip: not swept
        dts = np.array([g.step() for _ in range(20)])
        u = g.download()
    finally:
        g.close()
    assert np.array_equal(dts, dts_ref) and np.array_equal(u, u_ref)


def test_set_tune_rejects_unbuilt_shapes():
    g = gpu.Step(*setups.disk(nr=32, nphi=32, nd=16))
    try:
        with pytest.raises(gpu.SharkGPUError, match="post_tw"):
            g.set_tune(post_tw=8)
        with pytest.raises(gpu.SharkGPUError, match="strip_h"):
            g.set_tune(strip_h=0)
        with pytest.raises(gpu.SharkGPUError, match="strip_gas"):
            g.set_tune(strip_gas=5000)
        with pytest.raises(gpu.SharkGPUError, match="xcd_order"):
            g.set_tune(xcd_order=2)
        assert g.get_tune()["post_tw"] == 32   # unchanged after the refusals
    finally:
        g.close()


def test_kernel_timers_without_growth():
    g = gpu.Step(*setups.disk(nr=64, nphi=64, nd=4))
    try:
        g.set_timing(True)
        g.run(5)
        for k in ("bc", "hydro", "post", "dt"):
            ms, n = g.kernel_time(k)
            assert n == 5 and ms > 0, (k, ms, n)
        assert g.kernel_time("growth") == (0.0, 0)
    finally:
        g.close()
