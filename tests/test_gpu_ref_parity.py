"""GPU step vs the flang-built reference (fixtures tests/golden/ref_*.npz), bit-exact.

Each case uploads the reference's own initial state and grid to libshark_gpu, takes the
same number of steps, and requires every dt and every active-cell value of u_prim to equal
the reference's bit for bit."""
import numpy as np
import pytest

from refcases import REF_CASES, load_ref, active_mask

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", REF_CASES)
def test_gpu_matches_reference(name):
    from shark_amd import gpu
    cfg, geom, state, final, dts, meta = load_ref(name)
    print(f"case {name}: {cfg.nx}x{cfg.ny} nd={cfg.ndust} steps={len(dts)}", flush=True)
    g = gpu.Step(cfg, geom, state)
    got = np.array([g.step() for _ in range(len(dts))])
    np.testing.assert_array_equal(got, dts)
    u = g.download()
    m = active_mask(cfg)
    np.testing.assert_array_equal(u[:, m], final[:, m])
    g.close()
