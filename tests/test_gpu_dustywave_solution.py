"""GPU dusty waves of the reference's test suite (K = 50, 100, 1000 to t = 4.5): bit-exact with the
oracle and within the band of the test-suite solution files that the oracle meets."""
import numpy as np
import pytest

from test_dustywave_solution import BAND, errors, wave

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K", [50, 100, 1000])
def test_gpu_dustywave_against_solution_file(K):
    import pyoracle
    from shark_amd import gpu
    cfg, g, st = wave(K)
    print(f"dusty wave K={K} to t=4.5", flush=True)
    o = pyoracle.Oracle(cfg, g, st)
    n_o = o.run(1000000, 4.5)
    s = gpu.Step(cfg, g, st)
    n_g, t_g = s.run(1000000, 4.5)
    u = s.download()
    s.close()
    assert (n_g, t_g) == (n_o, o.time)
    np.testing.assert_array_equal(u[:, 2:130], o.u()[:, 2:130])
    err = errors(u, K)
    for k, b in BAND.items():
        assert err[k] < b, (K, k, err[k])
