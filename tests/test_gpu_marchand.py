"""Marchand grain charging on the GPU (shark_charge_marchand) against the reference's own
`charge` (tests/golden/ref_marchand.npz) and against the oracle on 4096 random cells.  The device's
exp, pow and sqrt may differ from the host libm by an ulp, so 1e-12 relative."""
import numpy as np
import pytest

import pyoracle
from shark_amd import growth as G
from test_marchand_oracle import OUT, fixture, ref_out

pytestmark = pytest.mark.gpu


def close(a, b, rtol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= rtol * np.abs(b) + 1e-300)


def test_gpu_marchand_matches_reference():
    p, z = fixture()
    args = (z["u_rho"], z["rhod"], z["T"], z["cs"], z["sdust"], z["mdust"])
    psi = None
    for call in (1, 2):
        out = G.charge_marchand(p, *args, psi_old=psi)
        ref = ref_out(z, call)
        for k in OUT:
            assert close(out[k], ref[k]), (call, k)
        psi = out["psi"]


def test_gpu_marchand_matches_oracle_on_random_cells():
    p, z = fixture()
    rng = np.random.default_rng(5)
    nc, nb = 4096, z["rhod"].shape[0]
    rho = z["u_rho"][0] * 10.0 ** rng.uniform(0, 16, nc)
    rhod = rho[None, :] * 0.01 * rng.dirichlet(np.ones(nb), nc).T
    T = 10.0 ** rng.uniform(1, 4, nc)
    cs = z["cs"][0] * np.sqrt(T / z["T"][0])
    sd = np.repeat(z["sdust"][:, :1], nc, axis=1)
    md = np.repeat(z["mdust"][:, :1], nc, axis=1)
    print(f"marchand {nc} random cells, {nb} bins", flush=True)
    for resist in (1, 0):
        p.resistivities = resist
        o = pyoracle.charge_marchand(p, rho, rhod, T, cs, sd, md)
        g = G.charge_marchand(p, rho, rhod, T, cs, sd, md)
        for k in o:
            assert close(g[k], o[k]), (resist, k)
