"""Marchand grain charging on the GPU (shark_charge_marchand) against the reference's own
`charge` (tests/golden/ref_marchand.npz) and against the oracle on 4096 random cells.

The device's exp and pow may differ from the host libm by an ulp.  The psi equation amplifies
that where psi -> 0 (its terms cancel: the fixture's cell 8, psi = -1.9e-9, moves by 3e-11
relative), so the bound is 1e-9 of each field's scale: the value itself, the cell's largest
|value| over the bins for per-bin fields (Z_k is psi*tau_k plus an O(1) term, and near-zero
charges are cancellation residues), and the larger of |eta_a| and eta_o for eta_a (a difference of
two terms)."""
import numpy as np
import pytest

import pyoracle
from shark_amd import growth as G
from test_marchand_oracle import OUT, fixture, ref_out

pytestmark = pytest.mark.gpu


def close(k, a, b, out_ref, rtol=1e-9):
    a, b = np.asarray(a), np.asarray(b)
    if k in ("zd", "gamma_d"):
        scale = np.abs(b).max(axis=0, keepdims=True)
    elif k == "eta_a":
        scale = np.maximum(np.abs(b), np.abs(out_ref["eta_o"]))
    else:
        scale = np.abs(b)
    rel = np.abs(a - b) / np.maximum(scale, 1e-300)
    if not np.all(rel <= rtol):
        i = np.unravel_index(np.argmax(rel), rel.shape)
        print(f"{k}: max scaled diff {rel.max():.3e} at {i}: {a[i]!r} vs {b[i]!r}", flush=True)
    return bool(np.all(rel <= rtol))


def test_gpu_marchand_matches_reference():
    p, z = fixture()
    args = (z["u_rho"], z["rhod"], z["T"], z["cs"], z["sdust"], z["mdust"])
    psi = None
    for call in (1, 2):
        out = G.charge_marchand(p, *args, psi_old=psi)
        ref = ref_out(z, call)
        ok = {k: close(k, out[k], ref[k], ref) for k in OUT}
        assert all(ok.values()), (call, ok)
        psi = out["psi"]


def test_gpu_marchand_matches_oracle_on_random_cells():
    p, z = fixture()
    rng = np.random.default_rng(5)
    nc, nb = 4096, z["rhod"].shape[0]
    # six decades of density and 10-300 K keep |psi| >= 5e-5, where psi is well determined (above,
    # psi -> 0 and the charges become cancellation residues of the O(1) terms)
    rho = z["u_rho"][0] * 10.0 ** rng.uniform(0, 6, nc)
    rhod = rho[None, :] * 0.01 * rng.dirichlet(np.ones(nb), nc).T
    T = 10.0 ** rng.uniform(1, 2.5, nc)
    cs = z["cs"][0] * np.sqrt(T / z["T"][0])
    sd = np.repeat(z["sdust"][:, :1], nc, axis=1)
    md = np.repeat(z["mdust"][:, :1], nc, axis=1)
    print(f"marchand {nc} random cells, {nb} bins", flush=True)
    for resist in (1, 0):
        p.resistivities = resist
        o = pyoracle.charge_marchand(p, rho, rhod, T, cs, sd, md)
        g = G.charge_marchand(p, rho, rhod, T, cs, sd, md)
        ok = {k: close(k, g[k], o[k], o) for k in o}
        assert all(ok.values()), (resist, ok)
