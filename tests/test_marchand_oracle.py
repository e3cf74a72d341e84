"""The Marchand grain-charging oracle (oracle/marchand_oracle.c) against the flang-built
reference's own `charge` (src/chemistry/ionisation.f90:4-214) on tests/golden/ref_marchand.npz:
twelve Collapse_1D cells whose densities span twenty decades (cold neutral to the low-temperature
branch), 16 grain sizes, called twice (cold start, then from psi_old).  Bit for bit."""
import json
import os

import numpy as np
import pytest

import pyoracle
from refcases import GOLDEN
from shark_amd import growth as G

OUT = ("zd", "ni", "ne", "psi", "eta_o", "eta_h", "eta_a", "sigma_o", "sigma_p", "sigma_h", "gamma_d")


def fixture():
    z = np.load(os.path.join(GOLDEN, "ref_marchand.npz"))
    m = json.loads(str(z["meta"]))
    p = G.marchand_params(zeta=m["zeta"], mu_ions=m["mu_ions"], stickeff_el=m["stickeff_el"],
                          epsilon_ionis=m["epsilon_ionis"], nitermax_ionis=m["nitermax_ionis"],
                          resistivities=int(m["res_Marchand"]), B_0_lee=m["B_0_lee"], rhograin=m["rhograin"],
                          gamma=m["gamma"], unit_l=m["unit_l"], unit_d=m["unit_d"], unit_m=m["unit_m"],
                          unit_v=m["unit_v"], unit_nh=m["unit_nh"])
    return p, z


def ref_out(z, call):
    return {k: z[f"{k}_{call}"] for k in OUT}


def test_marchand_oracle_matches_reference_bit_for_bit():
    p, z = fixture()
    args = (z["u_rho"], z["rhod"], z["T"], z["cs"], z["sdust"], z["mdust"])
    psi = None
    for call in (1, 2):
        out = pyoracle.charge_marchand(p, *args, psi_old=psi)
        ref = ref_out(z, call)
        for k in OUT:
            a, b = out[k], ref[k]
            bad = np.flatnonzero(np.asarray(a).view(np.uint64).ravel() != np.asarray(b).view(np.uint64).ravel())
            assert bad.size == 0, (call, k, bad[:4], np.asarray(a).ravel()[bad[:4]], np.asarray(b).ravel()[bad[:4]])
        psi = out["psi"]
