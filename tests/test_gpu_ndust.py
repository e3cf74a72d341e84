"""Numbers of dust species without an exact kernel instantiation (the post pass and the Courant
kernel take the species count at run time inside the next larger instantiation): the GPU step
equals the oracle bit for bit on the shearing box and on the FARGO disk, for both post segment
widths."""
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
        return np.array([g.step() for _ in range(nsteps)]), g.download()
    finally:
        g.close()


@pytest.mark.parametrize("nd", [5, 7, 12, 20, 33, 50])
@pytest.mark.parametrize("which", ["si", "disk"])
def test_any_species_count_matches_oracle(nd, which):
    if which == "si":
        case, n = setups.si_box(n=32, nd=nd), 8
    else:
        case, n = setups.disk(nr=40, nphi=64, nd=nd), 8
    print(f"{which} nd={nd}", flush=True)
    cfg, geom, state = case
    o = pyoracle.Oracle(cfg, geom, state)
    dto = np.array([o.step() for _ in range(n)])
    uo = o.u()
    m = active_mask(cfg)
    for tw in ((16, 32) if nd <= 16 else (8, 16)):
        dts, u = _run(case, n, post_tw=tw)
        assert np.array_equal(dts, dto), tw
        assert np.array_equal(u[:, m], uo[:, m]), tw
