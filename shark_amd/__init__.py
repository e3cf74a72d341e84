"""shark_amd — MI355X-native implementation of SHARK's dusty-gas Godunov step.

The step (gas + N dust fluids: MUSCL-Hancock predictor, HLL/HLLC/LLF gas and Huang-Bai
dust Riemann fluxes, conservative update, polar sources, FARGO, implicit drag) runs as
hand-written gfx950 HIP kernels in libshark_gpu.so behind the C-ABI of
include/shark_gpu.h.  This package holds the Python view of that boundary (``gpu``), the
configuration struct (``cfg``) and the setups that build initial states (``setups``).
"""
from . import cfg  # noqa: F401

__all__ = ["cfg", "gpu", "setups"]
