"""vivsim_amd — MI355X-native (gfx950 HIP) implementation of VIVSIM's IB-LBM time step.

Drop-in modules with the reference's function names and array layouts:
``vivsim_amd.lbm`` (D2Q9), ``vivsim_amd.lbm3d`` (D3Q19), ``vivsim_amd.ib``,
``vivsim_amd.ib3d``, ``vivsim_amd.dyn``, ``vivsim_amd.post`` (field diagnostics)
and ``vivsim_amd.multigrid`` (refinement transfers); fused multi-step drivers
in ``vivsim_amd.sim``.  All arithmetic runs in the HIP library
``libvivsim_amd.so`` (C ABI: include/vivsim_amd.h).
"""

from . import dyn, ib, ib3d, lbm, lbm3d, multigrid, post, sim  # noqa: F401
from ._lib import lib  # noqa: F401

__version__ = "0.1.0"
