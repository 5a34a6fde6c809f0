"""ct_amd -- MI355X (gfx950) implementation of Ct's FDK reconstruction hot path.

The compute lives in ``libctg.so`` (hand-written HIP, C ABI in ``include/ctg.h``);
this package holds the ctypes binding (:mod:`ct_amd._ctg`), synthetic scan inputs
(:mod:`ct_amd.phantom`) and a ``torch.autograd`` harness (:mod:`ct_amd.autograd`).
There is no CPU fallback: every operation raises if the HIP library is missing.
"""
from . import _ctg
from ._ctg import CtgError, Context, Geometry, geometry_ctvolume, geometry_grid, load

__all__ = ["CtgError", "Context", "Geometry", "geometry_ctvolume", "geometry_grid", "load", "_ctg"]
