"""MI355X-native solvent-excluded-surface builder (hand-written gfx950 HIP kernels behind a C ABI).

Public API mirrors the reference host classes bcs::Sesdf / bcs::Ssesdf; see ``surface.py``.
"""
from .contour_buildup import ContourBuildup  # noqa: F401
from .loader import element_radius, load  # noqa: F401
from .surface import Sesdf, Ssesdf, get_aabb  # noqa: F401

__all__ = ["Sesdf", "Ssesdf", "ContourBuildup", "get_aabb", "load", "element_radius"]
