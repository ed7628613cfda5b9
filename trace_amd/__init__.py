"""trace_amd -- the rendering hot path of ChSchmidt99/trace (LBVH build, PHR
refinement, closest-hit traversal, path tracing) as hand-written gfx950 HIP
kernels behind a C-ABI (include/trace_hip.h)."""
from . import _lib  # noqa: F401
from ._lib import KEY_M30, KEY_REF63, TraceError, load  # noqa: F401

__all__ = ["load", "TraceError", "KEY_REF63", "KEY_M30"]
