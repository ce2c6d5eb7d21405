"""xenon_amd — MI355X (gfx950) wavefront path-tracing backend for Xenon.

The product is libxenon_hip.so (C ABI in include/xenon_hip.h); this package
is its Python host mirror: ``HipRenderer`` (the reference's CudaRenderer
contract), ``HipScene`` (the reference's Scene after build_acceleration)
and the seeded workload generators in ``scenes``.
"""
from . import scenes  # noqa: F401
from .native import XnhError  # noqa: F401
from .renderer import HIT_DTYPE, NODE_DTYPE, GpuScene, HipRenderer, HipScene  # noqa: F401
from .swapchain import TripleSwapchain, progressive  # noqa: F401

__all__ = ["GpuScene", "HipRenderer", "HipScene", "XnhError", "scenes", "HIT_DTYPE", "NODE_DTYPE",
           "TripleSwapchain", "progressive"]
