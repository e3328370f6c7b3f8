"""axi_visualizer_amd -- MI355X (gfx950) rendering hot path of AXI Visualizer.

``CLDraw`` keeps the reference's draw-call and shader-binding API
(PyOpenCL27C/Ops.py) and runs hand-written HIP kernels through the C ABI of
libaxr.so (include/axr.h).
"""
from .ops import CLDraw, shader_flags  # noqa: F401
from .scenes import SH_CULL, SH_MIP, SH_PHONG, SH_SHADOW  # noqa: F401

__all__ = ["CLDraw", "shader_flags", "SH_MIP", "SH_SHADOW", "SH_PHONG", "SH_CULL"]
