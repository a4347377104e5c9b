"""RenderCore_HIP: a MI355X (gfx950) render core behind Lighthouse 2's CoreAPI.

The compute path is librendercore_hip.so (hand-written HIP kernels, C ABI in
include/lh2hip.h); this package holds its Python binding (`core`), the boundary
struct layouts (`formats`) and the host-side scene producers (`scenes`).
"""
from . import formats, scenes  # noqa: F401
from .core import RenderCore, RenderServer, Svgf, LH2HIPError, load_scene, library, LH2_CONVERGE, LH2_RESTART  # noqa: F401
