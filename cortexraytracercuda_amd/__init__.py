"""MI355X (gfx950) render hot path of CortexRaytracerCuda: HIP wavefront path tracer behind a C-ABI.

The compute path is libcortex_rt.so (csrc/): host BVH build + hand-written HIP kernels. This
package holds the ctypes front end, the scene description and the synthetic workloads.
"""
from . import _abi
from .scene import Scene
from .renderer import Renderer, SceneFile, launch_gpu, read_scene_file, render_desc, write_png, pixel_lines
from . import scenes

__all__ = ["Scene", "SceneFile", "read_scene_file", "Renderer", "launch_gpu", "render_desc", "write_png", "pixel_lines", "scenes", "_abi"]
