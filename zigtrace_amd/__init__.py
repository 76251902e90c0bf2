"""zigtrace_amd: zigtrace's path-tracing core on MI355X (gfx950).

Mirrors the reference library surface for the hot path (src/root.zig):
``Scene`` (scene.zig), ``bvh_build`` (bvh.zig), ``scenefile.parse``,
``render``/``Options`` (render.zig) -- with rendering done by hand-written
HIP kernels in ``lib/libzigtrace_amd.so``.
"""
from . import _native, image
from ._native import ZTError
from .render import Context, Options, Tuning, render
from .scene import CameraConfig, Scene, bvh_build, dielectric, emissive, lambertian, metal, plane, sky, solid, \
    sphere, triangle
from .scenefile import SceneSyntaxError, parse

__all__ = ["image", "Scene", "CameraConfig", "Options", "Tuning", "Context", "render", "parse", "bvh_build", "ZTError",
           "SceneSyntaxError", "lambertian", "metal", "dielectric", "emissive", "sphere", "plane", "triangle", "solid",
           "sky"]
