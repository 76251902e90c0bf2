"""ctypes binding of ``libzigtrace_amd.so`` (include/zt_types.h, zt_gpu.h, zt_host.h).

The library is built in-tree (``__graft_entry__.build()`` or
``make -C zigtrace_amd/csrc``). There is no fallback: importing this module
without the library raises, so nothing can silently run another path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB_PATH = os.path.join(HERE, "lib", "libzigtrace_amd.so")
# ZIGTRACE_AMD_LIB selects another build of the same library (an instrumented
# or differently configured one, exporting the same ABI). It is honoured only
# together with ZIGTRACE_AMD_ALLOW_OTHER_LIB=1, so a stray variable cannot
# silently swap the library that tests and benchmarks load.
_override = os.environ.get("ZIGTRACE_AMD_LIB")
if _override and os.environ.get("ZIGTRACE_AMD_ALLOW_OTHER_LIB") != "1":
    raise ImportError(f"ZIGTRACE_AMD_LIB={_override} is set but ZIGTRACE_AMD_ALLOW_OTHER_LIB=1 is not: "
                      "refusing to load a library other than the in-tree build")
LIB_PATH = _override or DEFAULT_LIB_PATH

# ---------------------------------------------------------------- status codes (zt_types.h)
OK = 0
ERR_INVALID_OPTIONS = -1
ERR_OUT_OF_MEMORY = -2
ERR_HIP = -3
ERR_BAD_SCENE = -4
ERR_INVALID_ARGUMENT = -5
ERR_UNBOUNDED_SHAPE = -6
ERR_TOO_MANY_PRIMITIVES = -7
ERR_SYNTAX = -8
ERR_INTERNAL = -9
ERR_INVALID_DIMENSIONS = -10
ERR_IMAGE_TOO_LARGE = -11
ERR_BUFFER_TOO_SMALL = -12
ERROR_NAMES = {
    ERR_INVALID_OPTIONS: "InvalidOptions", ERR_OUT_OF_MEMORY: "OutOfMemory", ERR_HIP: "HipError",
    ERR_BAD_SCENE: "BadScene", ERR_INVALID_ARGUMENT: "InvalidArgument", ERR_UNBOUNDED_SHAPE: "UnboundedShape",
    ERR_TOO_MANY_PRIMITIVES: "TooManyPrimitives", ERR_SYNTAX: "SyntaxError", ERR_INTERNAL: "Internal",
    ERR_INVALID_DIMENSIONS: "InvalidDimensions", ERR_IMAGE_TOO_LARGE: "ImageTooLarge",
    ERR_BUFFER_TOO_SMALL: "BufferTooSmall",
}

LAMBERTIAN, METAL, DIELECTRIC, EMISSIVE = 0, 1, 2, 3
SPHERE, TRIANGLE, PLANE = 0, 1, 2
BG_SOLID, BG_SKY = 0, 1
METHOD_AUTO, METHOD_MEGAKERNEL, METHOD_WAVEFRONT = 0, 1, 2
PRECISION_FP32, PRECISION_FP64 = 0, 1

# ---------------------------------------------------------------- records
MATERIAL = np.dtype([("kind", "<u4"), ("_pad", "<u4"), ("color", "<f8", (3,)), ("param", "<f8")])
SHAPE = np.dtype([("kind", "<u4"), ("material", "<u4"), ("v", "<f8", (9,))])
CAMERA = np.dtype([("look_from", "<f8", (3,)), ("look_at", "<f8", (3,)), ("up", "<f8", (3,)), ("vfov", "<f8"),
                   ("aperture", "<f8"), ("focus_dist", "<f8"), ("has_focus_dist", "<u4"), ("_pad", "<u4")])
BACKGROUND = np.dtype([("kind", "<u4"), ("_pad", "<u4"), ("a", "<f8", (3,)), ("b", "<f8", (3,))])
NODE = np.dtype([("min", "<f8", (3,)), ("max", "<f8", (3,)), ("first", "<u4"), ("count", "<u4")])
LIGHT = np.dtype([("shape", "<u4"), ("_pad", "<u4"), ("radiance", "<f8", (3,)), ("weight", "<f8")])
OPTIONS = np.dtype([("width", "<u8"), ("height", "<u8"), ("spp", "<u8"), ("max_depth", "<u8"), ("threads", "<u8"),
                    ("seed", "<u8"), ("tile_size", "<u8"), ("light_sampling", "<u4"), ("_pad", "<u4")])
RAY = np.dtype([("origin", "<f8", (3,)), ("dir", "<f8", (3,))])
HIT = np.dtype([("t", "<f8"), ("point", "<f8", (3,)), ("normal", "<f8", (3,)), ("front_face", "<u4"),
                ("material", "<u4"), ("hit", "<i4"), ("prim", "<i4")])
TUNING = np.dtype([("method", "<u4"), ("precision", "<u4"), ("tile_w", "<u4"), ("tile_h", "<u4"),
                   ("block_size", "<u4"), ("queue_paths", "<u4"), ("variant", "<u4"), ("stack_entries", "<u4"),
                   ("bvh_width", "<u4"), ("schedule", "<u4"), ("refill", "<u4"), ("node_bytes", "<u4"),
                   ("tile_order", "<u4")])
OUTPUT = np.dtype([("tonemap", "<u4"), ("_pad", "<u4"), ("gamma", "<f8"), ("exposure", "<f8")])
TONEMAP_NONE, TONEMAP_REINHARD, TONEMAP_ACES = 0, 1, 2
STATS = np.dtype([("kernel_ms", "<f8"), ("upload_ms", "<f8"), ("launches", "<u8"), ("paths", "<u8"),
                  ("extend_rays", "<u8"), ("shadow_rays", "<u8")])
SCENE_INFO = np.dtype([("num_materials", "<u8"), ("num_shapes", "<u8"), ("num_planes", "<u8"), ("num_nodes", "<u8"),
                       ("num_lights", "<u8"), ("bvh_depth", "<u8"), ("width", "<u8"), ("height", "<u8"),
                       ("spp", "<u8"), ("max_depth", "<u8"), ("built", "<u4"), ("_pad", "<u4")])

for _dt, _size in ((MATERIAL, 40), (SHAPE, 80), (CAMERA, 104), (BACKGROUND, 56), (NODE, 56), (LIGHT, 40),
                   (OPTIONS, 64), (RAY, 48), (HIT, 72), (TUNING, 52), (STATS, 48), (OUTPUT, 24), (SCENE_INFO, 88)):
    assert _dt.itemsize == _size, (_dt, _size)


class SceneView(C.Structure):
    """zt_scene_view (zt_gpu.h)."""
    _fields_ = [
        ("materials", C.c_void_p), ("num_materials", C.c_uint64),
        ("shapes", C.c_void_p), ("num_shapes", C.c_uint64),
        ("planes", C.c_void_p), ("num_planes", C.c_uint64),
        ("nodes", C.c_void_p), ("num_nodes", C.c_uint64),
        ("indices", C.c_void_p),
        ("lights", C.c_void_p), ("num_lights", C.c_uint64),
        ("light_cdf", C.c_void_p), ("light_total", C.c_double),
        ("camera", C.c_byte * 104), ("background", C.c_byte * 56),
    ]


assert C.sizeof(SceneView) == 104 + 56 + 8 * 13


class ZTError(RuntimeError):
    """A non-zero status from the library; `name` is the reference's error name."""

    def __init__(self, code: int, message: str = ""):
        self.code = code
        self.name = ERROR_NAMES.get(code, str(code))
        super().__init__(f"{self.name}" + (f": {message}" if message else ""))


_lib = None

_SIGS = {
    # zt_gpu.h
    "zt_gpu_create": (C.c_int, [C.c_int, C.c_void_p]),
    "zt_gpu_destroy": (None, [C.c_void_p]),
    "zt_gpu_last_error": (C.c_char_p, [C.c_void_p]),
    "zt_gpu_stream": (C.c_void_p, [C.c_void_p]),
    "zt_gpu_get_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "zt_gpu_device_alloc": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "zt_gpu_device_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "zt_gpu_copy_to_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "zt_gpu_copy_to_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "zt_gpu_scene_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "zt_gpu_scene_release": (C.c_int, [C.c_void_p, C.c_uint32]),
    "zt_gpu_render": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "zt_gpu_render_batch": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int]),
    "zt_gpu_encode_rgb8": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_int,
                                     C.c_int]),
    "zt_gpu_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                  C.c_void_p]),
    "zt_gpu_trace_rays": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                    C.c_uint32, C.c_uint32, C.c_void_p]),
    "zt_gpu_intersect": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.c_double, C.c_double, C.c_int,
                                   C.c_uint32, C.c_void_p]),
    "zt_gpu_rng": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_void_p]),
    "zt_gpu_param_count": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "zt_gpu_param_offsets": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "zt_gpu_get_params": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "zt_gpu_set_params": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "zt_gpu_forward": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "zt_gpu_backward": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                  C.c_void_p]),
    "zt_gpu_backward_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    # zt_host.h
    "zt_scene_create": (C.c_void_p, []),
    "zt_scene_destroy": (None, [C.c_void_p]),
    "zt_scene_add_material": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "zt_scene_add": (C.c_int, [C.c_void_p, C.c_void_p]),
    "zt_scene_add_many": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "zt_scene_add_quad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "zt_scene_set_camera": (C.c_int, [C.c_void_p, C.c_void_p]),
    "zt_scene_set_background": (C.c_int, [C.c_void_p, C.c_void_p]),
    "zt_scene_set_settings": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]),
    "zt_scene_build": (C.c_int, [C.c_void_p]),
    "zt_scene_get_info": (C.c_int, [C.c_void_p, C.c_void_p]),
    "zt_scene_get_view": (C.c_int, [C.c_void_p, C.c_void_p]),
    "zt_scene_parse": (C.c_int, [C.c_char_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_char_p, C.c_uint64]),
    "zt_bvh_build": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "zt_encode_rgb8": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]),
    "zt_write_ppm": (C.c_int, [C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "zt_write_png": (C.c_int, [C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "zt_bvh_build_threads": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "zt_render": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "zt_default_context": (C.c_int, [C.c_void_p]),
    "zt_scene_device_id": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
}


def lib() -> C.CDLL:
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"zigtrace_amd native library not built: {LIB_PATH} "
                              "(run __graft_entry__.build() or make -C zigtrace_amd/csrc)")
        _lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def loaded_library() -> dict:
    """Path and SHA-256 (first 16 hex digits) of the library this process loaded."""
    import hashlib
    lib()
    with open(LIB_PATH, "rb") as f:
        digest = hashlib.sha256(f.read()).hexdigest()[:16]
    return {"path": os.path.relpath(LIB_PATH, os.path.dirname(HERE)), "sha256_16": digest,
            "default": os.path.abspath(LIB_PATH) == os.path.abspath(DEFAULT_LIB_PATH)}


def declared_symbols() -> list[str]:
    """Every function the public headers declare (parsed from include/*.h)."""
    import re
    inc = os.path.join(os.path.dirname(HERE), "include")
    names = []
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(zt_\w+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def ptr(a: np.ndarray | None) -> C.c_void_p | None:
    # data_as keeps a reference to the array, so temporaries outlive the call.
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def check(code: int, ctx=None) -> None:
    if code != OK:
        msg = lib().zt_gpu_last_error(ctx).decode() if ctx else ""
        raise ZTError(code, msg)


def record(dtype, **fields) -> np.ndarray:
    a = np.zeros(1, dtype=dtype)
    for k, v in fields.items():
        a[k] = v
    return a
