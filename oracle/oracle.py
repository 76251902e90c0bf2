"""ctypes binding of the oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: the f64 CPU restatement of the zigtrace
reference path tracer. Only the test-suite, the smoke check and the
benchmark's CPU-baseline leg load it, and only as the checker or the
baseline -- never as the product path.

The flat record layouts below are declared independently of the product
library; both happen to share them so a test can feed one scene to both.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MATERIAL = np.dtype([("kind", "<u4"), ("_pad", "<u4"), ("color", "<f8", (3,)), ("param", "<f8")])
SHAPE = np.dtype([("kind", "<u4"), ("material", "<u4"), ("v", "<f8", (9,))])
CAMERA = np.dtype([("from", "<f8", (3,)), ("at", "<f8", (3,)), ("up", "<f8", (3,)), ("vfov", "<f8"),
                   ("aperture", "<f8"), ("focus", "<f8"), ("has_focus", "<u4"), ("_pad", "<u4")])
BACKGROUND = np.dtype([("kind", "<u4"), ("_pad", "<u4"), ("a", "<f8", (3,)), ("b", "<f8", (3,))])
NODE = np.dtype([("min", "<f8", (3,)), ("max", "<f8", (3,)), ("first", "<u4"), ("count", "<u4")])
RAY = np.dtype([("o", "<f8", (3,)), ("d", "<f8", (3,))])
HIT = np.dtype([("t", "<f8"), ("p", "<f8", (3,)), ("n", "<f8", (3,)), ("front", "<u4"), ("material", "<u4"),
                ("hit", "<i4"), ("_pad", "<i4")])
COUNTERS = np.dtype([(k, "<u8") for k in ("extend_rays", "shadow_rays", "node_pops", "leaf_pops", "tri_tests",
                                          "sph_tests", "plane_tests", "paths")])
OPTIONS = np.dtype([("width", "<u8"), ("height", "<u8"), ("spp", "<u8"), ("max_depth", "<u8"), ("threads", "<u8"),
                    ("seed", "<u8"), ("tile_size", "<u8"), ("light_sampling", "<u4"), ("_pad", "<u4")])
LIGHT_SAMPLE = np.dtype([("dir", "<f8", (3,)), ("distance", "<f8"), ("radiance", "<f8", (3,)), ("pdf", "<f8"),
                         ("valid", "<i4"), ("_pad", "<i4")])

assert MATERIAL.itemsize == 40 and SHAPE.itemsize == 80 and CAMERA.itemsize == 104
assert BACKGROUND.itemsize == 56 and NODE.itemsize == 56 and RAY.itemsize == 48 and HIT.itemsize == 72
assert OPTIONS.itemsize == 64 and LIGHT_SAMPLE.itemsize == 72

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_scene_new.restype = C.c_void_p
        _lib.orc_scene_free.argtypes = [C.c_void_p]
        _lib.orc_scene_info.argtypes = [C.c_void_p, C.c_void_p]
        _lib.orc_hardware_threads.restype = C.c_uint64
        _lib.orc_schlick.restype = C.c_double
        _lib.orc_schlick.argtypes = [C.c_double, C.c_double]
        _lib.orc_light_pdf_hit.restype = C.c_double
        _lib.orc_light_pdf_hit.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_aabb_op.restype = C.c_double
        _lib.orc_aabb_op.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
        _lib.orc_aabb_hit.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p]
        _lib.orc_shape_intersect.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p]
        _lib.orc_scene_intersect.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_double, C.c_double, C.c_int,
                                             C.c_void_p, C.c_void_p]
        _lib.orc_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_int, C.c_void_p,
                                   C.c_void_p]
        _lib.orc_render.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        _lib.orc_render_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p]
        _lib.orc_rng_state.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_void_p]
        _lib.orc_rng_draw.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, C.c_void_p]
        _lib.orc_splitmix64.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p]
        _lib.orc_camera_init.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
        _lib.orc_camera_rays.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                         C.c_void_p]
        _lib.orc_scene_parse.argtypes = [C.c_char_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        _lib.orc_bvh_build.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_scatter.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_emitted.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_light_sample.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_background.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_vec_refract.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
        _lib.orc_vec_reflect.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_shape_bounds.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_scene_flat.argtypes = [C.c_void_p] * 6
        _lib.orc_scene_bvh.argtypes = [C.c_void_p] * 3
        _lib.orc_scene_lights.argtypes = [C.c_void_p] * 6
        _lib.orc_set_perturbation.argtypes = [C.c_double]
    return _lib


def _p(a: np.ndarray) -> C.c_void_p:
    # data_as keeps a reference to the array, so temporaries outlive the call.
    return a.ctypes.data_as(C.c_void_p)


def _rec(dtype, **kw) -> np.ndarray:
    a = np.zeros(1, dtype=dtype)
    for k, v in kw.items():
        a[k] = v
    return a


# ----------------------------------------------------------------- RNG (rng.zig)
class Rng:
    """rng.zig:10-80 with its state held in a numpy u64[4]."""

    def __init__(self, seed: int = 0, stream: int | None = None, state=None):
        self.s = np.zeros(4, dtype=np.uint64)
        if state is not None:
            self.s[:] = state
        else:
            lib().orc_rng_state(seed, 0 if stream is None else stream, 0 if stream is None else 1, _p(self.s))

    @classmethod
    def for_stream(cls, seed: int, stream: int) -> "Rng":
        return cls(seed, stream)

    def _draw(self, kind, n, bound=0, dtype=np.uint64, width=1):
        out = np.zeros(n * width, dtype=dtype)
        lib().orc_rng_draw(_p(self.s), kind, bound, n, _p(out))
        return out.reshape(n, width) if width > 1 else out

    def next(self, n=1):
        return self._draw(0, n)

    def float(self, n=1):
        return self._draw(1, n, dtype=np.float64)

    def below(self, bound, n=1):
        return self._draw(2, n, bound=bound)

    def unit_vector(self, n=1):
        return self._draw(3, n, dtype=np.float64, width=3)

    def in_unit_sphere(self, n=1):
        return self._draw(4, n, dtype=np.float64, width=3)

    def in_unit_disk(self, n=1):
        return self._draw(5, n, dtype=np.float64, width=3)


def splitmix64(state: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_splitmix64(state, n, _p(out))
    return out


# ----------------------------------------------------------------- helpers
def rays(origins, dirs) -> np.ndarray:
    o = np.atleast_2d(np.asarray(origins, dtype=np.float64))
    d = np.atleast_2d(np.asarray(dirs, dtype=np.float64))
    n = max(len(o), len(d))
    r = np.zeros(n, dtype=RAY)
    r["o"] = o
    r["d"] = d
    return r


def shape_intersect(shape: np.ndarray, ray: np.ndarray, tmin: float, tmax: float) -> np.ndarray | None:
    h = np.zeros(1, dtype=HIT)
    ok = lib().orc_shape_intersect(_p(shape), _p(ray), tmin, tmax, _p(h))
    return h[0] if ok else None


def shape_bounds(shape: np.ndarray):
    mn = np.zeros(3)
    mx = np.zeros(3)
    return (mn, mx) if lib().orc_shape_bounds(_p(shape), _p(mn), _p(mx)) else None


def aabb_hit(mn, mx, ray: np.ndarray, tmin: float, tmax: float):
    t = np.zeros(1)
    ok = lib().orc_aabb_hit(_p(np.asarray(mn, float)), _p(np.asarray(mx, float)), _p(ray), tmin, tmax, _p(t))
    return float(t[0]) if ok else None


def aabb_surface_area(mn, mx) -> float:
    return lib().orc_aabb_op(0, _p(np.asarray(mn, float)), _p(np.asarray(mx, float)), 0.0, None)


def aabb_pad(mn, mx, delta):
    out = np.zeros(6)
    lib().orc_aabb_op(1, _p(np.asarray(mn, float)), _p(np.asarray(mx, float)), delta, _p(out))
    return out[:3], out[3:]


def aabb_longest_axis(mn, mx) -> int:
    return int(lib().orc_aabb_op(2, _p(np.asarray(mn, float)), _p(np.asarray(mx, float)), 0.0, None))


def refract(uv, n, eta):
    out = np.zeros(3)
    lib().orc_vec_refract(_p(np.asarray(uv, float)), _p(np.asarray(n, float)), eta, _p(out))
    return out


def reflect(v, n):
    out = np.zeros(3)
    lib().orc_vec_reflect(_p(np.asarray(v, float)), _p(np.asarray(n, float)), _p(out))
    return out


def schlick(cosine: float, eta: float) -> float:
    return lib().orc_schlick(cosine, eta)


def scatter(material: np.ndarray, ray_in: np.ndarray, hit: np.ndarray, rng: Rng):
    out = np.zeros(1, dtype=RAY)
    att = np.zeros(3)
    h = np.zeros(1, dtype=HIT)
    h[0] = hit
    ok = lib().orc_scatter(_p(material), _p(ray_in), _p(h), _p(rng.s), _p(out), _p(att))
    return (out[0], att) if ok else None


def emitted(material: np.ndarray, hit: np.ndarray):
    out = np.zeros(3)
    h = np.zeros(1, dtype=HIT)
    h[0] = hit
    lib().orc_emitted(_p(material), _p(h), _p(out))
    return out


def camera_init(cam: np.ndarray, aspect: float) -> dict:
    out = np.zeros(19)
    lib().orc_camera_init(_p(cam), aspect, _p(out))
    names = ("origin", "upper_left", "horizontal", "vertical", "u", "v")
    d = {k: out[3 * i:3 * i + 3].copy() for i, k in enumerate(names)}
    d["lens_radius"] = float(out[18])
    return d


def camera_rays(cam: np.ndarray, aspect: float, s, t, rng: Rng) -> np.ndarray:
    s = np.ascontiguousarray(s, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    out = np.zeros(len(s), dtype=RAY)
    lib().orc_camera_rays(_p(cam), aspect, _p(s), _p(t), len(s), _p(rng.s), _p(out))
    return out


def bvh_build(shapes: np.ndarray):
    """Bvh.build over `shapes` (bvh.zig:41-67). Returns (nodes, indices) or an error name."""
    n = len(shapes)
    nodes = np.zeros(max(2 * n - 1, 1), dtype=NODE)
    idx = np.zeros(max(n, 1), dtype=np.uint32)
    nn = np.zeros(1, dtype=np.uint64)
    rc = lib().orc_bvh_build(_p(shapes) if n else None, n, _p(nodes), _p(nn), _p(idx))
    if rc == 1:
        return "UnboundedShape"
    if rc == 2:
        return "TooManyPrimitives"
    return nodes[: int(nn[0]) if n else 0], idx[:n]


# ----------------------------------------------------------------- scenes
@dataclass
class Info:
    materials: int
    shapes: int
    planes: int
    nodes: int
    lights: int
    bvh_depth: int
    width: int
    height: int
    spp: int
    max_depth: int
    built: bool


class SyntaxErr(Exception):
    def __init__(self, line: int, message: str):
        super().__init__(f"{line}: {message}")
        self.line = line
        self.message = message


class OracleScene:
    """scene.zig's Scene after `build`, held by the oracle."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)

    @classmethod
    def from_arrays(cls, materials: np.ndarray, shapes: np.ndarray, camera: np.ndarray | None = None,
                    background: np.ndarray | None = None, build: bool = True) -> "OracleScene":
        materials = np.ascontiguousarray(materials, dtype=MATERIAL)
        shapes = np.ascontiguousarray(shapes, dtype=SHAPE)
        err = C.c_int(0)
        h = lib().orc_scene_new(_p(materials) if len(materials) else None, len(materials),
                                _p(shapes) if len(shapes) else None, len(shapes),
                                _p(camera) if camera is not None else None,
                                _p(background) if background is not None else None, int(build), C.byref(err))
        s = cls(h)
        if err.value:
            raise RuntimeError({1: "UnboundedShape", 2: "TooManyPrimitives"}.get(err.value, str(err.value)))
        return s

    @classmethod
    def from_view_arrays(cls, a: dict, build: bool = True) -> "OracleScene":
        """From the flattened scene inputs (materials, bounded shapes, planes,
        camera, background) of the product's Scene.view_arrays(): the oracle
        runs its own Bvh.build and Lights.build on them."""
        shapes = np.concatenate([a["shapes"], a["planes"]]).view(SHAPE)
        c = a["camera"][0]
        cam = _rec(CAMERA, **{"from": c["look_from"], "at": c["look_at"], "up": c["up"], "vfov": c["vfov"],
                              "aperture": c["aperture"], "focus": c["focus_dist"], "has_focus": c["has_focus_dist"]})
        bg = np.ascontiguousarray(a["background"]).view(BACKGROUND).copy()
        return cls.from_arrays(np.ascontiguousarray(a["materials"]).view(MATERIAL), shapes, cam, bg, build=build)

    @classmethod
    def parse(cls, text: str | bytes) -> "OracleScene":
        b = text.encode() if isinstance(text, str) else text
        out = C.c_void_p()
        line = np.zeros(1, dtype=np.uint64)
        msg = C.create_string_buffer(256)
        rc = lib().orc_scene_parse(b, len(b), C.byref(out), _p(line), msg, 256)
        if rc == 1:
            raise SyntaxErr(int(line[0]), msg.value.decode())
        if rc != 0:
            raise RuntimeError({2: "UnboundedShape", 3: "TooManyPrimitives"}.get(rc, str(rc)))
        return cls(out.value)

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().orc_scene_free(self.h)
            self.h = None

    def info(self) -> Info:
        a = np.zeros(11, dtype=np.uint64)
        lib().orc_scene_info(self.h, _p(a))
        return Info(*[int(x) for x in a[:10]], bool(a[10]))

    def flat(self):
        i = self.info()
        mats = np.zeros(i.materials, dtype=MATERIAL)
        shapes = np.zeros(i.shapes, dtype=SHAPE)
        planes = np.zeros(i.planes, dtype=SHAPE)
        cam = np.zeros(1, dtype=CAMERA)
        bg = np.zeros(1, dtype=BACKGROUND)
        lib().orc_scene_flat(self.h, _p(mats), _p(shapes), _p(planes), _p(cam), _p(bg))
        return mats, shapes, planes, cam, bg

    def bvh(self):
        i = self.info()
        nodes = np.zeros(i.nodes, dtype=NODE)
        idx = np.zeros(i.shapes, dtype=np.uint32)
        lib().orc_scene_bvh(self.h, _p(nodes), _p(idx))
        return nodes, idx

    def lights(self):
        n = self.info().lights
        si = np.zeros(n, dtype=np.uint32)
        rad = np.zeros((n, 3))
        w = np.zeros(n)
        cdf = np.zeros(n)
        tot = np.zeros(1)
        lib().orc_scene_lights(self.h, _p(si), _p(rad), _p(w), _p(cdf), _p(tot))
        return {"shape_index": si, "radiance": rad, "weight": w, "cdf": cdf, "total": float(tot[0])}

    def intersect(self, rays_: np.ndarray, tmin: float, tmax: float, mode: int = 0):
        out = np.zeros(len(rays_), dtype=HIT)
        ctr = np.zeros(1, dtype=COUNTERS)
        lib().orc_scene_intersect(self.h, _p(rays_), len(rays_), tmin, tmax, mode, _p(out), _p(ctr))
        return out, ctr[0]

    def background(self, d):
        out = np.zeros(3)
        lib().orc_background(self.h, _p(np.asarray(d, float)), _p(out))
        return out

    def light_sample(self, origin, rng: Rng):
        out = np.zeros(1, dtype=LIGHT_SAMPLE)
        ok = lib().orc_light_sample(self.h, _p(np.asarray(origin, float)), _p(rng.s), _p(out))
        return out[0] if ok else None

    def pdf_hit(self, origin, hit) -> float:
        h = np.zeros(1, dtype=HIT)
        h[0] = hit
        return lib().orc_light_pdf_hit(self.h, _p(np.asarray(origin, float)), _p(h))

    def trace(self, rays_: np.ndarray, states: np.ndarray, max_depth: int, light_sampling: bool = True):
        """traceWith per ray; `states` (n,4) u64 is advanced in place."""
        assert states.dtype == np.uint64 and states.shape == (len(rays_), 4) and states.flags.c_contiguous
        out = np.zeros((len(rays_), 3))
        ctr = np.zeros(1, dtype=COUNTERS)
        lib().orc_trace(self.h, _p(rays_), len(rays_), _p(states), max_depth, int(light_sampling), _p(out), _p(ctr))
        return out, ctr[0]

    def render(self, width, height, spp, max_depth, threads=None, seed=0, tile_size=16, light_sampling=True,
               pixel_stride=1):
        """render.zig:152-180; returns (H, W, 3) f64 and the ray/node counters."""
        if threads is None:
            threads = int(lib().orc_hardware_threads())
        o = _rec(OPTIONS, width=width, height=height, spp=spp, max_depth=max_depth, threads=threads, seed=seed,
                 tile_size=tile_size, light_sampling=int(light_sampling))
        out = np.zeros((max(height, 1), max(width, 1), 3))
        ctr = np.zeros(1, dtype=COUNTERS)
        rc = lib().orc_render(self.h, _p(o), _p(out), pixel_stride, _p(ctr))
        if rc:
            raise ValueError("InvalidOptions")
        return out, ctr[0]


def render_grad(scene: "OracleScene", width, height, spp, max_depth, offsets, dimg, seed=0, threads=None,
                light_sampling=True):
    """FP64 reference gradient (forward-mode dual numbers): (image, dLoss/dparams)."""
    if threads is None:
        threads = hardware_threads()
    o = _rec(OPTIONS, width=width, height=height, spp=spp, max_depth=max_depth, threads=threads, seed=seed,
             tile_size=16, light_sampling=int(light_sampling))
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    nparams = int(offsets.max()) + 3 if (offsets >= 0).any() else 0
    dimg = np.ascontiguousarray(dimg, dtype=np.float64).reshape(-1)
    assert dimg.size == width * height * 3
    img = np.zeros((height, width, 3))
    grad = np.zeros(max(nparams, 1))
    rc = lib().orc_render_grad(scene.h, _p(o), _p(offsets), nparams, _p(dimg), _p(img), _p(grad))
    if rc:
        raise ValueError("InvalidOptions" if rc == 1 else "too many parameters")
    return img, grad[:nparams]


class rounding_perturbation:
    """Context manager: renders inside it perturb every hit point and scattered
    direction by a relative (-rel, rel) (deterministic per path vertex, no RNG
    draw consumed) -- the image noise that rounding at that scale alone causes
    with common random numbers."""

    def __init__(self, rel: float):
        self.rel = rel

    def __enter__(self):
        lib().orc_set_perturbation(self.rel)
        return self

    def __exit__(self, *exc):
        lib().orc_set_perturbation(0.0)
        return False


def hardware_threads() -> int:
    return int(lib().orc_hardware_threads())
