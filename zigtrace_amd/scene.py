"""Scene construction (scene.zig), backed by the C++ host library (zt_host.h).

`Scene` mirrors the reference's `Scene` API: materials are added and
referenced by id, bounded shapes go into the BVH and planes are tested
linearly, `build()` runs the bit-identical binned-SAH `Bvh.build` and
`Lights.build` on the host.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N


@dataclass
class CameraConfig:
    """camera.zig:8-18."""
    look_from: tuple = (0.0, 0.0, 0.0)
    look_at: tuple = (0.0, 0.0, -1.0)
    up: tuple = (0.0, 1.0, 0.0)
    vfov: float = 60.0
    aperture: float = 0.0
    focus_dist: float | None = None

    def record(self) -> np.ndarray:
        return N.record(N.CAMERA, look_from=self.look_from, look_at=self.look_at, up=self.up, vfov=self.vfov,
                        aperture=self.aperture, focus_dist=self.focus_dist or 0.0,
                        has_focus_dist=0 if self.focus_dist is None else 1)

    @classmethod
    def from_record(cls, r) -> "CameraConfig":
        return cls(tuple(r["look_from"]), tuple(r["look_at"]), tuple(r["up"]), float(r["vfov"]), float(r["aperture"]),
                   float(r["focus_dist"]) if r["has_focus_dist"] else None)


def solid(color) -> np.ndarray:
    """Background.solid (scene.zig:16)."""
    return N.record(N.BACKGROUND, kind=N.BG_SOLID, a=color)


def sky(horizon, zenith) -> np.ndarray:
    """Background.sky (scene.zig:17)."""
    return N.record(N.BACKGROUND, kind=N.BG_SKY, a=horizon, b=zenith)


def material(kind: int, color=(0.0, 0.0, 0.0), param: float = 0.0) -> np.ndarray:
    return N.record(N.MATERIAL, kind=kind, color=color, param=param)


def lambertian(albedo) -> np.ndarray:
    return material(N.LAMBERTIAN, albedo)


def metal(albedo, fuzz: float = 0.0) -> np.ndarray:
    return material(N.METAL, albedo, fuzz)


def dielectric(ior: float) -> np.ndarray:
    return material(N.DIELECTRIC, (0, 0, 0), ior)


def emissive(color) -> np.ndarray:
    return material(N.EMISSIVE, color)


def sphere(center, radius: float, mat: int) -> np.ndarray:
    return N.record(N.SHAPE, kind=N.SPHERE, material=mat, v=list(center) + [radius] + [0.0] * 5)


def plane(point, normal, mat: int) -> np.ndarray:
    return N.record(N.SHAPE, kind=N.PLANE, material=mat, v=list(point) + list(normal) + [0.0] * 3)


def triangle(v0, v1, v2, mat: int) -> np.ndarray:
    return N.record(N.SHAPE, kind=N.TRIANGLE, material=mat, v=list(v0) + list(v1) + list(v2))


class Scene:
    """scene.zig:39-52: materials, bounded shapes, planes, camera, background, settings, BVH, lights."""

    def __init__(self, handle=None):
        self._h = C.c_void_p(handle if handle is not None else N.lib().zt_scene_create())
        if not self._h.value:
            raise N.ZTError(N.ERR_OUT_OF_MEMORY)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            N.lib().zt_scene_destroy(h)
            self._h = None

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    # ---- scene.zig:66-88
    def add_material(self, m: np.ndarray) -> int:
        out = C.c_uint32()
        N.check(N.lib().zt_scene_add_material(self._h, N.ptr(m), C.byref(out)))
        return out.value

    def add(self, shape: np.ndarray) -> None:
        N.check(N.lib().zt_scene_add(self._h, N.ptr(np.ascontiguousarray(shape, dtype=N.SHAPE))))

    def add_many(self, shapes: np.ndarray) -> None:
        shapes = np.ascontiguousarray(shapes, dtype=N.SHAPE)
        N.check(N.lib().zt_scene_add_many(self._h, N.ptr(shapes) if len(shapes) else None, len(shapes)))

    def add_quad(self, corner, u, v, mat: int) -> None:
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (corner, u, v)]
        N.check(N.lib().zt_scene_add_quad(self._h, *[N.ptr(a) for a in arr], mat))

    def add_sphere(self, center, radius, mat):
        self.add(sphere(center, radius, mat))

    def add_plane(self, point, normal, mat):
        self.add(plane(point, normal, mat))

    def add_triangle(self, v0, v1, v2, mat):
        self.add(triangle(v0, v1, v2, mat))

    # ---- fields
    @property
    def camera(self) -> CameraConfig:
        return CameraConfig.from_record(self.view_arrays()["camera"][0])

    @camera.setter
    def camera(self, cfg: CameraConfig) -> None:
        N.check(N.lib().zt_scene_set_camera(self._h, N.ptr(cfg.record())))

    @property
    def background(self) -> np.ndarray:
        return self.view_arrays()["background"]

    @background.setter
    def background(self, bg: np.ndarray) -> None:
        N.check(N.lib().zt_scene_set_background(self._h, N.ptr(np.ascontiguousarray(bg, dtype=N.BACKGROUND))))

    def set_settings(self, width, height, spp, max_depth) -> None:
        N.check(N.lib().zt_scene_set_settings(self._h, width, height, spp, max_depth))

    # ---- scene.zig:92-96
    def build(self) -> None:
        N.check(N.lib().zt_scene_build(self._h))

    def info(self) -> dict:
        a = np.zeros(1, dtype=N.SCENE_INFO)
        N.check(N.lib().zt_scene_get_info(self._h, N.ptr(a)))
        return {k: int(a[k][0]) for k in N.SCENE_INFO.names if not k.startswith("_")}

    def view(self) -> N.SceneView:
        v = N.SceneView()
        N.check(N.lib().zt_scene_get_view(self._h, C.byref(v)))
        return v

    def view_arrays(self) -> dict:
        """Copies of the flattened scene (materials, shapes, planes, nodes, indices, lights, cdf, camera, background)."""
        v = self.view()

        def arr(p, n, dt):
            if n == 0 or not p:
                return np.zeros(0, dtype=dt)
            buf = (C.c_byte * (n * dt.itemsize)).from_address(p)
            return np.frombuffer(buf, dtype=dt, count=n).copy()

        return {
            "materials": arr(v.materials, v.num_materials, N.MATERIAL),
            "shapes": arr(v.shapes, v.num_shapes, N.SHAPE),
            "planes": arr(v.planes, v.num_planes, N.SHAPE),
            "nodes": arr(v.nodes, v.num_nodes, N.NODE),
            "indices": arr(v.indices, v.num_shapes if v.num_nodes else 0, np.dtype("<u4")),
            "lights": arr(v.lights, v.num_lights, N.LIGHT),
            "cdf": arr(v.light_cdf, v.num_lights, np.dtype("<f8")),
            "light_total": float(v.light_total),
            "camera": np.frombuffer(bytes(v.camera), dtype=N.CAMERA).copy(),
            "background": np.frombuffer(bytes(v.background), dtype=N.BACKGROUND).copy(),
        }


def bvh_build(shapes: np.ndarray, threads: int | None = None):
    """Bvh.build (bvh.zig:41-67) over a shape list: (nodes, indices). `threads`: 1 = the
    sequential build, 0 or None = the parallel build on up to 16 threads (same bits)."""
    shapes = np.ascontiguousarray(shapes, dtype=N.SHAPE)
    n = len(shapes)
    nodes = np.zeros(max(2 * n - 1, 1), dtype=N.NODE)
    idx = np.zeros(max(n, 1), dtype=np.uint32)
    nn = C.c_uint64()
    N.check(N.lib().zt_bvh_build_threads(N.ptr(shapes) if n else None, n, threads or 0, N.ptr(nodes), C.byref(nn),
                                          N.ptr(idx)))
    return nodes[: nn.value], idx[:n]
