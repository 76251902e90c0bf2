"""Shared test helpers: scenes built identically for the product and the oracle."""
from __future__ import annotations

import numpy as np

import zigtrace_amd as zt
from zigtrace_amd import _native as N
from oracle import oracle as O


def oracle_of(scene: zt.Scene, build: bool = True) -> O.OracleScene:
    """The oracle's own Scene built from the product scene's flat inputs."""
    return O.OracleScene.from_view_arrays(scene.view_arrays(), build=build)


def random_shapes(rng: np.random.Generator, n: int) -> np.ndarray:
    """bvh.zig:272-289: half spheres (centre U(-10,10)^3, radius U(0.05,1.5)),
    half triangles (c + U(-2,2)^3 per vertex); shape i uses material i."""
    s = np.zeros(n, dtype=N.SHAPE)
    for i in range(n):
        s[i]["material"] = i
        if rng.random() < 0.5:
            s[i]["kind"] = N.SPHERE
            s[i]["v"][:3] = rng.uniform(-10, 10, 3)
            s[i]["v"][3] = rng.uniform(0.05, 1.5)
        else:
            c = rng.uniform(-10, 10, 3)
            s[i]["kind"] = N.TRIANGLE
            s[i]["v"] = np.concatenate([c + rng.uniform(-2, 2, 3) for _ in range(3)])
    return s


def random_scene(rng, n: int, planes: int = 0) -> zt.Scene:
    sc = zt.Scene()
    for i in range(n + planes):
        sc.add_material(zt.lambertian((0.5, 0.5, 0.5)))
    shapes = random_shapes(rng, n)
    sc.add_many(shapes)
    for j in range(planes):
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        sc.add(zt.plane(rng.uniform(-12, 12, 3), nrm, n + j))
    sc.build()
    return sc


def aimed_rays(rng, shapes: np.ndarray, count: int) -> np.ndarray:
    """bvh.zig:291-301: origins U(-15,15)^3; 80% aimed near a random primitive."""
    r = np.zeros(count, dtype=N.RAY)
    for k in range(count):
        o = rng.uniform(-15, 15, 3)
        if rng.random() < 0.8 and len(shapes):
            s = shapes[rng.integers(len(shapes))]
            if s["kind"] == N.SPHERE:
                c = s["v"][:3]
            else:
                v = s["v"].reshape(3, 3)
                c = (v.min(0) + v.max(0)) * 0.5
            d = c + rng.uniform(-0.5, 0.5, 3) - o
        else:
            d = rng.normal(size=3)
        r[k]["origin"] = o
        r[k]["dir"] = d / np.linalg.norm(d)
    return r


def to_oracle_rays(r: np.ndarray) -> np.ndarray:
    return r.view(O.RAY)


def block_mean(img: np.ndarray, b: int) -> np.ndarray:
    h, w = img.shape[:2]
    return img[: h // b * b, : w // b * b].reshape(h // b, b, w // b, b, -1).mean(axis=(1, 3))
