"""Seeded synthetic scenes for the benchmark workloads and tests.

The reference ships scene files but no generators, so these are this
package's own. Each returns a built `Scene` plus the render settings it is
benchmarked at. All randomness comes from numpy's PCG64 with a fixed seed.

  cornell()       examples/cornell.txt as shipped, 400x400, 64 spp, depth 16
  sphere_field()  ground plane + 3 feature spheres + ~320 small spheres, 600x340, 16 spp
  mesh_field()    64 displaced icospheres (subdivision 5 -> 1,310,720 triangles) over a
                  Lambertian plane, 64 small emissive spheres + sky, 1920x1080, 64 spp
  diff_box()      Cornell-style box, 10 triangles + 2 spheres, 8 albedo materials + 1 emitter
  batch_scenes()  many independent scenes of uneven size and settings
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .render import Options
from .scene import CameraConfig, Scene, dielectric, emissive, lambertian, metal, sky, solid

HERE = os.path.dirname(os.path.abspath(__file__))
SCENE_DIR = os.path.join(os.path.dirname(HERE), "examples")


@dataclass
class Workload:
    name: str
    scene: Scene
    options: Options
    description: str


def _shapes(n: int) -> np.ndarray:
    return np.zeros(n, dtype=N.SHAPE)


def cornell(spp: int = 64, max_depth: int = 16, width: int = 400, height: int = 400, seed: int = 0) -> Workload:
    from .scenefile import load
    s = load(os.path.join(SCENE_DIR, "cornell.txt"))
    return Workload("cornell", s, Options(width, height, spp, max_depth, seed=seed),
                    "examples/cornell.txt as shipped (12 triangles, 2 spheres)")


def sphere_field(seed: int = 2024, spp: int = 16, max_depth: int = 16, width: int = 600, height: int = 340,
                 ground_plane: bool = True) -> Workload:
    """Ground plane y=0 (Lambertian 0.5), feature spheres r=1 at (0,1,0) glass 1.5, (-4,1,0)
    Lambertian (0.4,0.2,0.1), (4,1,0) metal (0.7,0.6,0.5); small spheres r=0.2 at
    (a+0.9u, 0.2, b+0.9u), a,b in [-11,11), rejected within 1.2 of a feature sphere or
    0.75 of an earlier small sphere (about 320 remain); materials 80%
    Lambertian (albedo = u*u per channel), 15% metal (albedo U[0.5,1], fuzz U[0,0.5]),
    5% dielectric 1.5. Thin-lens camera (13,2,3)->(0,0,0), fov 20, aperture 0.1, focus 10."""
    rng = np.random.default_rng(seed)
    s = Scene()
    ground = s.add_material(lambertian((0.5, 0.5, 0.5)))
    glass = s.add_material(dielectric(1.5))
    brown = s.add_material(lambertian((0.4, 0.2, 0.1)))
    polished = s.add_material(metal((0.7, 0.6, 0.5), 0.0))
    if ground_plane:
        s.add_plane((0, 0, 0), (0, 1, 0), ground)
    else:
        s.add_sphere((0, -1000, 0), 1000, ground)
    feature = [np.array([0.0, 1.0, 0.0]), np.array([-4.0, 1.0, 0.0]), np.array([4.0, 1.0, 0.0])]
    centers = []
    shapes = []
    for a in range(-11, 11):
        for b in range(-11, 11):
            c = np.array([a + 0.9 * rng.random(), 0.2, b + 0.9 * rng.random()])
            if any(np.linalg.norm(c - f) < 1.2 for f in feature):
                continue
            if any(np.linalg.norm(c - p) < 0.75 for p in centers):
                continue
            u = rng.random()
            if u < 0.8:
                m = s.add_material(lambertian(tuple(rng.random(3) * rng.random(3))))
            elif u < 0.95:
                m = s.add_material(metal(tuple(0.5 + 0.5 * rng.random(3)), 0.5 * rng.random()))
            else:
                m = glass
            centers.append(c)
            shapes.append((c, m))
    arr = _shapes(len(shapes) + 3)
    for i, (c, m) in enumerate(shapes):
        arr[i]["kind"] = N.SPHERE
        arr[i]["material"] = m
        arr[i]["v"][:4] = [c[0], c[1], c[2], 0.2]
    for j, (c, m) in enumerate(zip(feature, (glass, brown, polished))):
        k = len(shapes) + j
        arr[k]["kind"] = N.SPHERE
        arr[k]["material"] = m
        arr[k]["v"][:4] = [c[0], c[1], c[2], 1.0]
    s.add_many(arr)
    s.camera = CameraConfig((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, 0.1, 10.0)
    s.background = sky((1, 1, 1), (0.5, 0.7, 1.0))
    s.build()
    return Workload("sphere_field", s, Options(width, height, spp, max_depth),
                    f"ground plane + 3 feature spheres + {len(shapes)} small spheres")


def icosphere(subdivisions: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: (vertices (V,3), faces (F,3) int), faces counter-clockwise seen from outside."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(subdivisions):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = base + inv[:nf], base + inv[nf:2 * nf], base + inv[2 * nf:]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    return v, f


def mesh_field(seed: int = 7, subdivisions: int = 5, meshes: int = 64, spp: int = 64, max_depth: int = 16,
               width: int = 1920, height: int = 1080, noise: float = 0.05) -> Workload:
    """64 displaced icospheres on an 8x8 grid over a Lambertian plane, mixed Lambertian /
    metal / dielectric per mesh, lit by 64 emissive spheres (r=0.05, radiance U[5,50]) and a sky."""
    rng = np.random.default_rng(seed)
    unit_v, faces = icosphere(subdivisions)
    s = Scene()
    floor = s.add_material(lambertian((0.6, 0.6, 0.58)))
    s.add_plane((0, 0, 0), (0, 1, 0), floor)
    side = int(round(meshes ** 0.5))
    tris = []
    for k in range(meshes):
        i, j = divmod(k, side)
        radius = 0.55 + 0.25 * rng.random()
        center = np.array([(i - (side - 1) / 2) * 2.0, radius * (1 + noise), (j - (side - 1) / 2) * 2.0])
        # Smooth radial noise: a few random plane waves over the sphere directions.
        freqs = rng.normal(size=(6, 3)) * 4.0
        phases = rng.random(6) * 2 * np.pi
        disp = np.sin(unit_v @ freqs.T + phases).mean(axis=1)
        disp /= max(np.abs(disp).max(), 1e-9)
        verts = center + unit_v * (radius * (1.0 + noise * disp))[:, None]
        u = rng.random()
        if u < 0.6:
            m = s.add_material(lambertian(tuple(0.2 + 0.7 * rng.random(3))))
        elif u < 0.85:
            m = s.add_material(metal(tuple(0.5 + 0.5 * rng.random(3)), 0.3 * rng.random()))
        else:
            m = s.add_material(dielectric(1.5))
        tri = _shapes(len(faces))
        tri["kind"] = N.TRIANGLE
        tri["material"] = m
        tri["v"] = np.concatenate([verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]], axis=1)
        tris.append(tri)
    lights = _shapes(meshes)
    for k in range(meshes):
        le = 5.0 + 45.0 * rng.random()
        m = s.add_material(emissive((le, le * (0.7 + 0.3 * rng.random()), le * (0.5 + 0.5 * rng.random()))))
        extent = side * 2.0 / 2
        lights[k]["kind"] = N.SPHERE
        lights[k]["material"] = m
        lights[k]["v"][:4] = [rng.uniform(-extent, extent), 2.2 + 1.5 * rng.random(), rng.uniform(-extent, extent),
                              0.05]
    s.add_many(np.concatenate(tris + [lights]))
    s.camera = CameraConfig((0.0, 7.5, 17.0), (0.0, 0.3, 0.0), (0, 1, 0), 42.0)
    s.background = sky((0.25, 0.28, 0.32), (0.05, 0.08, 0.16))
    s.build()
    n_tri = meshes * len(faces)
    return Workload("mesh_field", s, Options(width, height, spp, max_depth),
                    f"{meshes} displaced icospheres (subdivision {subdivisions}), {n_tri} triangles, "
                    f"{meshes} emissive spheres r=0.05, Lambertian plane, sky")


# Parameter layout of the differentiable box: the albedo of materials 0..7
# (RGB each) and the emitter radiance of material 8.
DIFF_ALBEDO_MATERIALS = 8


def diff_box(albedos=None, radiance=None, spp: int = 32, max_depth: int = 8, width: int = 256,
             height: int = 256) -> Workload:
    """Cornell-style box: floor and back wall split into two triangles with their own
    materials, left/right walls, a downward-facing quad light (10 triangles), a diffuse
    and a metal sphere. Materials 0-7 carry albedo parameters, material 8 the emitter."""
    if albedos is None:
        albedos = [(0.73, 0.73, 0.73), (0.70, 0.72, 0.70), (0.73, 0.70, 0.68), (0.68, 0.71, 0.73),
                   (0.65, 0.05, 0.05), (0.12, 0.45, 0.15), (0.75, 0.55, 0.25), (0.85, 0.85, 0.88)]
    if radiance is None:
        radiance = (15.0, 15.0, 15.0)
    s = Scene()
    ids = [s.add_material(lambertian(a)) for a in albedos[:7]] + [s.add_material(metal(albedos[7], 0.05))]
    lamp = s.add_material(emissive(radiance))
    L = 555.0
    tri = [
        ((0, 0, 0), (L, 0, 0), (L, 0, L), ids[0]), ((0, 0, 0), (L, 0, L), (0, 0, L), ids[1]),      # floor
        ((0, 0, L), (L, 0, L), (L, L, L), ids[2]), ((0, 0, L), (L, L, L), (0, L, L), ids[3]),      # back
        ((L, 0, 0), (L, L, 0), (L, L, L), ids[4]), ((L, 0, 0), (L, L, L), (L, 0, L), ids[4]),      # left (red)
        ((0, 0, 0), (0, 0, L), (0, L, L), ids[5]), ((0, 0, 0), (0, L, L), (0, L, 0), ids[5]),      # right (green)
    ]
    for a, b, c, m in tri:
        s.add_triangle(a, b, c, m)
    s.add_quad((343, 554, 332), (-130, 0, 0), (0, 0, -105), lamp)
    s.add_sphere((190, 90, 190), 90, ids[6])
    s.add_sphere((380, 120, 360), 120, ids[7])
    s.camera = CameraConfig((278, 278, -800), (278, 278, 0), (0, 1, 0), 40.0)
    s.background = solid((0, 0, 0))
    s.build()
    return Workload("diff_box", s, Options(width, height, spp, max_depth),
                    "Cornell-style box, 10 triangles + 2 spheres, 8 albedo materials + 1 emitter")


def batch_scenes(count: int = 256, seed: int = 11, min_prims: int = 100, max_prims: int = 200_000,
                 resolutions=(64, 128, 256, 512), spps=(4, 16, 64), depth_range=(8, 32)) -> list:
    """Independent scenes of uneven size and settings for the batched renderer: primitive
    counts log-uniform in [min_prims, max_prims], half sphere fields over a ground plane and
    half triangle soups, mixed Lambertian/metal/dielectric materials, 1-4 small emissive
    spheres plus a sky; square resolution, spp and depth drawn per scene."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(min_prims), np.log(max_prims)))))
        res = int(rng.choice(resolutions))
        spp = int(rng.choice(spps))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        s = Scene()
        mats = []
        for _ in range(12):
            u = rng.random()
            if u < 0.6:
                mats.append(s.add_material(lambertian(tuple(0.1 + 0.8 * rng.random(3)))))
            elif u < 0.85:
                mats.append(s.add_material(metal(tuple(0.5 + 0.5 * rng.random(3)), 0.4 * rng.random())))
            else:
                mats.append(s.add_material(dielectric(1.3 + 0.4 * rng.random())))
        mats = np.array(mats)
        half = 10.0
        kind = "spheres" if rng.random() < 0.5 else "soup"
        arr = _shapes(n)
        arr["material"] = mats[rng.integers(0, len(mats), n)]
        if kind == "spheres":
            ground = s.add_material(lambertian((0.5, 0.5, 0.5)))
            s.add_plane((0, -half, 0), (0, 1, 0), ground)
            r = 0.35 * half / n ** (1 / 3)
            arr["kind"] = N.SPHERE
            arr["v"][:, :3] = rng.uniform(-half, half, (n, 3))
            arr["v"][:, 3] = r * rng.uniform(0.5, 1.5, n)
        else:
            edge = 1.5 * half / n ** (1 / 3)
            c = rng.uniform(-half, half, (n, 3))
            arr["kind"] = N.TRIANGLE
            arr["v"] = np.concatenate([c + rng.uniform(-edge, edge, (n, 3)) for _ in range(3)], axis=1)
        nl = int(rng.integers(1, 5))
        lights = _shapes(nl)
        for k in range(nl):
            le = 5 + 45 * rng.random()
            m = s.add_material(emissive((le, le, le * (0.6 + 0.4 * rng.random()))))
            lights[k]["kind"] = N.SPHERE
            lights[k]["material"] = m
            lights[k]["v"][:4] = list(rng.uniform(-half, half, 3)) + [0.3 + 0.3 * rng.random()]
        s.add_many(np.concatenate([arr, lights]))
        eye = rng.normal(size=3)
        eye = eye / np.linalg.norm(eye) * half * 3.2
        s.camera = CameraConfig(tuple(eye), (0, 0, 0), (0, 1, 0), 40.0)
        s.background = sky((0.6, 0.6, 0.6), (0.2, 0.3, 0.6))
        s.build()
        out.append(Workload(f"batch{i}_{kind}", s, Options(res, res, spp, depth, seed=i),
                            f"{kind}, {n} primitives, {res}x{res}, {spp} spp, depth {depth}"))
    return out
