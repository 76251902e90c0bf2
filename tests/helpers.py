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


def seed_mean_test(gpu: list, cpu: list, block: int, z: float = 4.0):
    """Per-block means over independent seeds, GPU vs oracle: returns (beyond, cells, expected).

    `gpu`, `cpu`: renders (H, W, 3) with independent seeds. Each b x b block
    and channel is one cell; `beyond` counts cells whose mean difference
    exceeds z standard errors (Welch), `expected` is how many of them the
    t distribution of that statistic puts beyond z when both estimators have
    the same mean (the seed-level variances are estimated from few seeds, so
    that is not the Gaussian 6e-5 per cell)."""
    from scipy import stats
    g = np.stack([block_mean(x, block) for x in gpu])
    c = np.stack([block_mean(x, block) for x in cpu])
    k1, k2 = len(g), len(c)
    vg, vc = g.var(axis=0, ddof=1) / k1, c.var(axis=0, ddof=1) / k2
    se = np.sqrt(vg + vc)
    diff = np.abs(g.mean(axis=0) - c.mean(axis=0))
    zz = np.where(se > 0, diff / np.maximum(se, 1e-300), np.where(diff > 1e-6, np.inf, 0))
    den = vg ** 2 / (k1 - 1) + vc ** 2 / (k2 - 1)
    df = np.where(den > 0, (vg + vc) ** 2 / np.maximum(den, 1e-300), 1e9)
    expected = float((2 * stats.t.sf(z, df)).sum())
    return int((zz > z).sum()), int(zz.size), expected


def image_mean_within(gpu: list, cpu: list, z: float = 4.0) -> tuple[bool, np.ndarray]:
    """Whole-image mean per channel over seeds agrees within z standard errors."""
    g = np.stack([x.reshape(-1, 3).mean(axis=0) for x in gpu])
    c = np.stack([x.reshape(-1, 3).mean(axis=0) for x in cpu])
    se = np.sqrt(g.var(axis=0, ddof=1) / len(g) + c.var(axis=0, ddof=1) / len(c))
    zz = np.abs(g.mean(axis=0) - c.mean(axis=0)) / np.maximum(se, 1e-300)
    return bool((zz <= z).all()), zz


def strided(img: np.ndarray, stride: int) -> np.ndarray:
    """The pixels (y*W + x) % stride == 0 of an (H, W, 3) image, as the oracle's pixel_stride renders them."""
    flat = img.reshape(-1, 3)
    return flat[np.arange(len(flat)) % stride == 0]


def rel_rmse(got: np.ndarray, want: np.ndarray) -> float:
    return float(np.sqrt(np.mean((got - want) ** 2)) / np.sqrt(np.mean(want ** 2)))
