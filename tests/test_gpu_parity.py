"""Device path vs the oracle, through the C ABI (zt_gpu.h / zt_host.h).

* RNG streams: bit-exact (rng.zig:13-38).
* BVH traversal: exact t and material equal to the device brute force
  (bvh.zig:316-325), in both precisions; the fp64 path equals the oracle.
* fp64 verification path: the reference's f64 arithmetic on the device; its
  images and traceWith results must equal the oracle's to rounding of the
  transcendental functions (tolerance 1e-9 relative), pixel for pixel.
"""
import numpy as np
import pytest

import zigtrace_amd as zt
from zigtrace_amd import _native as N
from oracle import oracle as O

from .conftest import scene_text
from .helpers import aimed_rays, oracle_of, random_scene, to_oracle_rays

pytestmark = pytest.mark.gpu

SEEDS = [0, 1, 2**64 - 1]
STREAMS = [0, 1, 159999, 2**63]


def test_rng_streams_bit_exact(gpu_ctx):
    for seed in SEEDS:
        got = gpu_ctx.rng(seed, STREAMS, 8)
        for i, st in enumerate(STREAMS):
            want = O.Rng.for_stream(seed, st).next(8)
            assert np.array_equal(got[i], want), (seed, st)


def test_rng_many_streams_bit_exact(gpu_ctx):
    streams = np.arange(0, 2_000_000, 997, dtype=np.uint64)
    got = gpu_ctx.rng(12345, streams, 3)
    for i in range(0, len(streams), 101):
        assert np.array_equal(got[i], O.Rng.for_stream(12345, int(streams[i])).next(3))


def test_fp32_uniform_is_truncated_f64_uniform(gpu_ctx):
    bits = gpu_ctx.rng(7, [5, 6], 64, kind=1).astype(np.uint32).view(np.float32).astype(np.float64)
    for i, st in enumerate([5, 6]):
        f64 = O.Rng.for_stream(7, st).float(64)
        assert np.all(bits[i] <= f64) and np.all(f64 - bits[i] < 2.0**-24)
        assert np.all((bits[i] >= 0) & (bits[i] < 1))


@pytest.mark.parametrize("n", [1, 2, 3, 7, 50, 400, 5000])
@pytest.mark.parametrize("precision", [N.PRECISION_FP32, N.PRECISION_FP64])
@pytest.mark.parametrize("mode", [0, 2, 3])  # default (fp32: four-wide, 128-byte nodes); two-child; 64-byte
def test_bvh_traversal_equals_brute_force(gpu_ctx, n, precision, mode):
    rng = np.random.default_rng(2026 + n)
    sc = random_scene(rng, n)
    sid = gpu_ctx.upload(sc)
    rays = aimed_rays(rng, sc.view_arrays()["shapes"], 3000)
    bvh = gpu_ctx.intersect(sid, rays, 1e-6, 1e9, mode=mode, precision=precision)
    brute = gpu_ctx.intersect(sid, rays, 1e-6, 1e9, mode=1, precision=precision)
    gpu_ctx.release(sid)
    assert np.array_equal(bvh["hit"], brute["hit"])
    h = bvh["hit"] == 1
    assert np.array_equal(bvh["t"][h], brute["t"][h])
    assert np.array_equal(bvh["material"][h], brute["material"][h])
    assert h.sum() > 0
    if n >= 7:
        assert h.sum() > len(rays) // 4


@pytest.mark.parametrize("precision", [N.PRECISION_FP32, N.PRECISION_FP64])
@pytest.mark.parametrize("mode", [0, 2, 3])
def test_axis_parallel_rays_equal_brute_force(gpu_ctx, precision, mode):
    """Directions with components of exactly +0 or -0 (a uniform draw of exactly
    one half makes them, e.g. off an axis-aligned plane) have infinite
    reciprocals, which the four-wide walker holds to +-2^100; every node format
    must still find brute force's hits, in both precisions."""
    rng = np.random.default_rng(77)
    sc = random_scene(rng, 400)
    sid = gpu_ctx.upload(sc)
    shapes = sc.view_arrays()["shapes"]
    rays = np.zeros(3000, dtype=N.RAY)
    for k in range(len(rays)):
        s = shapes[rng.integers(len(shapes))]
        c = s["v"][:3] if s["kind"] == N.SPHERE else s["v"].reshape(3, 3).mean(0)
        d = np.zeros(3)
        axes = rng.permutation(3)[: 1 + k % 2]  # one axis, or a coordinate plane
        d[axes] = rng.choice([-1.0, 1.0], len(axes)) * rng.uniform(0.2, 1.0, len(axes))
        d /= np.linalg.norm(d)
        o = c - 20.0 * d + np.where(d == 0, rng.uniform(-0.3, 0.3, 3), 0.0)
        zero = d == 0
        d[zero] = np.where(rng.random(zero.sum()) < 0.5, 0.0, -0.0)
        rays[k]["origin"] = o
        rays[k]["dir"] = d
    bvh = gpu_ctx.intersect(sid, rays, 1e-6, 1e9, mode=mode, precision=precision)
    brute = gpu_ctx.intersect(sid, rays, 1e-6, 1e9, mode=1, precision=precision)
    gpu_ctx.release(sid)
    assert np.array_equal(bvh["hit"], brute["hit"])
    h = bvh["hit"] == 1
    assert h.sum() > len(rays) // 4
    assert np.array_equal(bvh["t"][h], brute["t"][h])
    assert np.array_equal(bvh["material"][h], brute["material"][h])


def test_fp64_intersect_equals_oracle(gpu_ctx):
    rng = np.random.default_rng(99)
    sc = random_scene(rng, 300, planes=2)
    ref = oracle_of(sc)
    sid = gpu_ctx.upload(sc)
    rays = aimed_rays(rng, sc.view_arrays()["shapes"], 4000)
    got = gpu_ctx.intersect(sid, rays, 1e-4, 1e9, precision=N.PRECISION_FP64)
    want, _ = ref.intersect(to_oracle_rays(rays), 1e-4, 1e9)
    gpu_ctx.release(sid)
    assert np.array_equal(got["hit"], want["hit"])
    h = got["hit"] == 1
    assert np.array_equal(got["t"][h], want["t"][h])
    assert np.array_equal(got["material"][h], want["material"][h])
    assert np.array_equal(got["normal"][h], want["n"][h])
    assert np.array_equal(got["front_face"][h], want["front"][h])


@pytest.mark.parametrize("name,size,spp,depth,seed", [
    ("cornell", (48, 48), 8, 16, 0),
    ("spheres", (60, 34), 4, 16, 1),
    ("glass", (48, 27), 6, 16, 2),
    ("lamps", (48, 27), 6, 12, 3),
])
def test_fp64_render_equals_oracle(name, size, spp, depth, seed):
    text = scene_text(name)
    sc = zt.parse(text)
    ref = O.OracleScene.parse(text)
    w, h = size
    got = zt.render(sc, zt.Options(w, h, spp, depth, seed=seed), zt.Tuning(precision=N.PRECISION_FP64))
    want, _ = ref.render(w, h, spp, depth, seed=seed)
    close = np.isclose(got, want, rtol=1e-9, atol=1e-12).all(axis=2)
    assert close.mean() >= 0.995, f"{name}: {close.mean():.4f} of pixels match"


def test_fp64_trace_rays_equals_oracle(gpu_ctx):
    # traceWith with caller-owned Rng.init streams (render.zig:290-302 style).
    text = scene_text("lamps")
    sc = zt.parse(text)
    ref = O.OracleScene.parse(text)
    sid = gpu_ctx.upload(sc)
    rng = np.random.default_rng(5)
    n = 2000
    rays = np.zeros(n, dtype=N.RAY)
    rays["origin"] = [0, 2.2, 6]
    d = np.stack([rng.uniform(-0.4, 0.4, n), rng.uniform(-0.5, -0.1, n), -np.ones(n)], 1)
    rays["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    states = np.stack([O.Rng(1000 + i).s for i in range(n)]).astype(np.uint64)
    s_dev = states.copy()
    s_ref = states.copy()
    got = gpu_ctx.trace_rays(sid, rays, s_dev, 12, True, N.PRECISION_FP64)
    want, _ = ref.trace(to_oracle_rays(rays), s_ref, 12, True)
    gpu_ctx.release(sid)
    same_state = (s_dev == s_ref).all(axis=1)
    close = np.isclose(got, want, rtol=1e-9, atol=1e-12).all(axis=1)
    assert same_state.mean() >= 0.995
    assert close.mean() >= 0.995
