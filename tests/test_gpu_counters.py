"""Device ray counters (zt_gpu_stats.extend_rays / shadow_rays) against the
oracle's instrumented traceWith.

An extension ray is one Scene.intersect call of the bounce loop
(render.zig:73), a shadow ray one of directLight (render.zig:51). The fp64
verification path traces the reference's paths, so its counts equal the
oracle's exactly wherever its pixels do; every fp32 schedule traces the same
paths as every other (their images are bit-identical), so their counts agree
exactly, and differ from the oracle's only by decision flips.
"""
import numpy as np
import pytest

import zigtrace_amd as zt
from zigtrace_amd import _native as N
from zigtrace_amd import workloads as W
from oracle import oracle as O

from .conftest import scene_text

pytestmark = pytest.mark.gpu

MK = N.METHOD_MEGAKERNEL
WF = N.METHOD_WAVEFRONT
F64 = N.PRECISION_FP64

CASES = [
    ("cornell", (48, 48), 8, 16, 0),
    ("spheres", (60, 34), 4, 16, 1),
    ("glass", (48, 27), 6, 16, 2),
    ("lamps", (48, 27), 6, 12, 3),
]


def counts(ctx):
    s = ctx.stats()
    return s["extend_rays"], s["shadow_rays"]


@pytest.mark.parametrize("name,size,spp,depth,seed", CASES)
def test_fp64_ray_counts_equal_oracle(gpu_ctx, name, size, spp, depth, seed):
    text = scene_text(name)
    sid = gpu_ctx.upload(zt.parse(text))
    ref = O.OracleScene.parse(text)
    w, h = size
    o = zt.Options(w, h, spp, depth, seed=seed)
    got = gpu_ctx.render(sid, o, zt.Tuning(method=MK, precision=F64))
    ext, sh = counts(gpu_ctx)
    assert gpu_ctx.stats()["paths"] == w * h * spp
    want, ctr = ref.render(w, h, spp, depth, seed=seed)
    close = np.isclose(got, want, rtol=1e-9, atol=1e-12).all(axis=2)
    assert ext > w * h * spp  # every sample traces at least its camera ray
    if close.all():
        assert (ext, sh) == (int(ctr["extend_rays"]), int(ctr["shadow_rays"])), name
    else:
        # A pixel whose path diverged (a last-ulp difference of a
        # transcendental) may trace other rays: at most 2*depth per sample.
        bound = int((~close).sum()) * spp * 2 * depth
        assert abs(ext - int(ctr["extend_rays"])) + abs(sh - int(ctr["shadow_rays"])) <= bound, name
    # The wavefront renderer traces the same paths.
    gpu_ctx.render(sid, o, zt.Tuning(method=WF, precision=F64, queue_paths=300))
    assert counts(gpu_ctx) == (ext, sh)
    gpu_ctx.release(sid)


def test_fp64_trace_rays_counts_equal_oracle(gpu_ctx):
    text = scene_text("lamps")
    sid = gpu_ctx.upload(zt.parse(text))
    ref = O.OracleScene.parse(text)
    rng = np.random.default_rng(8)
    n = 1500
    rays = np.zeros(n, dtype=N.RAY)
    rays["origin"] = [0, 2.2, 6]
    d = np.stack([rng.uniform(-0.4, 0.4, n), rng.uniform(-0.5, -0.1, n), -np.ones(n)], 1)
    rays["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    states = np.stack([O.Rng(77 + i).s for i in range(n)]).astype(np.uint64)
    s_dev, s_ref = states.copy(), states.copy()
    gpu_ctx.trace_rays(sid, rays, s_dev, 12, True, F64)
    ext, sh = counts(gpu_ctx)
    _, ctr = ref.trace(rays.view(O.RAY), s_ref, 12, True)
    gpu_ctx.release(sid)
    if (s_dev == s_ref).all():
        assert (ext, sh) == (int(ctr["extend_rays"]), int(ctr["shadow_rays"]))
    else:
        bound = int((~(s_dev == s_ref).all(axis=1)).sum()) * 2 * 12
        assert abs(ext - int(ctr["extend_rays"])) + abs(sh - int(ctr["shadow_rays"])) <= bound


@pytest.mark.parametrize("name,size,spp,depth,seed", CASES)
def test_fp32_counts_agree_across_schedules_and_with_oracle(gpu_ctx, name, size, spp, depth, seed):
    text = scene_text(name)
    sid = gpu_ctx.upload(zt.parse(text))
    w, h = size
    o = zt.Options(w, h, spp, depth, seed=seed)
    base = gpu_ctx.render(sid, o, zt.Tuning(method=MK))
    c0 = counts(gpu_ctx)
    for tu in (zt.Tuning(method=MK, schedule=1), zt.Tuning(method=WF, queue_paths=500),
               zt.Tuning(method=WF, queue_paths=4500, schedule=3), zt.Tuning(method=WF, schedule=1),
               zt.Tuning(method=WF, queue_paths=1, schedule=4), zt.Tuning(method=WF, queue_paths=1, schedule=5),
               zt.Tuning(method=MK, tile_order=2), zt.Tuning(method=MK, bvh_width=2, schedule=1)):
        img = gpu_ctx.render(sid, o, tu)
        assert counts(gpu_ctx) == c0, tu
        if tu.bvh_width != 2:  # the two-child walk may resolve an exact t tie the other way
            assert np.array_equal(img, base)
    _, ctr = O.OracleScene.parse(text).render(w, h, spp, depth, seed=seed)
    want = (int(ctr["extend_rays"]), int(ctr["shadow_rays"]))
    # Decision flips between fp32 and f64 are rare: the totals agree to 1%.
    assert abs(c0[0] - want[0]) <= 0.01 * want[0] + 2, (c0, want)
    assert abs(c0[1] - want[1]) <= 0.01 * max(want[1], 100) + 2, (c0, want)
    gpu_ctx.release(sid)


def test_batch_counts_are_the_sum_of_single_renders(gpu_ctx):
    ws = W.batch_scenes(count=5, seed=4, min_prims=40, max_prims=3000, resolutions=(16, 24), spps=(2, 3),
                        depth_range=(3, 10))
    ids = [gpu_ctx.upload(w.scene) for w in ws]
    opts = [w.options for w in ws]
    single = np.zeros(2, dtype=np.int64)
    for sid, o in zip(ids, opts):
        gpu_ctx.render(sid, o, zt.Tuning(method=MK))
        single += counts(gpu_ctx)
    gpu_ctx.render_batch(ids, opts, zt.Tuning())  # one wave-scheduled launch
    assert counts(gpu_ctx) == tuple(single)
    gpu_ctx.render_batch(ids, opts, zt.Tuning(method=WF, queue_paths=700))  # the shared path pool
    assert counts(gpu_ctx) == tuple(single)
    for sid in ids:
        gpu_ctx.release(sid)
