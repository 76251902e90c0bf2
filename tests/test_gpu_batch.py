"""The batched renderer: many scenes of uneven size, spp and depth in one
persistent pass (the wave-scheduled megakernel over every scene's tiles, or
the wavefront renderer's shared path pool) give each scene's single-render
image bit for bit."""
import numpy as np
import pytest

import zigtrace_amd as zt
from zigtrace_amd import _native as N
from zigtrace_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("queue,schedule", [(0, 0), (4096, 2), (0, 3), (4500, 3), (700, 1), (0, 4), (1, 4), (0, 5), (1, 5)])
def test_batch_equals_individual_renders(gpu_ctx, queue, schedule):
    ws = W.batch_scenes(count=10, seed=3, min_prims=50, max_prims=5000, resolutions=(16, 32, 48), spps=(1, 2, 4),
                        depth_range=(2, 12))
    ids = [gpu_ctx.upload(w.scene) for w in ws]
    opts = [w.options for w in ws]
    outs = gpu_ctx.render_batch(ids, opts, zt.Tuning(method=N.METHOD_WAVEFRONT, queue_paths=queue, schedule=schedule))
    for sid, o, img in zip(ids, opts, outs):
        single = gpu_ctx.render(sid, o, zt.Tuning(method=N.METHOD_MEGAKERNEL))
        assert np.array_equal(single, img)
        assert np.isfinite(img).all()
    for sid in ids:
        gpu_ctx.release(sid)


@pytest.mark.parametrize("schedule", [4, 5])
def test_batch_pool_smaller_than_the_pixels(gpu_ctx, schedule):
    """Segmented schedules over a batch with more pixels (about 88k) than the
    smallest pool (8 segments of 4096 slots): slots take pixels of other
    scenes as theirs finish, bands run out at different times, and the tail
    goes to the drain (which for sample passes starts each slot's next
    sample)."""
    ws = W.batch_scenes(count=5, seed=9, min_prims=50, max_prims=3000, resolutions=(112, 160), spps=(1, 3, 5),
                        depth_range=(2, 10))
    ids = [gpu_ctx.upload(w.scene) for w in ws]
    opts = [w.options for w in ws]
    assert sum(o.width * o.height for o in opts) > 8 * 4096
    outs = gpu_ctx.render_batch(ids, opts, zt.Tuning(method=N.METHOD_WAVEFRONT, queue_paths=1, schedule=schedule))
    for sid, o, img in zip(ids, opts, outs):
        assert np.array_equal(gpu_ctx.render(sid, o, zt.Tuning(method=N.METHOD_MEGAKERNEL)), img)
    for sid in ids:
        gpu_ctx.release(sid)


def test_batch_fp64_and_repeated_scene(gpu_ctx):
    ws = W.batch_scenes(count=3, seed=5, min_prims=20, max_prims=500, resolutions=(16,), spps=(2,),
                        depth_range=(4, 8))
    ids = [gpu_ctx.upload(w.scene) for w in ws]
    # The same scene twice with different seeds and sizes.
    opts = [ws[0].options, zt.Options(24, 8, 3, 6, seed=42), ws[1].options, ws[2].options]
    ids4 = [ids[0], ids[0], ids[1], ids[2]]
    tu = zt.Tuning(method=N.METHOD_WAVEFRONT, precision=N.PRECISION_FP64, block_size=64, queue_paths=100)
    outs = gpu_ctx.render_batch(ids4, opts, tu)
    for sid, o, img in zip(ids4, opts, outs):
        single = gpu_ctx.render(sid, o, zt.Tuning(method=N.METHOD_MEGAKERNEL, precision=N.PRECISION_FP64))
        assert np.array_equal(single, img)
    for sid in ids:
        gpu_ctx.release(sid)


@pytest.mark.parametrize("tuning", [
    dict(), dict(refill=8, tile_w=16, tile_h=8), dict(variant=1, stack_entries=8), dict(method=N.METHOD_MEGAKERNEL),
    dict(tile_order=2), dict(tile_order=2, tile_w=16, tile_h=16, refill=40),
])
def test_wave_megakernel_batch_equals_individual_renders(gpu_ctx, tuning):
    ws = W.batch_scenes(count=12, seed=9, min_prims=50, max_prims=20000, resolutions=(8, 16, 40, 64),
                        spps=(1, 3, 8), depth_range=(0, 12))
    ids = [gpu_ctx.upload(w.scene) for w in ws]
    opts = [w.options for w in ws]
    outs = gpu_ctx.render_batch(ids, opts, zt.Tuning(**tuning))
    for sid, o, img in zip(ids, opts, outs):
        single = gpu_ctx.render(sid, o, zt.Tuning(method=N.METHOD_MEGAKERNEL, schedule=1))
        assert np.array_equal(single, img), o
    for sid in ids:
        gpu_ctx.release(sid)


def test_batch_with_one_scene_and_mixed_layouts(gpu_ctx):
    """One scene alone; and a batch where one scene has no four-wide tree (rendered
    through the wavefront path as a whole) still matches single renders."""
    ws = W.batch_scenes(count=3, seed=21, min_prims=30, max_prims=3000, resolutions=(24,), spps=(2,),
                        depth_range=(3, 6))
    ids = [gpu_ctx.upload(w.scene) for w in ws]
    opts = [w.options for w in ws]
    one = gpu_ctx.render_batch(ids[:1], opts[:1])
    assert np.array_equal(one[0], gpu_ctx.render(ids[0], opts[0], zt.Tuning(method=N.METHOD_MEGAKERNEL)))
    u = zt.Scene()
    m = u.add_material(zt.lambertian((0.5, 0.5, 0.5)))
    u.add_sphere((0, 0, -3), 1, m)  # not built: brute force, no four-wide tree
    uid = gpu_ctx.upload(u)
    uo = zt.Options(16, 16, 2, 4)
    outs = gpu_ctx.render_batch(ids + [uid], opts + [uo])
    for sid, o, img in zip(ids + [uid], opts + [uo], outs):
        assert np.array_equal(img, gpu_ctx.render(sid, o, zt.Tuning(method=N.METHOD_MEGAKERNEL)))
    for sid in ids + [uid]:
        gpu_ctx.release(sid)
