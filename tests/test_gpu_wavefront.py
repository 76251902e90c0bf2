"""The wavefront renderer produces the megakernel's image bit for bit, for
any path-pool size (slot reuse across pixels, queues compacted over one or
many ranges of 4096 slots), block size, schedule (one ray per lane, the
queue-fed wave tracer with separate shading, the tracer shading each hit
itself, the whole bounce in one launch over XCD segments, whole samples per
launch) and precision."""
import ctypes as C

import numpy as np
import pytest

import zigtrace_amd as zt
from zigtrace_amd import _native as N

from .conftest import scene_text

pytestmark = pytest.mark.gpu

MEGA = N.METHOD_MEGAKERNEL
WAVE = N.METHOD_WAVEFRONT


@pytest.mark.parametrize("name,w,h,spp,depth", [
    ("cornell", 40, 40, 8, 16), ("spheres", 60, 34, 4, 16), ("glass", 48, 27, 6, 16), ("lamps", 48, 27, 6, 12),
])
@pytest.mark.parametrize("precision", [N.PRECISION_FP32, N.PRECISION_FP64])
def test_wavefront_equals_megakernel(name, w, h, spp, depth, precision):
    s = zt.parse(scene_text(name))
    o = zt.Options(w, h, spp, depth, seed=5)
    mega = zt.render(s, o, zt.Tuning(method=MEGA, precision=precision))
    schedules = (1, 2, 3, 4, 5) if precision == N.PRECISION_FP32 else (0, 4, 5)
    for block, queue in [(256, 0), (64, 64), (128, 1000), (192, 333), (256, 4097)]:
        for sched in schedules:
            wave = zt.render(s, o, zt.Tuning(method=WAVE, precision=precision, block_size=block, queue_paths=queue,
                                             schedule=sched))
            assert np.array_equal(mega, wave), (name, block, queue, sched)


def test_sample_passes_beyond_one_graph():
    """Sample passes run in replays of a 16-pass graph: 37 samples per pixel
    take three replays whose last has 11 passes with empty queues, and a pool
    smaller than the frame (queue_paths) claims pixels across replays and
    ends in the drain. Both give the megakernel's bits."""
    from zigtrace_amd import workloads
    wl = workloads.mesh_field(subdivisions=2, meshes=16, spp=37, width=48, height=40)
    s, o = wl.scene, wl.options
    mega = zt.render(s, o, zt.Tuning(method=MEGA, schedule=1))
    for queue, order in [(0, 0), (0, 1), (1, 1)]:
        wave = zt.render(s, o, zt.Tuning(method=WAVE, schedule=5, queue_paths=queue, tile_order=order))
        assert np.array_equal(mega, wave), (queue, order)
    big = workloads.mesh_field(subdivisions=2, meshes=16, spp=64, width=256, height=160)
    mega = zt.render(big.scene, big.options, zt.Tuning(method=MEGA, schedule=1))
    wave = zt.render(big.scene, big.options, zt.Tuning(method=WAVE, schedule=5, queue_paths=1))
    assert np.array_equal(mega, wave)


def test_wavefront_max_depth_zero_and_single_sample():
    s = zt.parse(scene_text("cornell"))
    for depth, spp in [(0, 3), (5, 1)]:
        o = zt.Options(17, 9, spp, depth, seed=1)
        mega = zt.render(s, o, zt.Tuning(method=MEGA))
        for sched in (0, 4, 5):
            wave = zt.render(s, o, zt.Tuning(method=WAVE, queue_paths=50, schedule=sched))
            assert np.array_equal(mega, wave), sched
            if depth == 0:
                assert not wave.any()


@pytest.mark.parametrize("w,h,queue", [(256, 160, 1), (256, 160, 40960), (301, 7, 0), (3, 2, 0), (37, 29, 1)])
@pytest.mark.parametrize("schedule,order", [(4, 0), (5, 1), (5, 2)])
def test_segmented_schedules(w, h, queue, schedule, order):
    """The one-launch bounce and the sample passes keep 8 segments of whole
    4096-slot ranges, each claiming pixels from its own band of the frame and
    then from the other bands: a 40,960-pixel frame on 32,768 slots reuses
    slots across bands (bands of 5,120 pixels on segments of 4,096 slots);
    bands of a few pixels, or empty ones (6 pixels), leave most slots idle.
    Sample passes in longest-first order (tile_order 2) claim 64 indices per
    8x8 tile of the probe's order, those past the right or bottom edge (301x7,
    37x29, 3x2) being no pixel."""
    from zigtrace_amd import workloads
    wl = workloads.mesh_field(subdivisions=2, meshes=16, spp=3, width=w, height=h)
    s, o = wl.scene, wl.options
    mega = zt.render(s, o, zt.Tuning(method=MEGA, schedule=1))
    wave = zt.render(s, o, zt.Tuning(method=WAVE, schedule=schedule, queue_paths=queue, tile_order=order))
    assert np.array_equal(mega, wave)
    assert np.isfinite(wave).all()


def test_wavefront_no_nee_and_unbuilt_scene():
    s = zt.parse(scene_text("lamps"))
    o = zt.Options(32, 18, 4, 12, light_sampling=False)
    assert np.array_equal(zt.render(s, o, zt.Tuning(method=MEGA)), zt.render(s, o, zt.Tuning(method=WAVE)))
    u = zt.Scene()
    m = u.add_material(zt.lambertian((0.5, 0.5, 0.5)))
    u.add_sphere((0, 0, -3), 1, m)  # not built: brute-force path, no lights
    o = zt.Options(16, 16, 4, 8)
    assert np.array_equal(zt.render(u, o, zt.Tuning(method=MEGA)), zt.render(u, o, zt.Tuning(method=WAVE)))


@pytest.mark.parametrize("precision", [N.PRECISION_FP32, N.PRECISION_FP64])
def test_kernel_variants_and_stack_sizes_are_bit_identical(precision):
    """Occupancy variants and LDS stack capacities change only where traversal
    entries live (registers, LDS or private spill), never the visiting order."""
    from zigtrace_amd import workloads
    wl = workloads.mesh_field(subdivisions=2, meshes=16, spp=2, width=64, height=36)
    s, o = wl.scene, wl.options
    base = zt.render(s, o, zt.Tuning(method=MEGA, precision=precision))
    for variant in range(5):
        for cap in (8, 16, 32, 64):
            img = zt.render(s, o, zt.Tuning(method=MEGA, precision=precision, variant=variant, stack_entries=cap))
            assert np.array_equal(base, img), (variant, cap)
    for cap in (8, 24, 64):
        img = zt.render(s, o, zt.Tuning(method=WAVE, precision=precision, stack_entries=cap, queue_paths=700))
        assert np.array_equal(base, img), cap


def test_wave_scheduled_megakernel_is_bit_identical():
    """The wave-scheduled megakernel (lanes shade and refill while others traverse)
    renders exactly what the one-pixel-per-lane loop renders, for any refill
    threshold, tile shape (including ragged tiles and tiles larger than a wave)
    and build variant; also with light sampling off and max_depth 0."""
    from zigtrace_amd import workloads
    wl = workloads.mesh_field(subdivisions=2, meshes=16, spp=3, width=70, height=37)
    s, o = wl.scene, wl.options
    base = zt.render(s, o, zt.Tuning(method=MEGA, schedule=1))
    for refill in (1, 17, 32, 64):
        for tw, th in ((8, 8), (16, 16), (13, 5), (64, 64)):
            img = zt.render(s, o, zt.Tuning(method=MEGA, schedule=2, refill=refill, tile_w=tw, tile_h=th))
            assert np.array_equal(base, img), (refill, tw, th)
    for variant in (1, 2, 3, 4):
        img = zt.render(s, o, zt.Tuning(method=MEGA, schedule=2, variant=variant, stack_entries=8))
        assert np.array_equal(base, img), variant
    for opts in (zt.Options(33, 20, 2, 12, light_sampling=False, seed=4), zt.Options(9, 7, 2, 0)):
        a = zt.render(s, opts, zt.Tuning(method=MEGA, schedule=1))
        b = zt.render(s, opts, zt.Tuning(method=MEGA, schedule=2, refill=24))
        assert np.array_equal(a, b)
    with pytest.raises(zt.ZTError):
        zt.render(s, o, zt.Tuning(method=MEGA, schedule=2, refill=65))
    with pytest.raises(zt.ZTError):
        zt.render(s, o, zt.Tuning(method=MEGA, schedule=2, tile_w=2, tile_h=2))


@pytest.mark.parametrize("precision,bvh_width", [(N.PRECISION_FP64, 0), (N.PRECISION_FP32, 2)])
def test_wave_scheduled_two_child_walk_is_bit_identical(precision, bvh_width):
    """Over the two-child tree (the fp64 path, or fp32 without the four-wide
    tree) the wave-scheduled kernel walks each ray in the reference's order,
    one step per pass, so it renders the one-pixel-per-lane kernel's bits for
    any refill threshold, tile shape, build variant, stack capacity and tile
    order; also with light sampling off and max_depth 0."""
    from zigtrace_amd import workloads
    wl = workloads.mesh_field(subdivisions=2, meshes=16, spp=3, width=70, height=37)
    s, o = wl.scene, wl.options

    def render(opts, **kw):
        return zt.render(s, opts, zt.Tuning(method=MEGA, precision=precision, bvh_width=bvh_width, **kw))

    base = render(o, schedule=1)
    assert np.array_equal(base, render(o))  # the default form
    for refill in (1, 17, 64):
        for tw, th in ((8, 8), (13, 5), (64, 64)):
            img = render(o, schedule=2, refill=refill, tile_w=tw, tile_h=th)
            assert np.array_equal(base, img), (refill, tw, th)
    for variant in (1, 2, 3, 4):
        for cap in (8, 64):
            img = render(o, schedule=2, variant=variant, stack_entries=cap)
            assert np.array_equal(base, img), (variant, cap)
    for order in (1, 2):
        assert np.array_equal(base, render(o, schedule=2, tile_order=order)), order
    for opts in (zt.Options(33, 20, 2, 12, light_sampling=False, seed=4), zt.Options(9, 7, 2, 0)):
        assert np.array_equal(render(opts, schedule=1), render(opts, schedule=2))


def test_two_child_and_four_wide_traversal_agree():
    """fp32 renders walk the four-wide BVH by default; the two-child walk finds the
    same closest primitives (exact t ties aside), so the images agree; fp64 always
    walks the two-child nodes, so bvh_width cannot change it."""
    from zigtrace_amd import workloads
    wl = workloads.mesh_field(subdivisions=2, meshes=16, spp=4, width=96, height=54)
    s, o = wl.scene, wl.options
    four = zt.render(s, o, zt.Tuning(method=MEGA))
    two = zt.render(s, o, zt.Tuning(method=MEGA, bvh_width=2))
    small = zt.render(s, o, zt.Tuning(method=MEGA, node_bytes=64))
    for other in (two, small):
        differ = np.any(four != other, axis=2).mean()
        assert differ <= 1e-3, differ
    assert np.array_equal(four, zt.render(s, o, zt.Tuning(method=MEGA, bvh_width=4, node_bytes=128)))
    assert np.array_equal(small, zt.render(s, o, zt.Tuning(method=MEGA, schedule=1, node_bytes=64)))
    assert np.array_equal(two, zt.render(s, o, zt.Tuning(method=WAVE, bvh_width=2, queue_paths=1000)))
    f64 = [zt.render(s, o, zt.Tuning(method=MEGA, precision=N.PRECISION_FP64, bvh_width=w)) for w in (2, 4)]
    assert np.array_equal(f64[0], f64[1])


def test_rejects_bad_tuning():
    s = zt.parse(scene_text("cornell"))
    o = zt.Options(8, 8, 1, 4)
    for tu in (zt.Tuning(method=MEGA, stack_entries=7), zt.Tuning(method=MEGA, stack_entries=66),
               zt.Tuning(method=MEGA, variant=9), zt.Tuning(method=WAVE, stack_entries=5),
               zt.Tuning(method=MEGA, bvh_width=3), zt.Tuning(method=MEGA, node_bytes=32),
               zt.Tuning(method=WAVE, schedule=6), zt.Tuning(method=MEGA, schedule=3)):
        with pytest.raises(zt.ZTError) as e:
            zt.render(s, o, tu)
        assert e.value.name == "InvalidArgument"


def test_autotune_records_a_configuration_with_identical_output():
    from zigtrace_amd import workloads
    wl = workloads.mesh_field(subdivisions=1, meshes=9, spp=3, width=48, height=32)
    ctx = zt.Context(0)
    try:
        sid = ctx.upload(wl.scene)
        o = wl.options
        default = ctx.render(sid, o, zt.Tuning(method=MEGA))
        for method in (MEGA, WAVE, N.METHOD_AUTO):
            tu, ms = ctx.autotune(sid, o, method=method, probe_spp=1)
            assert ms > 0 and tu.method in (MEGA, WAVE)
            if method != N.METHOD_AUTO:
                assert tu.method == method
            assert np.array_equal(ctx.render(sid, o, tu), default), (method, tu)
        # AUTO (and no tuning at all) now renders with the record: same bits.
        assert np.array_equal(ctx.render(sid, o, zt.Tuning()), default)
        assert np.array_equal(ctx.render(sid, o, None), default)
        # fp64: the pixel-per-lane and wave-scheduled candidates render the same bits.
        f64 = N.PRECISION_FP64
        default64 = ctx.render(sid, o, zt.Tuning(method=MEGA, precision=f64))
        tu, ms = ctx.autotune(sid, o, method=MEGA, precision=f64, probe_spp=1)
        assert ms > 0 and tu.precision == f64
        assert np.array_equal(ctx.render(sid, o, tu), default64), tu
    finally:
        ctx.close()


def test_longest_first_tile_order_is_bit_identical():
    """Tiles taken most expensive first (after the one-sample cost probe) or in scan
    order give the same image as the one-pixel-per-lane kernel: only the assignment of
    tiles to waves changes. At 20 spp the default orders by cost."""
    from zigtrace_amd import workloads
    wl = workloads.mesh_field(subdivisions=2, meshes=16, spp=20, width=70, height=37)
    s, o = wl.scene, wl.options
    base = zt.render(s, o, zt.Tuning(method=MEGA, schedule=1))
    for order in (0, 1, 2):
        for tw, th in ((8, 8), (13, 5), (64, 64)):
            img = zt.render(s, o, zt.Tuning(method=MEGA, schedule=2, tile_order=order, tile_w=tw, tile_h=th))
            assert np.array_equal(base, img), (order, tw, th)
    small = zt.Options(9, 7, 3, 6, seed=2)
    a = zt.render(s, small, zt.Tuning(method=MEGA, schedule=1))
    assert np.array_equal(a, zt.render(s, small, zt.Tuning(method=MEGA, schedule=2, tile_order=2)))
    with pytest.raises(zt.ZTError):
        zt.render(s, o, zt.Tuning(method=MEGA, schedule=2, tile_order=3))


MIRROR_BOX = """
image 64 64
samples 2
depth 400
camera from 0 0 3 to 0 0 -1 fov 70
background color 0 0 0
material mirror metal 1 1 1 0
material lamp emissive 4 4 4
quad -5 -5 -5  10 0 0  0 10 0  mirror
quad -5 -5 5   10 0 0  0 10 0  mirror
quad -5 -5 -5  0 10 0  0 0 10  mirror
quad 5 -5 -5   0 10 0  0 0 10  mirror
quad -5 -5 -5  10 0 0  0 0 10  mirror
quad -5 5 -5   10 0 0  0 0 10  mirror
sphere 2 2 -3 0.7 lamp
"""


def test_tile_order_with_deep_paths_and_the_largest_tiles():
    """A closed box of perfect mirrors: no path escapes, roulette keeps 95% per
    bounce, so with depth 400 some one-sample probe paths start more than 100
    rays and their cubed counts reach the per-pixel cap; with 64x64 tiles a
    tile's sum of 4,096 capped terms must still fit its 32-bit counter. Forced
    longest-first order gives the same bits as scan order and the one-pixel-
    per-lane kernel, in both precisions."""
    s = zt.parse(MIRROR_BOX)
    o = zt.Options(64, 64, 2, 400, seed=5)
    for prec in (N.PRECISION_FP32, N.PRECISION_FP64):
        base = zt.render(s, o, zt.Tuning(method=MEGA, schedule=1, precision=prec))
        assert np.isfinite(base).all() and base.max() > 0
        for tw, th in ((8, 8), (64, 64)):
            for order in (1, 2):
                img = zt.render(s, o, zt.Tuning(method=MEGA, schedule=2, tile_order=order, tile_w=tw, tile_h=th,
                                                precision=prec))
                assert np.array_equal(base, img), (prec, tw, th, order)


def _with_large_leaves(nodes: np.ndarray, cap: int) -> np.ndarray:
    """The host BVH with every subtree of at most `cap` primitives turned into
    one leaf over its (contiguous) primitive range -- the shape of the leaves
    the reference's build emits at its depth cap (bvh.zig:150-152)."""
    out = nodes.copy()
    first = out["first"].astype(np.int64).copy()
    count = out["count"].astype(np.int64).copy()
    for i in range(len(out) - 1, -1, -1):  # children follow their parent
        if out["count"][i] == 0:
            left = int(out["first"][i])
            first[i] = min(first[left], first[left + 1])
            count[i] = count[left] + count[left + 1]
    for i in range(len(out)):
        if out["count"][i] == 0 and count[i] <= cap:
            out["first"][i] = first[i]
            out["count"][i] = count[i]
    return out


@pytest.mark.parametrize("cap", [9, 15, 40])
def test_large_leaves_trace_like_brute_force_and_render_identically(gpu_ctx, cap):
    """Leaves of up to 15 primitives (longer ones are split into synthetic
    nodes at upload) go through the wave kernel's parallel leaf pass: closest
    hits equal brute force and every schedule renders the same bits."""
    from zigtrace_amd import workloads
    from .helpers import aimed_rays
    wl = workloads.mesh_field(subdivisions=2, meshes=6, spp=2, width=40, height=24)
    arrays = wl.scene.view_arrays()
    nodes = _with_large_leaves(arrays["nodes"], cap)
    assert nodes["count"].max() > 4
    view = wl.scene.view()
    view.nodes = nodes.ctypes.data
    sid_c = C.c_uint32()
    N.check(N.lib().zt_gpu_scene_upload(gpu_ctx.handle, C.byref(view), C.byref(sid_c)), gpu_ctx.handle)
    sid = sid_c.value
    rng = np.random.default_rng(cap)
    rays = aimed_rays(rng, arrays["shapes"], 2000)
    brute = gpu_ctx.intersect(sid, rays, 1e-4, 1e9, mode=1)
    for mode in (0, 2):
        got = gpu_ctx.intersect(sid, rays, 1e-4, 1e9, mode=mode)
        assert np.array_equal(got["hit"], brute["hit"])
        h = got["hit"] == 1
        assert h.sum() > 100
        assert np.array_equal(got["t"][h], brute["t"][h]) and np.array_equal(got["material"][h], brute["material"][h])
    o = wl.options
    base = gpu_ctx.render(sid, o, zt.Tuning(method=MEGA, schedule=1))
    for refill in (1, 32):
        img = gpu_ctx.render(sid, o, zt.Tuning(method=MEGA, schedule=2, refill=refill))
        assert np.array_equal(base, img), refill
    gpu_ctx.release(sid)
