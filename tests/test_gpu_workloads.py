"""fp32 product path vs the oracle on the benchmark workloads.

Each workload is compared two ways, the reference's per-pixel estimator
(render.zig:121-147) being the quantity under test:

* seed means: per-block means over 16 independent seeds on each side agree
  within 4 standard errors (cells beyond 4 SE may not exceed what the t
  distribution of the statistic predicts, plus slack), and so does the
  whole-image mean;
* common random numbers: with the same seed the per-pixel streams and draw
  order are the reference's (render.zig:133-144), so at 4096 samples per
  pixel the fp32 image's relative RMSE against the oracle's f64 image stays
  below 1e-3. Stream ids are y*W + x, so the GPU renders the whole frame and
  the oracle the pixels (y*W + x) % stride == 0 of the same frame. What is
  left is decision flips (a ray that grazes a silhouette, a roulette draw
  on a threshold), whose effect depends on the scene: the oracle's rounding
  perturbation measures it, and where rounding alone moves the oracle's own
  image by more than 1e-3 the bound is 3x that floor instead.

Workloads: the 1.3M-triangle mesh field (1920 wide), the sphere field over a
ground plane (600x340), the shipped spheres.txt (r = 1000 ground sphere),
scenes of the batch workload rendered through zt_gpu_render_batch, and the
wavefront renderer on its own; plus fp64 wavefront and batch renders
against the oracle pixel for pixel.
"""
import numpy as np
import pytest

import zigtrace_amd as zt
from zigtrace_amd import _native as N
from zigtrace_amd import workloads as W
from oracle import oracle as O

from .conftest import scene_text
from .helpers import image_mean_within, oracle_of, rel_rmse, seed_mean_test, strided

pytestmark = pytest.mark.gpu

SEEDS = 16
RMSE_BAR = 1e-3  # relative RMSE of converged 4096-spp images with common random numbers


def assert_seed_means(gpu, cpu, block, what):
    beyond, cells, expected = seed_mean_test(gpu, cpu, block)
    allowed = expected + 4 * np.sqrt(expected) + 3
    assert beyond <= allowed, f"{what}: {beyond} of {cells} cells beyond 4 SE (t-expected {expected:.1f})"
    ok, z = image_mean_within(gpu, cpu)
    assert ok, f"{what}: image mean off by {z} SE"


@pytest.fixture(scope="module")
def mesh_field():
    wl = W.mesh_field()
    return wl, oracle_of(wl.scene)


def test_mesh_field_seed_means(gpu_ctx, mesh_field):
    wl, ref = mesh_field
    sid = gpu_ctx.upload(wl.scene)
    w, h, spp, depth = 384, 216, 16, 16
    gpu = [gpu_ctx.render(sid, zt.Options(w, h, spp, depth, seed=s)).astype(np.float64) for s in range(SEEDS)]
    gpu_ctx.release(sid)
    cpu = [ref.render(w, h, spp, depth, seed=500 + s)[0] for s in range(SEEDS)]
    assert_seed_means(gpu, cpu, 8, "mesh field")


def test_mesh_field_full_size_wavefront_equals_megakernel(gpu_ctx, mesh_field):
    """The benchmark frame (1920x1080, 1.3M triangles, depth 16) at 16 spp
    through the wavefront's default form -- the one-sample cost probe and the
    longest-first pixel order, a pool of every pixel, 16 sample passes -- and
    through the default wave-scheduled megakernel: the same bits, and the
    same extension and shadow ray counts."""
    wl, _ = mesh_field
    o = wl.options
    o16 = zt.Options(o.width, o.height, 16, o.max_depth, seed=11)
    sid = gpu_ctx.upload(wl.scene)
    mega = gpu_ctx.render(sid, o16, zt.Tuning(method=N.METHOD_MEGAKERNEL))
    mega_rays = {k: gpu_ctx.stats()[k] for k in ("extend_rays", "shadow_rays")}
    wave = gpu_ctx.render(sid, o16, zt.Tuning(method=N.METHOD_WAVEFRONT))
    wave_rays = {k: gpu_ctx.stats()[k] for k in ("extend_rays", "shadow_rays")}
    gpu_ctx.release(sid)
    assert np.array_equal(mega, wave)
    assert mega_rays == wave_rays and mega_rays["extend_rays"] > o.width * o.height * 16


def test_mesh_field_full_width_crn_rmse(gpu_ctx, mesh_field):
    wl, ref = mesh_field
    o = wl.options  # 1920x1080, depth 16, NEE + MIS
    spp, stride = 4096, 4099
    sid = gpu_ctx.upload(wl.scene)
    got = gpu_ctx.render(sid, zt.Options(o.width, o.height, spp, o.max_depth, seed=3)).astype(np.float64)
    gpu_ctx.release(sid)
    want, _ = ref.render(o.width, o.height, spp, o.max_depth, seed=3, pixel_stride=stride)
    g, c = strided(got, stride), strided(want, stride)
    assert len(g) > 400
    err = rel_rmse(g, c)
    assert err < RMSE_BAR, f"mesh field: relative RMSE {err:.2e} over {len(g)} pixels"


@pytest.fixture(scope="module")
def sphere_field():
    wl = W.sphere_field()
    return wl, oracle_of(wl.scene)


def test_sphere_field_seed_means(gpu_ctx, sphere_field):
    wl, ref = sphere_field
    o = wl.options  # 600x340, 16 spp, depth 16
    sid = gpu_ctx.upload(wl.scene)
    gpu = [gpu_ctx.render(sid, zt.Options(o.width, o.height, o.spp, o.max_depth, seed=s)).astype(np.float64)
           for s in range(SEEDS)]
    gpu_ctx.release(sid)
    cpu = [ref.render(o.width, o.height, o.spp, o.max_depth, seed=900 + s)[0] for s in range(SEEDS)]
    assert_seed_means(gpu, cpu, 10, "sphere field")


@pytest.mark.parametrize("method", [N.METHOD_MEGAKERNEL, N.METHOD_WAVEFRONT])
def test_sphere_field_crn_rmse(gpu_ctx, sphere_field, method):
    wl, ref = sphere_field
    o = wl.options
    spp, stride = 4096, 97
    sid = gpu_ctx.upload(wl.scene)
    got = gpu_ctx.render(sid, zt.Options(o.width, o.height, spp, o.max_depth, seed=5),
                         zt.Tuning(method=method)).astype(np.float64)
    gpu_ctx.release(sid)
    want, _ = ref.render(o.width, o.height, spp, o.max_depth, seed=5, pixel_stride=stride)
    err = rel_rmse(strided(got, stride), strided(want, stride))
    assert err < RMSE_BAR, f"sphere field ({method}): relative RMSE {err:.2e}"


def test_shipped_spheres_file_fp32(gpu_ctx):
    """examples/spheres.txt: its ground is a sphere of radius 1000, a hard case for fp32."""
    text = scene_text("spheres")
    sid = gpu_ctx.upload(zt.parse(text))
    ref = O.OracleScene.parse(text)
    w, h, spp, depth = 300, 170, 16, 16
    gpu = [gpu_ctx.render(sid, zt.Options(w, h, spp, depth, seed=s)).astype(np.float64) for s in range(SEEDS)]
    cpu = [ref.render(w, h, spp, depth, seed=300 + s)[0] for s in range(SEEDS)]
    assert_seed_means(gpu, cpu, 10, "spheres.txt")
    stride = 53
    got = gpu_ctx.render(sid, zt.Options(w, h, 4096, depth, seed=8)).astype(np.float64)
    want, _ = ref.render(w, h, 4096, depth, seed=8, pixel_stride=stride)
    gpu_ctx.release(sid)
    err = rel_rmse(strided(got, stride), strided(want, stride))
    assert err < RMSE_BAR, f"spheres.txt: relative RMSE {err:.2e}"


@pytest.fixture(scope="module")
def batch4():
    """The first four scenes of the 256-scene batch workload (same generator and seed)."""
    ws = W.batch_scenes(count=4)
    return ws, [oracle_of(w.scene) for w in ws]


BATCH_SIDE, BATCH_CRN_SPP = 64, 4096
FLOOR_REL, FLOOR_FACTOR = 2.0 ** -22, 3.0


@pytest.fixture(scope="module")
def batch4_crn(batch4):
    """Per scene: the oracle's 4096-spp image (seed 40 + i) and its rounding-noise floor, the
    relative RMSE that moving every hit point (within its surface) and every scattered
    direction by a relative 2^-22 -- two fp32 ulps -- causes in the oracle itself."""
    ws, refs = batch4
    out = []
    for i, (w, ref) in enumerate(zip(ws, refs)):
        args = (BATCH_SIDE, BATCH_SIDE, BATCH_CRN_SPP, w.options.max_depth)
        want, _ = ref.render(*args, seed=40 + i)
        with O.rounding_perturbation(FLOOR_REL):
            moved, _ = ref.render(*args, seed=40 + i)
        out.append((want, rel_rmse(moved, want)))
    return out


@pytest.mark.parametrize("method", [N.METHOD_AUTO, N.METHOD_WAVEFRONT])
def test_batch_scenes_crn_rmse_through_render_batch(gpu_ctx, batch4, batch4_crn, method):
    """Scenes of uneven size in one batch launch. The 1e-3 bar holds where rounding allows it:
    some of these scenes (thousands of overlapping small spheres, glass, depth 27) are so
    sensitive that two-ulp rounding moves the oracle's own converged image by more than 1e-3;
    there the fp32 image must stay within 3x that floor."""
    ws, _ = batch4
    ids = [gpu_ctx.upload(w.scene) for w in ws]
    opts = [zt.Options(BATCH_SIDE, BATCH_SIDE, BATCH_CRN_SPP, w.options.max_depth, seed=40 + i)
            for i, w in enumerate(ws)]
    outs = gpu_ctx.render_batch(ids, opts, zt.Tuning(method=method))
    for sid in ids:
        gpu_ctx.release(sid)
    for w, (want, floor), img in zip(ws, batch4_crn, outs):
        err = rel_rmse(img.astype(np.float64), want)
        assert err < max(RMSE_BAR, FLOOR_FACTOR * floor), \
            f"{w.description}: relative RMSE {err:.2e} (rounding floor {floor:.2e})"


def test_batch_scenes_seed_means_through_render_batch(gpu_ctx, batch4):
    ws, refs = batch4
    ids = [gpu_ctx.upload(w.scene) for w in ws]
    side = 64
    gpu = [[] for _ in ws]
    for s in range(SEEDS):
        opts = [zt.Options(side, side, w.options.spp, w.options.max_depth, seed=s) for w in ws]
        for i, img in enumerate(gpu_ctx.render_batch(ids, opts, zt.Tuning())):
            gpu[i].append(img.astype(np.float64))
    for sid in ids:
        gpu_ctx.release(sid)
    for i, (w, ref) in enumerate(zip(ws, refs)):
        cpu = [ref.render(side, side, w.options.spp, w.options.max_depth, seed=700 + s)[0] for s in range(SEEDS)]
        assert_seed_means(gpu[i], cpu, 8, w.description)


FP64_CASES = [("cornell", (48, 48), 8, 16, 0), ("spheres", (60, 34), 4, 16, 1), ("glass", (48, 27), 6, 16, 2),
              ("lamps", (48, 27), 6, 12, 3)]


def fp64_close(got, want):
    return np.isclose(got, want, rtol=1e-9, atol=1e-12).all(axis=2).mean()


@pytest.mark.parametrize("name,size,spp,depth,seed", FP64_CASES)
def test_fp64_wavefront_equals_oracle(gpu_ctx, name, size, spp, depth, seed):
    text = scene_text(name)
    sid = gpu_ctx.upload(zt.parse(text))
    w, h = size
    got = gpu_ctx.render(sid, zt.Options(w, h, spp, depth, seed=seed),
                         zt.Tuning(method=N.METHOD_WAVEFRONT, precision=N.PRECISION_FP64, queue_paths=512))
    gpu_ctx.release(sid)
    want, _ = O.OracleScene.parse(text).render(w, h, spp, depth, seed=seed)
    assert fp64_close(got, want) >= 0.995, name


def test_fp64_batch_equals_oracle(gpu_ctx):
    """An fp64 batch runs through the wavefront renderer's shared path pool."""
    texts = [scene_text(n) for n, *_ in FP64_CASES]
    ids = [gpu_ctx.upload(zt.parse(t)) for t in texts]
    opts = [zt.Options(w, h, spp, depth, seed=seed) for _, (w, h), spp, depth, seed in FP64_CASES]
    outs = gpu_ctx.render_batch(ids, opts, zt.Tuning(precision=N.PRECISION_FP64, queue_paths=2048))
    for sid in ids:
        gpu_ctx.release(sid)
    for t, o, img in zip(texts, opts, outs):
        want, _ = O.OracleScene.parse(t).render(o.width, o.height, o.spp, o.max_depth, seed=o.seed)
        assert fp64_close(img, want) >= 0.995
