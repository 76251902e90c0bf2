"""The path-replay adjoint (zt_gpu_backward) against
  * the oracle's FP64 forward-mode (dual-number) gradient of the same
    estimator, same streams: the device fp64 path to 1e-9, the fp32 product
    path to 2e-3 (rounding and rare decision flips);
  * central finite differences of the device forward with common random
    numbers: exactly (1e-3) when no Russian roulette can occur (depth <= 3),
    and within 1% in gradient norm at depth 8 once the roulette decision
    flips that FD picks up are averaged out (256x256, 4096 spp, h = 1e-2);
  * the same at the step h = 1e-3: an fp32 forward cannot resolve a change of
    1e-3 (its accumulation noise divided by 2h swamps the difference), so the
    differences come from the fp64 forward, and the roulette flips, whose FD
    noise grows as 1/sqrt(paths * h), are averaged over 16384 spp on the
    emitter and one channel of each material;
and the Adam driver on the differentiable box."""
import numpy as np
import pytest

import zigtrace_amd as zt
from zigtrace_amd import _native as N
from zigtrace_amd import workloads as W
from zigtrace_amd.adjoint import DiffRenderer, inverse_render
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def box_params(dr):
    p = dr.get_params().astype(np.float64)
    # Per-channel-distinct values away from the shipped ones (no channel ties).
    rng = np.random.default_rng(4)
    return (p * rng.uniform(0.85, 1.15, p.size)).astype(np.float32)


def oracle_grad(w, dr, params, res, spp, depth, dimg, seed):
    a = w.scene.view_arrays()
    mats = a["materials"].copy()
    for m, off in enumerate(dr.offsets):
        if off >= 0:
            mats[m]["color"] = params[off:off + 3].astype(np.float64)
    a["materials"] = mats
    ref = O.OracleScene.from_view_arrays(a)
    return O.render_grad(ref, res, res, spp, depth, dr.offsets, dimg, seed=seed)


@pytest.mark.parametrize("depth", [3, 8])
def test_fp64_adjoint_equals_oracle_dual_numbers(gpu_ctx, depth):
    w = W.diff_box(width=48, height=48, spp=8, max_depth=depth)
    dr = DiffRenderer(gpu_ctx, w.scene, w.options, precision=N.PRECISION_FP64)
    p = box_params(dr)
    rng = np.random.default_rng(1)
    dimg = rng.uniform(-1, 1, (48, 48, 3)).astype(np.float32)
    img = dr.forward(p)
    g = dr.backward(p, dimg)
    ref_img, ref_g = oracle_grad(w, dr, p, 48, 8, depth, dimg.astype(np.float64), seed=0)
    dr.release()
    assert np.isclose(img, ref_img, rtol=1e-9, atol=1e-12).all(axis=2).mean() > 0.995
    assert np.allclose(g, ref_g, rtol=1e-7, atol=1e-9 * np.abs(ref_g).max())


def test_fp32_adjoint_close_to_oracle(gpu_ctx):
    w = W.diff_box(width=64, height=64, spp=16, max_depth=8)
    dr = DiffRenderer(gpu_ctx, w.scene, w.options)
    p = box_params(dr)
    dimg = np.ones((64, 64, 3), np.float32)
    g = dr.backward(p, dimg)
    _, ref_g = oracle_grad(w, dr, p, 64, 16, 8, dimg.astype(np.float64), seed=0)
    dr.release()
    assert np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g) < 2e-3


@pytest.mark.parametrize("precision", [N.PRECISION_FP32, N.PRECISION_FP64])
def test_backward_bits_identical_across_calls_and_tile_shapes(gpu_ctx, precision):
    """No atomics: per-lane columns, fixed 8x8 pixel groups and a fixed-order reduction.
    The 16x16 fp64 tile keeps its columns in global memory instead of LDS: same bits."""
    w = W.diff_box(width=72, height=56, spp=16, max_depth=8)  # groups straddle the tile grid's edge
    dr = DiffRenderer(gpu_ctx, w.scene, w.options, precision=precision)
    p = box_params(dr)
    rng = np.random.default_rng(12)
    dimg = rng.uniform(-1, 1, (56, 72, 3)).astype(np.float32)
    base = dr.backward(p, dimg)
    assert np.any(base != 0)
    for tile in (None, (8, 8), (16, 8), (8, 16), (16, 16), (32, 8)):
        g = dr.backward(p, dimg, tile=tile)
        assert np.array_equal(g.view(np.uint64), base.view(np.uint64)), tile
    dr.release()


def test_fp64_backward_with_global_columns_equals_oracle(gpu_ctx):
    """25 parameter materials (75 parameters): the fp64 columns do not fit in LDS and live
    in global memory; the gradient still equals the oracle's dual numbers."""
    rng = np.random.default_rng(3)
    s = zt.Scene()
    mats = [s.add_material(zt.lambertian(tuple(rng.uniform(0.2, 0.9, 3)))) for _ in range(24)]
    lamp = s.add_material(zt.emissive((6.0, 5.0, 4.0)))
    for k in range(24):
        s.add_sphere((rng.uniform(-3, 3), rng.uniform(0.3, 2), rng.uniform(-3, 0)), 0.35, mats[k])
    s.add_plane((0, 0, 0), (0, 1, 0), mats[0])
    s.add_sphere((0, 4, -1.5), 0.6, lamp)
    s.camera = zt.CameraConfig((0, 1.5, 6), (0, 1, -1.5), (0, 1, 0), 45.0)
    s.background = zt.solid((0.1, 0.1, 0.12))
    s.build()
    o = zt.Options(40, 32, 8, 6)
    dr = DiffRenderer(gpu_ctx, s, o, precision=N.PRECISION_FP64)
    assert dr.num_params == 75
    p = dr.get_params()
    dimg = rng.uniform(-1, 1, (32, 40, 3)).astype(np.float32)
    g = dr.backward(p, dimg)
    a = s.view_arrays()
    _, ref_g = O.render_grad(O.OracleScene.from_view_arrays(a), 40, 32, 8, 6, dr.offsets, dimg.astype(np.float64))
    dr.release()
    # fp64 on both sides; a path whose sine or cosine differs in the last ulp from the
    # oracle's may end elsewhere, and the random-sign dL/dimage makes some components sums
    # of cancelling terms: 1e-6 relative (fp32 agrees to 2e-3 at best).
    assert np.allclose(g, ref_g, rtol=1e-6, atol=1e-7 * np.abs(ref_g).max())


def fd_grad(dr, p, o, dimg, h):
    g = np.zeros(p.size)
    for i in range(p.size):
        up, dn = p.copy(), p.copy()
        up[i] += h
        dn[i] -= h
        g[i] = (np.sum(dimg * dr.forward(up, o).astype(np.float64)) -
                np.sum(dimg * dr.forward(dn, o).astype(np.float64))) / (2 * h)
    return g


def test_adjoint_equals_central_fd_without_roulette(gpu_ctx):
    w = W.diff_box(width=64, height=64, spp=16, max_depth=3)
    dr = DiffRenderer(gpu_ctx, w.scene, w.options)
    p = box_params(dr)
    dimg = np.ones((64, 64, 3))
    g = dr.backward(p, dimg.astype(np.float32))
    fd = fd_grad(dr, p, w.options, dimg, 1e-3 * np.maximum(p, 1.0).astype(np.float32).min())
    dr.release()
    assert np.linalg.norm(g - fd) / np.linalg.norm(fd) < 1e-3


def test_adjoint_within_1pct_of_central_fd_at_depth_8(gpu_ctx):
    w = W.diff_box(width=256, height=256, spp=4096, max_depth=8)
    dr = DiffRenderer(gpu_ctx, w.scene, w.options)
    p = box_params(dr)
    dimg = np.ones((256, 256, 3))
    g = dr.backward(p, dimg.astype(np.float32))
    fd = fd_grad(dr, p, w.options, dimg, 1e-2)
    dr.release()
    rel = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    assert rel < 0.01, rel


def test_adjoint_within_1pct_of_fp64_central_fd_at_h_1e3_depth_8(gpu_ctx):
    w = W.diff_box(width=256, height=256, spp=16384, max_depth=8)
    dr = DiffRenderer(gpu_ctx, w.scene, w.options)
    f64 = DiffRenderer(gpu_ctx, w.scene, w.options, precision=N.PRECISION_FP64)
    p = box_params(dr)
    dimg = np.ones((256, 256, 3))
    g = dr.backward(p, dimg.astype(np.float32))
    emitter = p.size - 1
    idx = [3 * m + m % 3 for m in range(emitter // 3)] + [emitter]
    fd = np.zeros(len(idx))
    for j, i in enumerate(idx):
        up, dn = p.copy(), p.copy()
        up[i] = np.float32(p[i] + 1e-3)
        dn[i] = np.float32(p[i] - 1e-3)
        step = float(up[i]) - float(dn[i])  # the parameters are fp32: divide by the step taken
        fd[j] = (np.sum(dimg * f64.forward(up)) - np.sum(dimg * f64.forward(dn))) / step
    f64.release()
    dr.release()
    rel = np.linalg.norm(g[idx] - fd) / np.linalg.norm(fd)
    assert rel < 0.01, rel
    # The emitter cannot flip a decision (light selection ignores Le): exact up to fp32.
    assert abs(g[emitter] - fd[-1]) < 2e-3 * abs(fd[-1])


def test_inverse_rendering_recovers_parameters(gpu_ctx):
    w = W.diff_box()  # 256x256, 32 spp, depth 8
    dr = DiffRenderer(gpu_ctx, w.scene, w.options)
    truth = dr.get_params().astype(np.float64)
    o = w.options
    target = dr.forward(truth.astype(np.float32), zt.Options(o.width, o.height, 1024, o.max_depth, seed=777))
    rng = np.random.default_rng(30)
    init = truth * rng.uniform(0.7, 1.3, truth.size)
    res = inverse_render(dr, target.astype(np.float64), init, steps=200, lr=1e-2)
    dr.release()
    assert res.losses[-1] < 0.2 * res.losses[0]
    err0 = np.abs(init - truth) / truth
    err1 = np.abs(res.params - truth) / truth
    assert np.median(err1) < 0.5 * np.median(err0)
    print(f"adam: {res.iters_per_s:.1f} fwd+bwd iters/s, loss {res.losses[0]:.4g} -> {res.losses[-1]:.4g}, "
          f"median rel param error {np.median(err0):.3f} -> {np.median(err1):.3f}")
