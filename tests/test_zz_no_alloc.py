"""zt_gpu_backward_noalloc (through the C ABI) against zt_gpu_backward: the same
gradient bits, on two calls in a row with the second one smaller (the shared
scratch keeps the first call's rows; nothing may leak into the second), in both
precisions; sizes beyond its scratch are rejected before any work is launched."""
import ctypes as C

import numpy as np
import pytest

import zigtrace_amd as zt
from zigtrace_amd import _native as N
from zigtrace_amd import workloads as W
from zigtrace_amd.adjoint import DiffRenderer

pytestmark = pytest.mark.gpu


def noalloc_backward(ctx, sid, opts, tuning, dimg, nparams):
    """dL/dparams through zt_gpu_backward_noalloc, with device buffers the caller owns."""
    d_img = ctx.alloc(dimg.nbytes)
    d_grad = ctx.alloc(max(nparams, 1) * 8)
    try:
        ctx.to_device(d_img, dimg)
        rc = N.lib().zt_gpu_backward_noalloc(ctx.handle, sid, N.ptr(opts.record()), N.ptr(tuning.record()),
                                             C.c_void_p(d_img), C.c_void_p(d_grad))
        N.check(rc, ctx.handle)
        return ctx.to_host(d_grad, np.zeros(max(nparams, 1)))[:nparams]
    finally:
        ctx.free(d_img)
        ctx.free(d_grad)


@pytest.mark.parametrize("precision", [N.PRECISION_FP32, N.PRECISION_FP64])
def test_noalloc_backward_equals_backward_on_two_calls(precision):
    ctx = zt.Context(0)
    w = W.diff_box(width=96, height=80, spp=8, max_depth=8)
    dr = DiffRenderer(ctx, w.scene, w.options, precision=precision)
    rng = np.random.default_rng(21)
    p = (dr.get_params() * rng.uniform(0.9, 1.1, dr.num_params)).astype(np.float32)
    tu = zt.Tuning(precision=precision)
    for (width, height, spp, seed) in ((96, 80, 8, 3), (40, 24, 4, 9)):  # the second call is smaller
        o = zt.Options(width, height, spp, 8, seed=seed)
        dimg = rng.uniform(-1, 1, (height, width, 3)).astype(np.float32)
        want = dr.backward(p, dimg, o)  # also sets p as the scene's parameters
        got = noalloc_backward(ctx, dr.sid, o, tu, dimg, dr.num_params)
        assert np.any(want != 0)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (width, height)
    dr.release()
    ctx.close()


def test_noalloc_backward_rejects_sizes_beyond_its_scratch():
    ctx = zt.Context(0)
    w = W.diff_box(width=16, height=16, spp=1, max_depth=4)
    dr = DiffRenderer(ctx, w.scene, w.options)
    # 27 parameters x (4096/8)^2 groups = 7.1M partial sums > 1048576: rejected on the host.
    o = zt.Options(4096, 4096, 1, 4)
    d = ctx.alloc(64)
    rc = N.lib().zt_gpu_backward_noalloc(ctx.handle, dr.sid, N.ptr(o.record()), N.ptr(zt.Tuning().record()),
                                         C.c_void_p(d), C.c_void_p(d))
    assert rc == N.ERR_INVALID_ARGUMENT
    assert "8 MiB" in N.lib().zt_gpu_last_error(ctx.handle).decode()
    ctx.free(d)
    dr.release()
    ctx.close()
