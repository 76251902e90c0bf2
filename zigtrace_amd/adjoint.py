"""Differentiable rendering through the C ABI's forward/backward pair.

`DiffRenderer` has the contract of an autograd Function: `forward(params)`
renders the image for the given material colours, `backward(params,
dL_dimage)` returns dL/dparams by re-tracing every path from its pixel's RNG
stream (nothing is stored between the calls but the parameters, options and
seed). Parameters: 3 floats per Lambertian/metal albedo and per emitter
radiance, in material order (`offsets` maps material -> first index).

`inverse_render` runs Adam on an L2 image loss.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .render import Context, Options, Tuning
from .scene import Scene

def _lib():
    return N.lib()


class DiffRenderer:
    def __init__(self, ctx: Context, scene: Scene, options: Options, precision: int = N.PRECISION_FP32):
        self.ctx = ctx
        self.sid = ctx.upload(scene)
        self.options = options
        self.tuning = Tuning(method=N.METHOD_MEGAKERNEL, precision=precision)
        n = C.c_uint32()
        self._check(_lib().zt_gpu_param_count(ctx.handle, self.sid, C.byref(n)))
        self.num_params = n.value
        nm = scene.info()["num_materials"]
        self.offsets = np.zeros(max(nm, 1), dtype=np.int32)
        self._check(_lib().zt_gpu_param_offsets(ctx.handle, self.sid, N.ptr(self.offsets)))
        self.offsets = self.offsets[:nm]

    def _check(self, rc):
        N.check(rc, self.ctx.handle)

    def get_params(self) -> np.ndarray:
        p = np.zeros(max(self.num_params, 1), dtype=np.float32)
        self._check(_lib().zt_gpu_get_params(self.ctx.handle, self.sid, N.ptr(p)))
        return p[: self.num_params]

    def forward(self, params: np.ndarray, options: Options | None = None) -> np.ndarray:
        o = options or self.options
        p = np.ascontiguousarray(params, dtype=np.float32)
        dt = np.float64 if self.tuning.precision == N.PRECISION_FP64 else np.float32
        out = np.zeros((o.height, o.width, 3), dtype=dt)
        self._check(_lib().zt_gpu_forward(self.ctx.handle, self.sid, N.ptr(o.record()), N.ptr(self.tuning.record()),
                                          N.ptr(p), N.ptr(out), 0))
        return out

    def backward(self, params: np.ndarray, dl_dimage: np.ndarray, options: Options | None = None,
                 tile: tuple[int, int] | None = None) -> np.ndarray:
        """dLoss/dparams; `tile` (multiples of 8) sets the backward's workgroup tile, which
        does not change the result's bits."""
        o = options or self.options
        p = np.ascontiguousarray(params, dtype=np.float32)
        g = np.ascontiguousarray(dl_dimage, dtype=np.float32)
        assert g.size == o.width * o.height * 3
        out = np.zeros(max(self.num_params, 1), dtype=np.float64)
        t = self.tuning
        if tile is not None:
            t = Tuning(**{**t.__dict__, "tile_w": tile[0], "tile_h": tile[1]})
        self._check(_lib().zt_gpu_backward(self.ctx.handle, self.sid, N.ptr(o.record()), N.ptr(t.record()),
                                           N.ptr(p), N.ptr(g), 0, N.ptr(out)))
        return out[: self.num_params]

    def release(self):
        self.ctx.release(self.sid)


@dataclass
class AdamResult:
    params: np.ndarray
    losses: list = field(default_factory=list)
    seconds: float = 0.0
    iters_per_s: float = 0.0


def inverse_render(dr: DiffRenderer, target: np.ndarray, init: np.ndarray, steps: int = 200, lr: float = 1e-2,
                   beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, seed_stride: int = 1,
                   lower: float = 0.0, upper: float | None = None) -> AdamResult:
    """Adam on L = mean((render(params) - target)^2); one forward + one backward per step.
    Each step uses a fresh seed (seed + step*seed_stride) so the noise averages out.
    Adam runs on x = params / init (each parameter relative to its starting value), so one
    learning rate serves albedos (~0.5) and emitter radiances (~15) alike."""
    init = np.array(init, dtype=np.float64)
    scale = np.where(init > 0, init, 1.0)
    x = init / scale
    theta = init.copy()
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    res = AdamResult(params=theta)
    o0 = dr.options
    npx = target.size
    t0 = time.perf_counter()
    for it in range(steps):
        o = Options(o0.width, o0.height, o0.spp, o0.max_depth, o0.threads, o0.seed + it * seed_stride, o0.tile_size,
                    o0.light_sampling)
        p32 = theta.astype(np.float32)
        img = dr.forward(p32, o)
        diff = img.astype(np.float64) - target
        res.losses.append(float(np.mean(diff * diff)))
        g = dr.backward(p32, (2.0 / npx) * diff, o) * scale
        m = beta1 * m + (1 - beta1) * g
        v = beta2 * v + (1 - beta2) * g * g
        mh = m / (1 - beta1 ** (it + 1))
        vh = v / (1 - beta2 ** (it + 1))
        x = x - lr * mh / (np.sqrt(vh) + eps)
        theta = np.maximum(x * scale, lower)
        if upper is not None:
            theta = np.minimum(theta, upper)
        x = theta / scale
    res.seconds = time.perf_counter() - t0
    res.iters_per_s = steps / res.seconds if res.seconds > 0 else float("inf")
    res.params = theta
    return res
