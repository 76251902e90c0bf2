"""render.render (render.zig:152-180) on the GPU, plus the GPU context API.

`render(scene, Options)` is the drop-in: same options, same validation
(`InvalidOptions` before anything else), same f64 HDR image layout
(row-major y*width+x RGB), computed by the gfx950 kernels of
libzigtrace_amd.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .scene import Scene


@dataclass
class Options:
    """render.zig:12-23, field for field."""
    width: int
    height: int
    spp: int
    max_depth: int
    threads: int = 1
    seed: int = 0
    tile_size: int = 16
    light_sampling: bool = True

    def record(self) -> np.ndarray:
        return N.record(N.OPTIONS, width=self.width, height=self.height, spp=self.spp, max_depth=self.max_depth,
                        threads=self.threads, seed=self.seed, tile_size=self.tile_size,
                        light_sampling=int(bool(self.light_sampling)))


@dataclass
class Tuning:
    """GPU-only knobs (zt_gpu_tuning); none of them changes the image."""
    method: int = N.METHOD_AUTO
    precision: int = N.PRECISION_FP32
    tile_w: int = 0
    tile_h: int = 0
    block_size: int = 0
    queue_paths: int = 0
    variant: int = 0
    stack_entries: int = 0
    bvh_width: int = 0
    schedule: int = 0
    refill: int = 0
    node_bytes: int = 0
    tile_order: int = 0

    def record(self) -> np.ndarray:
        return N.record(N.TUNING, method=self.method, precision=self.precision, tile_w=self.tile_w,
                        tile_h=self.tile_h, block_size=self.block_size, queue_paths=self.queue_paths,
                        variant=self.variant, stack_entries=self.stack_entries, bvh_width=self.bvh_width,
                        schedule=self.schedule, refill=self.refill, node_bytes=self.node_bytes,
                        tile_order=self.tile_order)


def render(scene: Scene, opts: Options, tuning: Tuning | None = None) -> np.ndarray:
    """render.render: returns the HDR image as an (height, width, 3) f64 array."""
    o = opts.record()
    out = np.zeros((max(opts.height, 1), max(opts.width, 1), 3), dtype=np.float64)
    t = tuning.record() if tuning is not None else None
    rc = N.lib().zt_render(scene.handle, N.ptr(o), N.ptr(t), N.ptr(out))
    if rc != N.OK:
        ctx = C.c_void_p()
        msg = ""
        if N.lib().zt_default_context(C.byref(ctx)) == N.OK:
            msg = N.lib().zt_gpu_last_error(ctx).decode()
        raise N.ZTError(rc, msg)
    return out


class Context:
    """A zt_ctx: one HIP stream on one device, scenes uploaded to it."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        N.check(N.lib().zt_gpu_create(device, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().zt_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    def _check(self, rc):
        N.check(rc, self._h)

    def stream(self) -> int:
        return N.lib().zt_gpu_stream(self._h)

    def stats(self) -> dict:
        a = np.zeros(1, dtype=N.STATS)
        self._check(N.lib().zt_gpu_get_stats(self._h, N.ptr(a)))
        return {k: a[k][0].item() for k in N.STATS.names}

    def alloc(self, nbytes: int) -> int:
        """Device memory (an int address); free with `free`."""
        p = C.c_void_p()
        self._check(N.lib().zt_gpu_device_alloc(self._h, nbytes, C.byref(p)))
        return p.value

    def free(self, address: int) -> None:
        self._check(N.lib().zt_gpu_device_free(self._h, C.c_void_p(address)))

    def to_host(self, address: int, out: np.ndarray) -> np.ndarray:
        self._check(N.lib().zt_gpu_copy_to_host(self._h, N.ptr(out), C.c_void_p(address), out.nbytes))
        return out

    def to_device(self, address: int, data: np.ndarray) -> None:
        data = np.ascontiguousarray(data)
        self._check(N.lib().zt_gpu_copy_to_device(self._h, C.c_void_p(address), N.ptr(data), data.nbytes))

    def upload(self, scene: Scene) -> int:
        view = scene.view()
        sid = C.c_uint32()
        self._check(N.lib().zt_gpu_scene_upload(self._h, C.byref(view), C.byref(sid)))
        return sid.value

    def release(self, scene_id: int) -> None:
        self._check(N.lib().zt_gpu_scene_release(self._h, scene_id))

    def render(self, scene_id: int, opts: Options, tuning: Tuning | None = None, out=None) -> np.ndarray:
        """Renders into a host array ((h, w, 3) float32, or float64 for the fp64 path)."""
        t = tuning or Tuning()
        dt = np.float64 if t.precision == N.PRECISION_FP64 else np.float32
        if out is None:
            out = np.zeros((max(opts.height, 1), max(opts.width, 1), 3), dtype=dt)
        self._check(N.lib().zt_gpu_render(self._h, scene_id, N.ptr(opts.record()), N.ptr(t.record()), N.ptr(out), 0))
        return out

    def render_device(self, scene_id: int, opts: Options, tuning: Tuning | None, device_ptr: int) -> None:
        """Renders into device memory (no host copy): the timed form."""
        t = tuning or Tuning()
        self._check(N.lib().zt_gpu_render(self._h, scene_id, N.ptr(opts.record()), N.ptr(t.record()),
                                          C.c_void_p(device_ptr), 1))

    def autotune(self, scene_id: int, opts: Options, method: int = N.METHOD_AUTO,
                 precision: int = N.PRECISION_FP32, probe_spp: int = 0) -> tuple[Tuning, float]:
        """Times the candidate launch configurations on a reduced-spp probe of this frame and
        records the fastest in the context (used by later AUTO renders). Returns (tuning, ms)."""
        rec = np.zeros(1, dtype=N.TUNING)
        ms = C.c_double()
        self._check(N.lib().zt_gpu_autotune(self._h, scene_id, N.ptr(opts.record()), method, precision, probe_spp,
                                            N.ptr(rec), C.byref(ms)))
        return Tuning(**{k: int(rec[k][0]) for k in N.TUNING.names}), ms.value

    def encode_rgb8(self, hdr, out=None, count: int | None = None, precision: int = N.PRECISION_FP32,
                    rgb_device_ptr: int | None = None) -> np.ndarray | None:
        """Output.encode on the device. `hdr` is a host array or a device address (then give
        `count` and `precision`); returns host RGB8 unless `rgb_device_ptr` is given."""
        from .image import Output
        rec = (out or Output()).record()
        on_dev = not isinstance(hdr, np.ndarray)
        if on_dev:
            src, n = C.c_void_p(hdr), int(count)
        else:
            arr = np.ascontiguousarray(hdr)
            precision = N.PRECISION_FP64 if arr.dtype == np.float64 else N.PRECISION_FP32
            arr = arr.astype(np.float64 if precision == N.PRECISION_FP64 else np.float32, copy=False)
            src, n = N.ptr(arr), arr.size
        if rgb_device_ptr is not None:
            self._check(N.lib().zt_gpu_encode_rgb8(self._h, src, precision, n, N.ptr(rec), C.c_void_p(rgb_device_ptr),
                                                   int(on_dev), 1))
            return None
        rgb = np.zeros(n, dtype=np.uint8)
        self._check(N.lib().zt_gpu_encode_rgb8(self._h, src, precision, n, N.ptr(rec), N.ptr(rgb), int(on_dev), 0))
        return rgb.reshape(hdr.shape) if not on_dev else rgb

    def render_batch(self, scene_ids, options, tuning: Tuning | None = None, outs=None):
        """Renders several scenes in one wavefront pass; returns host images."""
        t = tuning or Tuning(method=N.METHOD_WAVEFRONT)
        dt = np.float64 if t.precision == N.PRECISION_FP64 else np.float32
        n = len(scene_ids)
        if outs is None:
            outs = [np.zeros((o.height, o.width, 3), dtype=dt) for o in options]
        ids = np.ascontiguousarray(scene_ids, dtype=np.uint32)
        recs = np.concatenate([o.record() for o in options])
        ptrs = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
        self._check(N.lib().zt_gpu_render_batch(self._h, n, N.ptr(ids), N.ptr(recs), N.ptr(t.record()), ptrs, 0))
        return outs

    def render_batch_device(self, scene_ids, options, tuning: Tuning | None, device_ptrs) -> None:
        t = tuning or Tuning(method=N.METHOD_WAVEFRONT)
        n = len(scene_ids)
        ids = np.ascontiguousarray(scene_ids, dtype=np.uint32)
        recs = np.concatenate([o.record() for o in options])
        ptrs = (C.c_void_p * n)(*device_ptrs)
        self._check(N.lib().zt_gpu_render_batch(self._h, n, N.ptr(ids), N.ptr(recs), N.ptr(t.record()), ptrs, 1))

    def trace_rays(self, scene_id: int, rays: np.ndarray, states: np.ndarray, max_depth: int,
                   light_sampling: bool = True, precision: int = N.PRECISION_FP32) -> np.ndarray:
        """traceWith per ray; `states` ((n, 4) uint64) is advanced in place."""
        rays = np.ascontiguousarray(rays, dtype=N.RAY)
        assert states.dtype == np.uint64 and states.shape == (len(rays), 4) and states.flags.c_contiguous
        out = np.zeros((len(rays), 3))
        self._check(N.lib().zt_gpu_trace_rays(self._h, scene_id, N.ptr(rays), len(rays), N.ptr(states), max_depth,
                                              int(light_sampling), precision, N.ptr(out)))
        return out

    def intersect(self, scene_id: int, rays: np.ndarray, t_min: float, t_max: float, mode: int = 0,
                  precision: int = N.PRECISION_FP32) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=N.RAY)
        out = np.zeros(len(rays), dtype=N.HIT)
        self._check(N.lib().zt_gpu_intersect(self._h, scene_id, N.ptr(rays), len(rays), t_min, t_max, mode, precision,
                                             N.ptr(out)))
        return out

    def rng(self, seed: int, streams, draws: int, kind: int = 0) -> np.ndarray:
        streams = np.ascontiguousarray(streams, dtype=np.uint64)
        out = np.zeros((len(streams), draws), dtype=np.uint64)
        self._check(N.lib().zt_gpu_rng(self._h, seed, N.ptr(streams), len(streams), draws, kind, N.ptr(out)))
        return out
