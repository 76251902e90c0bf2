"""The output stage (image.zig, png.zig) through the C ABI: tone mapping and
RGB8 encoding (`Output`, `to_rgb8`), PPM and PNG byte streams, and the same
encoding on the device (`Context.encode_rgb8`)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N

NONE, REINHARD, ACES = N.TONEMAP_NONE, N.TONEMAP_REINHARD, N.TONEMAP_ACES


@dataclass
class Output:
    """image.zig:24-38 `Output`: tonemap (NONE/REINHARD/ACES), gamma, exposure."""
    tonemap: int = NONE
    gamma: float = 2.2
    exposure: float = 1.0

    def record(self) -> np.ndarray:
        return N.record(N.OUTPUT, tonemap=self.tonemap, gamma=self.gamma, exposure=self.exposure)


def to_rgb8(image: np.ndarray, out: Output | None = None) -> np.ndarray:
    """Image.toRgb8 (image.zig:64-72): linear HDR (..., 3) -> uint8 of the same shape."""
    lin = np.ascontiguousarray(image, dtype=np.float64)
    rgb = np.zeros(lin.shape, dtype=np.uint8)
    N.check(N.lib().zt_encode_rgb8(N.ptr(lin), lin.size, N.ptr((out or Output()).record()), N.ptr(rgb)))
    return rgb


def _encode(fn, rgb: np.ndarray, width: int, height: int) -> bytes:
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    n = C.c_uint64()
    N.check(fn(width, height, N.ptr(rgb), None, 0, C.byref(n)))
    buf = np.zeros(max(n.value, 1), dtype=np.uint8)
    N.check(fn(width, height, N.ptr(rgb), N.ptr(buf), n.value, C.byref(n)))
    return buf[: n.value].tobytes()


def ppm_bytes(rgb: np.ndarray, width: int, height: int) -> bytes:
    """writePpm (image.zig:75-80) of packed RGB8."""
    return _encode(N.lib().zt_write_ppm, rgb, width, height)


def png_bytes(rgb: np.ndarray, width: int, height: int) -> bytes:
    """writePng (png.zig:10-57): 8-bit RGB, filter none, stored deflate blocks."""
    return _encode(N.lib().zt_write_png, rgb, width, height)
