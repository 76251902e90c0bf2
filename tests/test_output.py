"""The output stage (image.zig, png.zig): the reference's own known answers,
PNG decoded by Python's zlib (an independent inflate and CRC-32), and the
device encoder equal to the host encoder."""
import struct
import zlib

import numpy as np
import pytest

import zigtrace_amd as zt
from zigtrace_amd import image as I


def test_tonemap_curves():  # image.zig:87-101
    lin = I.Output(I.NONE, gamma=1.0)
    assert I.to_rgb8(np.array([-3.0]), lin)[0] == 0
    # reinhard(1) = 0.5 -> 128 at gamma 1 (round half away from zero of 127.5)
    assert I.to_rgb8(np.array([1.0]), I.Output(I.REINHARD, gamma=1.0))[0] == 128
    assert I.to_rgb8(np.array([0.0]), I.Output(I.ACES, gamma=1.0))[0] == 0
    assert I.to_rgb8(np.array([100.0]), I.Output(I.ACES, gamma=1.0))[0] >= 253  # aces(100) > 0.99
    for tm in (I.NONE, I.REINHARD, I.ACES):
        x = np.arange(0, 4, 0.05)
        y = I.to_rgb8(x, I.Output(tm, gamma=1.0)).astype(int)
        assert np.all(np.diff(y) >= 0)


def test_encode_applies_gamma_and_clamps():  # image.zig:103-112
    lin = I.Output(gamma=1.0)
    got = I.to_rgb8(np.array([-1.0, 7.0, 0.5, np.nan]), lin)
    assert got.tolist() == [0, 255, 128, 0]
    assert I.to_rgb8(np.array([0.5]), I.Output(gamma=2.2))[0] == 186  # 0.5^(1/2.2) ~= 0.7297
    assert I.to_rgb8(np.array([0.25]), I.Output(gamma=1.0, exposure=2.0))[0] == 128


def test_ppm_output():  # image.zig:114-126
    img = np.array([[[1.0, 0, 0], [0, 0, 1.0]]])
    rgb = I.to_rgb8(img, I.Output())
    assert I.ppm_bytes(rgb, 2, 1) == b"P6\n2 1\n255\n\xff\x00\x00\x00\x00\xff"


def _chunks(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, out = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body)
        out.append((kind, body))
        pos += 12 + n
    assert pos == len(data)
    return out


@pytest.mark.parametrize("w,h", [(3, 2), (200, 150), (21845, 1), (1, 1)])
def test_png_round_trips_through_zlib(w, h):  # png.zig:136-180
    i = np.arange(w * h * 3, dtype=np.uint64)
    rgb = ((i * 31 + (i >> 7)) & 0xFF).astype(np.uint8)
    data = I.png_bytes(rgb, w, h)
    ch = _chunks(data)
    assert [k for k, _ in ch] == [b"IHDR", b"IDAT", b"IEND"]
    assert struct.unpack(">II", ch[0][1][:8]) == (w, h) and ch[0][1][8:] == bytes([8, 2, 0, 0, 0])
    raw = zlib.decompress(ch[1][1])
    assert len(raw) == (w * 3 + 1) * h
    lines = np.frombuffer(raw, np.uint8).reshape(h, w * 3 + 1)
    assert not lines[:, 0].any()
    assert np.array_equal(lines[:, 1:].reshape(-1), rgb)
    blocks = -(-((w * 3 + 1) * h) // 65535)
    assert len(ch[1][1]) == 2 + (w * 3 + 1) * h + 5 * blocks + 4  # stored blocks, as png.zig:29-31


def test_png_rejects_bad_dimensions():  # png.zig:12-13
    with pytest.raises(zt.ZTError) as e:
        I.png_bytes(np.zeros(0, np.uint8), 0, 5)
    assert e.value.name == "InvalidDimensions"


@pytest.mark.gpu
def test_device_encoder_equals_host_encoder():
    rng = np.random.default_rng(3)
    hdr = np.concatenate([rng.exponential(0.7, 300000), [0, -1, np.inf, np.nan, 1e30, 0.5]]).astype(np.float32)
    ctx = zt.Context(0)
    try:
        for out in (I.Output(I.ACES), I.Output(I.NONE, gamma=1.0), I.Output(I.REINHARD, 2.4, 1.7)):
            host = I.to_rgb8(hdr.astype(np.float64), out)
            dev = ctx.encode_rgb8(hdr, out)
            diff = np.abs(host.astype(int) - dev.astype(int))
            assert diff.max() <= 1 and (diff > 0).mean() < 1e-5, (out, diff.max(), (diff > 0).mean())
        # A framebuffer that never leaves the device: render, then encode in place.
        from .conftest import scene_text
        sid = ctx.upload(zt.parse(scene_text("cornell")))
        o = zt.Options(40, 30, 4, 8)
        d = ctx.alloc(o.width * o.height * 12)
        try:
            ctx.render_device(sid, o, zt.Tuning(), d)
            img = ctx.to_host(d, np.zeros((o.height, o.width, 3), np.float32))
            dev = ctx.encode_rgb8(d, I.Output(I.ACES), count=img.size).reshape(img.shape)
            host = I.to_rgb8(img.astype(np.float64), I.Output(I.ACES))
            assert np.abs(host.astype(int) - dev.astype(int)).max() <= 1
        finally:
            ctx.free(d)
    finally:
        ctx.close()
