"""scenefile.parse (scenefile.zig:35-105) through the C++ reader."""
from __future__ import annotations

import ctypes as C

from . import _native as N
from .scene import Scene


class SceneSyntaxError(Exception):
    """error.SyntaxError with the reference's Diagnostic (line, message)."""

    def __init__(self, line: int, message: str):
        super().__init__(f"line {line}: {message}")
        self.line = line
        self.message = message


def parse(source: str | bytes) -> Scene:
    """Parses a scene file into a built Scene."""
    b = source.encode() if isinstance(source, str) else bytes(source)
    out = C.c_void_p()
    line = C.c_uint64()
    msg = C.create_string_buffer(256)
    rc = N.lib().zt_scene_parse(b, len(b), C.byref(out), C.byref(line), msg, 256)
    if rc == N.ERR_SYNTAX:
        raise SceneSyntaxError(line.value, msg.value.decode())
    N.check(rc)
    return Scene(out.value)


def load(path: str) -> Scene:
    with open(path, "rb") as f:
        return parse(f.read())
