"""DNG NoiseProfile ingest (include/cpipe_dng.h) for Python hosts and tests.

noise_profile(path_or_bytes) returns the file's NoiseProfile as (beta, lambda)
pairs -- variance beta + lambda * x per colour plane, the order
CalibrationBlock::noise_profile holds them in -- or [] when the file has none;
host.Buffer.set_noise_profile(pairs) then hands them to the nodes as a cpipe
host's DngReader would (src/cpipe/ingest/dng/DngReader.cpp).
"""
from __future__ import annotations

import ctypes

from . import engine


class DngError(RuntimeError):
    def __init__(self, status: int):
        super().__init__(f"cpipe_dng_noise_profile status {status}: not a TIFF/DNG or a malformed NoiseProfile")
        self.status = status


def noise_profile(src, max_pairs: int = 16):
    """src: a path or the file's bytes. Returns [(beta, lambda), ...]."""
    data = src if isinstance(src, (bytes, bytearray)) else open(src, "rb").read()
    buf = (ctypes.c_char * max(1, len(data))).from_buffer_copy(bytes(data) or b"\0")
    beta = (ctypes.c_float * max_pairs)()
    lam = (ctypes.c_float * max_pairs)()
    n = ctypes.c_int()
    rc = engine.lib().cpipe_dng_noise_profile(buf, len(data), beta, lam, max_pairs, ctypes.byref(n))
    if rc != 0:
        raise DngError(rc)
    return [(float(beta[i]), float(lam[i])) for i in range(min(n.value, max_pairs))]
