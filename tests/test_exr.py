"""The minimal OpenEXR codec of the golden fixtures (cpipe_amd/exr.py), no GPU.

Pins: the byte layout of a hand-assembled 1x1 file against the OpenEXR file
layout (magic 20000630, version 2, attribute header with the eight required
attributes, an offset table, one `y, size, channel rows` block per scanline,
channel list sorted by name); round trips for every pixel type and channel
count; refusal of what the codec does not read; and the regenerated golden
pairs equal to what their generator computes.
"""
import os
import struct
import sys

import numpy as np
import pytest

from cpipe_amd import exr

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_layout_of_a_one_pixel_file():
    px = np.array([[[1.0, 2.0, 3.0, 0.5]]], np.float32)
    b = exr.encode(px, "half")
    assert struct.unpack_from("<ii", b, 0) == (20000630, 2)
    # first attribute: channels, chlist, sorted A, B, G, R, HALF, sampling 1
    assert b[8:24] == b"channels\0chlist\0"
    (size,) = struct.unpack_from("<i", b, 24)
    chl = b[28:28 + size]
    assert size == 4 * (2 + 16) + 1 and chl[-1] == 0
    for k, name in enumerate(b"ABGR"):
        e = chl[18 * k:18 * (k + 1)]
        assert e[0] == name and e[1] == 0
        assert struct.unpack("<iB3xii", e[2:]) == (1, 0, 1, 1)
    names = []
    pos = 8
    while b[pos] != 0:
        end = b.index(b"\0", pos)
        names.append(b[pos:end].decode())
        tend = b.index(b"\0", end + 1)
        (n,) = struct.unpack_from("<i", b, tend + 1)
        pos = tend + 5 + n
    assert names == ["channels", "compression", "dataWindow", "displayWindow", "lineOrder",
                     "pixelAspectRatio", "screenWindowCenter", "screenWindowWidth"]
    (off,) = struct.unpack_from("<Q", b, pos + 1)
    assert off == pos + 1 + 8 and len(b) == off + 8 + 4 * 2
    assert struct.unpack_from("<ii", b, off) == (0, 8)
    # channel rows in list order: A, B, G, R
    np.testing.assert_array_equal(np.frombuffer(b, "<f2", 4, off + 8), [0.5, 3.0, 2.0, 1.0])


@pytest.mark.parametrize("ptype,dtype", [("half", np.float16), ("float", np.float32), ("uint", np.uint32)])
@pytest.mark.parametrize("shape", [(1, 1, 4), (5, 7, 3), (3, 2), (13, 9, 4)])
def test_round_trip(ptype, dtype, shape):
    rng = np.random.default_rng(len(shape) * 7 + shape[0])
    if dtype == np.uint32:
        a = rng.integers(0, 2 ** 32, shape, dtype=np.uint64).astype(np.uint32)
    else:
        a = rng.uniform(-3, 3, shape).astype(dtype)
    got = exr.decode(exr.encode(a, ptype))
    assert got.dtype == np.dtype(dtype).newbyteorder("<")
    np.testing.assert_array_equal(got.reshape(a.shape), a)


def test_refusals():
    b = bytearray(exr.encode(np.zeros((2, 2, 4), np.float32)))
    with pytest.raises(ValueError):
        exr.decode(b"\0" * 16)
    tiled = bytearray(b)
    struct.pack_into("<i", tiled, 4, 2 | 0x200)
    with pytest.raises(ValueError, match="scanline"):
        exr.decode(bytes(tiled))
    i = bytes(b).index(b"compression\0compression\0") + len(b"compression\0compression\0") + 4
    zip_ = bytearray(b)
    zip_[i] = 3
    with pytest.raises(ValueError, match="compressed"):
        exr.decode(bytes(zip_))
    with pytest.raises(ValueError):
        exr.encode(np.zeros((2, 2, 2), np.float32))


def test_golden_pairs_match_their_generator():
    sys.path.insert(0, GOLDEN)
    try:
        import make_golden
    finally:
        sys.path.remove(GOLDEN)
    inp = make_golden.denoise_input()
    for name, out in (("denoise.bm3d", make_golden.bm3d_output(inp)),
                      ("denoise.wavelet_bayes_shrink", make_golden.wavelet_output(inp))):
        gi = exr.read(os.path.join(GOLDEN, name, "in.exr"))
        go = exr.read(os.path.join(GOLDEN, name, "out.exr"))
        assert gi.shape == (8, 8, 4) and gi.dtype == np.float32
        np.testing.assert_array_equal(gi, inp)
        np.testing.assert_array_equal(go, out)
