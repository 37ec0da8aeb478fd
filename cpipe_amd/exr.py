"""Minimal OpenEXR reader/writer for the golden fixtures (in.exr / out.exr).

The reference's golden gate reads each node's fixture pair with OpenImageIO
(tests/unit/test_golden_isp_nodes.cpp:86-108: FLOAT pixels, 1 or 4 channels)
and its generator writes them (tests/fixtures/gen_golden_isp_nodes.cpp). The
pairs in the reference tree are Git LFS pointer stubs, and no EXR library is
installed here, so the regenerated fixtures are written and read by this
module: single-part scanline files, NO_COMPRESSION, HALF / FLOAT / UINT
channels, INCREASING_Y line order -- the subset of the OpenEXR file layout
(magic 20000630, version 2, attribute header, one offset per scanline block,
blocks of `y, byte count, channel rows in channel-list order`) that such
fixtures need. Compressed, tiled, deep and multi-part files are refused.

Channel order: arrays are (H, W, C) in R, G, B, A order (or one channel "Y");
the file's channel list is sorted by name (A, B, G, R) as the format requires.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = 20000630
_TYPES = {0: np.dtype("<u4"), 1: np.dtype("<f2"), 2: np.dtype("<f4")}
_CODES = {np.dtype("<u4"): 0, np.dtype("<f2"): 1, np.dtype("<f4"): 2}
_NAMES = {1: ["Y"], 3: ["R", "G", "B"], 4: ["R", "G", "B", "A"]}


def _attr(name: str, kind: str, payload: bytes) -> bytes:
    return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(payload)) + payload


def encode(pixels: np.ndarray, pixel_type: str = "float") -> bytes:
    """EXR bytes of an (H, W, C) or (H, W) array; pixel_type 'half', 'float' or 'uint'."""
    a = np.asarray(pixels)
    if a.ndim == 2:
        a = a[..., None]
    h, w, c = a.shape
    if c not in _NAMES or h <= 0 or w <= 0:
        raise ValueError(f"unsupported image shape {a.shape}")
    dt = {"half": np.dtype("<f2"), "float": np.dtype("<f4"), "uint": np.dtype("<u4")}[pixel_type]
    names = _NAMES[c]
    order = sorted(range(c), key=lambda i: names[i])
    chlist = b"".join(names[i].encode() + b"\0" + struct.pack("<iB3xii", _CODES[dt], 0, 1, 1) for i in order)
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = b"".join([
        _attr("channels", "chlist", chlist + b"\0"),
        _attr("compression", "compression", b"\0"),
        _attr("dataWindow", "box2i", box),
        _attr("displayWindow", "box2i", box),
        _attr("lineOrder", "lineOrder", b"\0"),
        _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)),
        _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)),
        _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)),
    ]) + b"\0"
    head = struct.pack("<ii", MAGIC, 2) + header
    line_bytes = w * c * dt.itemsize
    first = len(head) + 8 * h
    offsets = struct.pack(f"<{h}Q", *[first + y * (8 + line_bytes) for y in range(h)])
    data = a.astype(dt)
    blocks = []
    for y in range(h):
        rows = b"".join(np.ascontiguousarray(data[y, :, i]).tobytes() for i in order)
        blocks.append(struct.pack("<ii", y, line_bytes) + rows)
    return head + offsets + b"".join(blocks)


def decode(buf: bytes) -> np.ndarray:
    """(H, W, C) array (channels R, G, B, A or Y; the file's pixel type) of EXR bytes."""
    magic, version = struct.unpack_from("<ii", buf, 0)
    if magic != MAGIC:
        raise ValueError("not an OpenEXR file")
    if version & 0xFF != 2 or version & ~0xFF:
        raise ValueError(f"only single-part scanline EXR is supported (version field {version:#x})")
    pos, attrs = 8, {}
    while buf[pos] != 0:
        end = buf.index(b"\0", pos)
        name = buf[pos:end].decode()
        tend = buf.index(b"\0", end + 1)
        kind = buf[end + 1:tend].decode()
        (size,) = struct.unpack_from("<i", buf, tend + 1)
        attrs[name] = (kind, buf[tend + 5:tend + 5 + size])
        pos = tend + 5 + size
    pos += 1
    if attrs.get("compression", ("", b"\0"))[1] != b"\0":
        raise ValueError("compressed EXR is not supported")
    if attrs.get("lineOrder", ("", b"\0"))[1] != b"\0":
        raise ValueError("only INCREASING_Y line order is supported")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chl, p, chans = attrs["channels"][1], 0, []
    while chl[p] != 0:
        end = chl.index(b"\0", p)
        name = chl[p:end].decode()
        ptype, _, xs, ys = struct.unpack_from("<iB3xii", chl, end + 1)
        if xs != 1 or ys != 1:
            raise ValueError("subsampled channels are not supported")
        chans.append((name, _TYPES[ptype]))
        p = end + 17
    names = [n for n, _ in chans]
    want = next((v for v in _NAMES.values() if sorted(v) == sorted(names)), None)
    if want is None or len({t for _, t in chans}) != 1:
        raise ValueError(f"unsupported channel set {names}")
    dt = chans[0][1]
    offsets = struct.unpack_from(f"<{h}Q", buf, pos)
    out = np.empty((h, w, len(names)), dt)
    line_bytes = w * len(names) * dt.itemsize
    for off in offsets:
        y, size = struct.unpack_from("<ii", buf, off)
        if size != line_bytes or not y0 <= y <= y1:
            raise ValueError("unexpected scanline block")
        for k, (name, _) in enumerate(chans):
            row = np.frombuffer(buf, dt, w, off + 8 + k * w * dt.itemsize)
            out[y - y0, :, want.index(name)] = row
    return out


def write(path: str, pixels: np.ndarray, pixel_type: str = "float") -> None:
    with open(path, "wb") as f:
        f.write(encode(pixels, pixel_type))


def read(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        return decode(f.read())
