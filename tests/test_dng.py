"""DNG NoiseProfile ingest (include/cpipe_dng.h, cpipe_amd/csrc/dng_noise.cpp), no GPU.

Pins: the committed fixtures (tests/golden/make_dng.py) read back to the pairs
their generator wrote, and to what an independent TIFF reader (Pillow's IFD
parser) finds in them; the raw image's NoiseProfile wins over a preview's; the
DNG [S, O] order becomes (beta, lambda) = (O, S); malformed files are refused;
and no truncation or byte mutation of a fixture makes the parser read outside
the buffer (AddressSanitizer build of the parser, host code only).
"""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

from cpipe_amd import dng, engine

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dng")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FILES = ["le_ifd0.dng", "be_subifd.dng", "le_chain.dng", "none.dng"]


def _expected():
    with open(os.path.join(GOLDEN, "expected.json")) as f:
        return json.load(f)


def _raw(data: bytes, max_pairs=16):
    buf = (ctypes.c_char * max(1, len(data))).from_buffer_copy(data or b"\0")
    beta = (ctypes.c_float * max(1, max_pairs))()
    lam = (ctypes.c_float * max(1, max_pairs))()
    n = ctypes.c_int(-1)
    rc = engine.lib().cpipe_dng_noise_profile(buf, len(data), beta, lam, max_pairs, ctypes.byref(n))
    return rc, n.value, [(beta[i], lam[i]) for i in range(min(max(n.value, 0), max_pairs))]


def _pil_noise_profiles(path):
    """(NewSubFileType, NoiseProfile) of every IFD, by Pillow's TIFF IFD reader."""
    from PIL.TiffImagePlugin import ImageFileDirectory_v2
    out = []
    with open(path, "rb") as fp:
        hdr = fp.read(8)
        todo, seen = [ImageFileDirectory_v2(hdr).next], set()
        while todo:
            off = todo.pop(0)
            if not off or off in seen:
                continue
            seen.add(off)
            d = ImageFileDirectory_v2(hdr)
            fp.seek(off)
            d.load(fp)
            t = dict(d)
            out.append((t.get(254, 0), t.get(51041)))
            sub = t.get(330)
            if sub is not None:
                todo += list(sub) if isinstance(sub, tuple) else [sub]
            todo.append(d.next)
    return out


@pytest.mark.parametrize("name", FILES)
def test_fixtures_read_back(name):
    want = _expected()[name]
    got = dng.noise_profile(os.path.join(GOLDEN, name))
    if want is None:
        assert got == []
        return
    assert len(got) == len(want)
    for (gb, gl), (wb, wl) in zip(got, want):
        assert gb == np.float32(wb) and gl == np.float32(wl)  # the doubles rounded once to float


@pytest.mark.parametrize("name", FILES)
def test_fixtures_agree_with_an_independent_tiff_reader(name):
    ifds = _pil_noise_profiles(os.path.join(GOLDEN, name))
    main = [np_ for sub, np_ in ifds if sub == 0 and np_ is not None]
    other = [np_ for sub, np_ in ifds if sub != 0 and np_ is not None]
    chosen = main[0] if main else (other[0] if other else None)
    got = dng.noise_profile(os.path.join(GOLDEN, name))
    if chosen is None:
        assert got == []
        return
    s_o = list(zip(chosen[0::2], chosen[1::2]))  # the file stores [S, O] pairs
    assert got == [(float(np.float32(o)), float(np.float32(s))) for s, o in s_o]
    if name == "be_subifd.dng":
        assert other and other[0] != chosen  # the preview's decoy was not taken


def test_pairs_feed_the_metadata_sigma_rule():
    """bm3d.cpp:49-54: sigma = sqrt(mean_c max(0, beta_c + lambda_c / 4))."""
    from cpipe_amd import synth
    pairs = dng.noise_profile(os.path.join(GOLDEN, "le_ifd0.dng"))
    want = float(np.float32(np.sqrt(np.mean([max(0.0, b + 0.25 * l) for b, l in pairs]))))
    assert abs(synth.sigma_from_noise_profile(pairs) - want) <= 1e-7


def test_max_pairs_and_bad_arguments():
    data = open(os.path.join(GOLDEN, "le_chain.dng"), "rb").read()
    rc, n, pairs = _raw(data, max_pairs=2)
    assert rc == 0 and n == 4 and len(pairs) == 2  # the count of the file, the first two written
    rc, n, _ = _raw(data, max_pairs=0)
    assert rc == 0 and n == 4
    L = engine.lib()
    cnt = ctypes.c_int()
    assert L.cpipe_dng_noise_profile(None, 10, None, None, 0, ctypes.byref(cnt)) == 5
    buf = ctypes.create_string_buffer(data)
    assert L.cpipe_dng_noise_profile(buf, len(data), None, None, 4, ctypes.byref(cnt)) == 5
    assert L.cpipe_dng_noise_profile(buf, len(data), None, None, 0, None) == 5


def _patch_noise_entry(data: bytes, field: str, value: int) -> bytes:
    """Rewrites the type or count of the (little-endian) NoiseProfile entry of IFD0."""
    import struct
    ifd = struct.unpack_from("<I", data, 4)[0]
    n = struct.unpack_from("<H", data, ifd)[0]
    out = bytearray(data)
    for e in range(n):
        at = ifd + 2 + 12 * e
        if struct.unpack_from("<H", data, at)[0] == 51041:
            if field == "type":
                struct.pack_into("<H", out, at + 2, value)
            else:
                struct.pack_into("<I", out, at + 4, value)
    return bytes(out)


def test_malformed_files_are_refused():
    data = open(os.path.join(GOLDEN, "le_ifd0.dng"), "rb").read()
    assert _raw(b"not a tiff at all")[0] == 1
    assert _raw(b"II*\0")[0] == 1                              # header cut short
    assert _raw(b"II\x2b\0\x08\0\0\0" + data[8:])[0] == 1      # BigTIFF magic: not classic TIFF
    assert _raw(_patch_noise_entry(data, "type", 4))[0] == 1   # LONG instead of DOUBLE
    assert _raw(_patch_noise_entry(data, "count", 5))[0] == 1  # odd count
    assert _raw(_patch_noise_entry(data, "count", 0))[0] == 1
    assert _raw(_patch_noise_entry(data, "count", 1 << 20))[0] == 1  # data past the end
    nan = bytearray(data)
    # the first double of the NoiseProfile data (S0) -> NaN
    import struct
    ifd = struct.unpack_from("<I", data, 4)[0]
    n = struct.unpack_from("<H", data, ifd)[0]
    for e in range(n):
        at = ifd + 2 + 12 * e
        if struct.unpack_from("<H", data, at)[0] == 51041:
            struct.pack_into("<d", nan, struct.unpack_from("<I", data, at + 8)[0], float("nan"))
    assert _raw(bytes(nan))[0] == 1


def test_truncations_and_mutations_stay_in_bounds(tmp_path):
    """Every prefix of every fixture and 3000 seeded byte mutations through the
    parser built with AddressSanitizer: a read outside the buffer aborts the run."""
    driver = tmp_path / "drive.cpp"
    driver.write_text(r'''
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "cpipe_dng.h"
int main(int argc, char** argv) {
    int seen[3] = {0, 0, 0};
    for (int f = 1; f < argc; ++f) {
        FILE* fp = std::fopen(argv[f], "rb");
        std::vector<unsigned char> all;
        int ch;
        while ((ch = std::fgetc(fp)) != EOF) all.push_back((unsigned char)ch);
        std::fclose(fp);
        float b[16], l[16];
        int n = 0;
        for (size_t len = 0; len <= all.size(); ++len) {  // exact-size heap copies
            unsigned char* p = (unsigned char*)std::malloc(len ? len : 1);
            std::memcpy(p, all.data(), len);
            int rc = cpipe_dng_noise_profile(p, len, b, l, 16, &n);
            seen[rc == 0 ? 0 : rc == 1 ? 1 : 2]++;
            std::free(p);
        }
        unsigned s = 12345u + f;
        for (int k = 0; k < 1000; ++k) {
            std::vector<unsigned char> m = all;
            for (int j = 0; j < 1 + k % 4; ++j) {
                s = s * 1103515245u + 12345u;
                m[(s >> 8) % m.size()] = (unsigned char)(s >> 20);
            }
            unsigned char* p = (unsigned char*)std::malloc(m.size());
            std::memcpy(p, m.data(), m.size());
            int rc = cpipe_dng_noise_profile(p, m.size(), b, l, 16, &n);
            seen[rc == 0 ? 0 : rc == 1 ? 1 : 2]++;
            std::free(p);
        }
    }
    std::printf("%d %d %d\n", seen[0], seen[1], seen[2]);
    return 0;
}
''')
    exe = tmp_path / "drive"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    f"-I{os.path.join(ROOT, 'include')}", str(driver),
                    os.path.join(ROOT, "cpipe_amd", "csrc", "dng_noise.cpp"), "-o", str(exe)], check=True)
    files = [os.path.join(GOLDEN, f) for f in FILES[:3]]
    r = subprocess.run([str(exe), *files], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    ok, failed, other = (int(v) for v in r.stdout.split())
    assert other == 0 and ok > 0 and failed > 0
