"""Pins for the precision-convert oracle (runs without a GPU).

The oracle restates src/cpipe/nodes/precision/precision_convert_generator.cpp.
Pins: the reference's own unit test (tests/unit/test_node_precision_convert.cpp:80-120),
numpy's independent float16 and integer conversions, and an exhaustive
comparison of the FP32 quantisation with the exact rational one.
"""
from fractions import Fraction

import numpy as np


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def test_reference_unit_test(oracle_lib):
    """rgba16f -> rgba8 within one LSB: 0 -> 0, 0.5 -> 128, 1 -> 255, 0.25 -> 64,
    0.75 -> 191, 1.25 -> 255 (clamp), -0.5 -> 0 (clamp)."""
    px = _f16([[[0.0, 0.5, 1.0, 1.0], [0.25, 0.75, 1.25, -0.5]]])
    b = oracle_lib.convert(px, "R8G8B8A8_UNORM").reshape(-1).astype(int)
    assert b[0] == 0 and abs(b[1] - 128) <= 1 and b[2] == 255 and b[3] == 255
    assert abs(b[4] - 64) <= 1 and abs(b[5] - 191) <= 1 and b[6] == 255 and b[7] == 0


def test_unorm_quantisation_exhaustive(oracle_lib):
    """Every finite binary16 value: the FP32 (v * scale) + 0.5 then truncation agrees with
    round-half-up of the exact product clamp(v) * scale except where the FP32 product
    rounds across a half (there it is off by one at most), and is monotone."""
    bits = np.arange(65536, dtype=np.uint16)
    vals = bits.view(np.float16).astype(np.float64)
    fin = np.isfinite(vals)
    img = bits.reshape(256, 64, 4)
    for target, scale in (("R8G8B8A8_UNORM", 255), ("R16G16B16A16_UNORM", 65535)):
        got = oracle_lib.convert(img, target).reshape(-1).astype(np.int64)
        exact = np.array([int(Fraction(min(max(v, 0.0), 1.0)) * scale + Fraction(1, 2)) if f else -1
                          for v, f in zip(vals, fin)])
        diff = np.abs(got[fin] - exact[fin])
        assert diff.max() <= 1 and np.mean(diff == 0) > 0.999, target
        order = np.argsort(vals[fin], kind="stable")
        assert np.all(np.diff(got[fin][order]) >= 0), target
        assert got[bits == _f16(1.0)[()]][0] == scale and got[bits == 0][0] == 0


def test_f32_to_rgba16f_matches_numpy_and_sets_alpha(oracle_lib):
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.uniform(-3, 3, 4000), rng.uniform(-1e-6, 1e-6, 1000),
                        [0.0, -0.0, 65504.0, 70000.0, -1e-9, 6.0e-8]]).astype(np.float32)
    src = v.reshape(1, -1)
    out = oracle_lib.convert(src, "R16G16B16A16_SFLOAT")
    want = v.astype(np.float16).view(np.uint16)
    assert np.array_equal(out[0, :, 0], want) and np.array_equal(out[0, :, 1], want)
    assert np.array_equal(out[0, :, 2], want) and np.all(out[..., 3] == _f16(1.0)[()])


def test_r16u_to_f32_is_exact(oracle_lib):
    src = np.arange(65536, dtype=np.uint16).reshape(256, 256)
    assert np.array_equal(oracle_lib.convert(src, "R32_SFLOAT"), src.astype(np.float32))
