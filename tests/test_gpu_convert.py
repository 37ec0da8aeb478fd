"""The precision-convert node (com.cpipe.precision_convert) on the GPU.

Bar: output bit-identical to the CPU oracle (oracle/convert_oracle.c, the
reference Halide generators' arithmetic) for all four conversions, with every
binary16 input pattern for the UNORM ones, padded strides and odd sizes; the
reference's own unit test (tests/unit/test_node_precision_convert.cpp:80-120)
through the plugin ABI with the minimal host.
"""
import numpy as np
import pytest

from cpipe_amd import engine, host

pytestmark = pytest.mark.gpu

NODE = "com.cpipe.precision_convert"


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def _padded(torch, arr, pad, fill=0):
    """A CUDA copy of arr in a buffer `pad` pixels wider (the view keeps the width)."""
    h, w = arr.shape[:2]
    shape = (h, w + pad) + arr.shape[2:]
    dt = {np.uint16: torch.int16, np.uint8: torch.uint8, np.float32: torch.float32}[arr.dtype.type]
    full = torch.full(shape, fill, dtype=dt, device="cuda")
    view = full[:, :w]
    src = arr.view(np.int16) if arr.dtype == np.uint16 else arr
    view.copy_(torch.from_numpy(np.ascontiguousarray(src)))
    return full, view


def _run(torch, arr, src_fmt, dst_fmt, dst_dtype, dst_ch):
    h, w = arr.shape[:2]
    _, src = _padded(torch, arr, 3)
    out_shape = (h, w, dst_ch) if dst_ch > 1 else (h, w)
    fill = 255 if dst_dtype == np.uint8 else -1
    full, dst = _padded(torch, np.zeros(out_shape, dst_dtype), 2, fill=fill)
    engine.convert(src, src_fmt, dst, dst_fmt)
    torch.cuda.synchronize()
    got = full.cpu().numpy()
    assert np.all(got[:, w:] == fill), "wrote outside the image"
    if dst_dtype == np.uint16:
        got = got.view(np.uint16)
    return got[:, :w]


@pytest.mark.parametrize("target,dtype", [("R8G8B8A8_UNORM", np.uint8), ("R16G16B16A16_UNORM", np.uint16)])
def test_unorm_every_half_pattern(cuda, oracle_lib, target, dtype):
    bits = np.arange(65536, dtype=np.uint16).reshape(256, 64, 4)
    print(f"convert rgba16f -> {target}", flush=True)
    got = _run(cuda, bits, "R16G16B16A16_SFLOAT", target, dtype, 4)
    assert np.array_equal(got, oracle_lib.convert(bits, target))


@pytest.mark.parametrize("h,w", [(1, 1), (5, 7), (37, 300)])
def test_f32_to_rgba16f(cuda, oracle_lib, h, w):
    rng = np.random.default_rng(h * w)
    v = rng.uniform(-2, 2, (h, w)).astype(np.float32)
    v.flat[0] = 70000.0
    v.flat[-1] = 3.0e-8
    got = _run(cuda, v, "R32_SFLOAT", "R16G16B16A16_SFLOAT", np.uint16, 4)
    assert np.array_equal(got, oracle_lib.convert(v, "R16G16B16A16_SFLOAT"))


def test_r16u_to_f32_every_value(cuda, oracle_lib):
    src = np.arange(65536, dtype=np.uint16).reshape(128, 512)
    got = _run(cuda, src, "R16_UINT", "R32_SFLOAT", np.float32, 1)
    assert np.array_equal(got, oracle_lib.convert(src, "R32_SFLOAT"))


def test_reference_unit_test_through_the_node(cuda, oracle_lib):
    """tests/unit/test_node_precision_convert.cpp:80-120: rgba16f -> rgba8 within one LSB."""
    d = host.find_node(NODE)
    inp = host.Buffer(host.FORMAT_RGBA16F, 2, 1)
    px = _f16([[[0.0, 0.5, 1.0, 1.0], [0.25, 0.75, 1.25, -0.5]]])
    inp.write(px)
    inp.set_cs_role("display")
    out = host.Buffer(host.FORMAT_RGBA8, 2, 1)
    assert host.run_node(d, [inp], [out], {"target_format": "R8G8B8A8_UNORM"}) == host.CPIPE_OK, host.host_log()
    b = out.read().reshape(-1).astype(int)
    assert b[0] == 0 and abs(b[1] - 128) <= 1 and b[2] == 255 and b[3] == 255
    assert abs(b[4] - 64) <= 1 and abs(b[5] - 191) <= 1 and b[6] == 255 and b[7] == 0
    assert np.array_equal(out.read(), oracle_lib.convert(px, "R8G8B8A8_UNORM"))
    assert out.cs_role == "display" and out.applied_steps == ["precision_convert"]


def test_node_every_pair_with_padded_rows(cuda, oracle_lib):
    d = host.find_node(NODE)
    rng = np.random.default_rng(5)
    h, w = 7, 11
    cases = [
        (host.FORMAT_R16U, rng.integers(0, 65536, (h, w), dtype=np.uint16), "R32_SFLOAT", host.FORMAT_R32F),
        (host.FORMAT_R32F, rng.uniform(-1, 2, (h, w)).astype(np.float32), "R16G16B16A16_SFLOAT",
         host.FORMAT_RGBA16F),
        (host.FORMAT_RGBA16F, _f16(rng.uniform(-0.5, 1.5, (h, w, 4))), "R8G8B8A8_UNORM", host.FORMAT_RGBA8),
        (host.FORMAT_RGBA16F, _f16(rng.uniform(-0.5, 1.5, (h, w, 4))), "R16G16B16A16_UNORM",
         host.FORMAT_RGBA16U),
    ]
    for fin, arr, target, fout in cases:
        inp = host.Buffer(fin, w, h, row_bytes=(w + 3) * arr.itemsize * (4 if arr.ndim == 3 else 1))
        inp.write(arr)
        out = host.Buffer(fout, w, h, row_bytes=0)
        assert host.run_node(d, [inp], [out], {"target_format": target}) == host.CPIPE_OK, (target, host.host_log())
        want = oracle_lib.convert(arr, target)
        assert np.array_equal(out.read().reshape(want.shape), want), target
