"""Device-resident buffers through the plugin ABI on the GPU.

The minimal host offers the "hip_buffer" suite (include/cpipe_node_abi.h):
buffers in HIP device memory reach the nodes as device addresses, so the chain
bm3d -> wavelet_bayes_shrink (the classic pipeline's edge,
examples/pipelines/full-classic-pipeline.cpipe.json:127-129) -> precision_convert
(the conversion the reference's precision planner inserts before an 8-bit
output) runs with no host copy between nodes.

Bar: every node's output bit-identical to the CPU oracle and to the same node on
host buffers; mixed residency (some buffers on the host) and a host without the
suite go through lock_cpu and give the same bits.
"""
import numpy as np
import pytest

from cpipe_amd import host, synth

pytestmark = pytest.mark.gpu

WAVELET = "com.cpipe.denoise.wavelet_bayes_shrink"
CONVERT = "com.cpipe.precision_convert"


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def _from_f16(b):
    return np.asarray(b, np.uint16).view(np.float16).astype(np.float32)


def _buf(fmt, w, h, pad_px=0, device=0, data=None):
    bpp = {host.FORMAT_RGBA16F: 8, host.FORMAT_R32F: 4, host.FORMAT_RGBA8: 4, host.FORMAT_R16U: 2,
           host.FORMAT_RGBA16U: 8}[fmt]
    b = host.Buffer(fmt, w, h, row_bytes=(w + pad_px) * bpp, device=device)
    if data is not None:
        b.write(data)
    return b


def _noisy_rgba(name, seed):
    clean, noisy, sigma = synth.workload(name, seed=seed)
    h, w = noisy.shape[1:]
    rgba = np.ones((h, w, 4), np.float32)
    rgba[..., :3] = np.moveaxis(noisy, 0, -1)
    rgba[..., 3] = np.linspace(0.2, 1.0, w, dtype=np.float32)[None, :]
    return _f16(rgba), float(sigma)


def _bm3d_want(oracle_lib, bits, sigma):
    planes = np.moveaxis(_from_f16(bits)[..., :3], -1, 0).copy()
    want = oracle_lib.bm3d_f32(planes, np.float32(sigma))
    out = bits.copy()
    out[..., :3] = _f16(np.moveaxis(want, 0, -1))
    return out


def test_bm3d_rgba16f_on_device_matches_oracle(cuda, oracle_lib):
    bits, sigma = _noisy_rgba("rgb64x48", 21)
    d = host.find_node()
    inp = _buf(host.FORMAT_RGBA16F, 64, 48, pad_px=3, data=bits)
    inp.add_step("color_matrix")
    out = _buf(host.FORMAT_RGBA16F, 64, 48, pad_px=1)
    assert inp.on_device and out.on_device
    print("bm3d rgba16f 64x48 device buffers", flush=True)
    assert host.run_node(d, [inp], [out], {"mode": "bm3d", "sigma": sigma}) == host.CPIPE_OK, host.host_log()
    np.testing.assert_array_equal(out.read(), _bm3d_want(oracle_lib, bits, sigma))
    assert out.applied_steps == ["color_matrix", "denoise.bm3d"]


@pytest.mark.parametrize("channels", [1, 3])
def test_bm3d_planes_on_device_match_oracle(cuda, oracle_lib, channels):
    clean, noisy, sigma = synth.workload("gray70x52" if channels == 1 else "rgb70x52", seed=5)
    d = host.find_node()
    ins = [_buf(host.FORMAT_R32F, 70, 52, pad_px=10, data=noisy[c]) for c in range(channels)]
    outs = [_buf(host.FORMAT_R32F, 70, 52, pad_px=2) for _ in range(channels)]
    print(f"bm3d planes x{channels} 70x52 device buffers", flush=True)
    assert host.run_node(d, ins, outs, {"mode": "bm3d", "sigma": float(sigma)}) == host.CPIPE_OK, host.host_log()
    got = np.stack([o.read() for o in outs])
    want = oracle_lib.bm3d_f32(noisy, np.float32(sigma))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_compat_mode_on_device_matches_oracle(cuda, oracle_lib):
    bits, _ = _noisy_rgba("rgb33x21", 3)
    d = host.find_node()
    inp = _buf(host.FORMAT_RGBA16F, 33, 21, pad_px=2, data=bits)
    out = _buf(host.FORMAT_RGBA16F, 33, 21)
    assert host.run_node(d, [inp], [out], {"sigma": 0.04}) == host.CPIPE_OK, host.host_log()
    np.testing.assert_array_equal(out.read(), oracle_lib.compat_rgba16f(bits, 0.04, 1))


def _chain(bits, sigma, device):
    """bm3d -> wavelet -> convert to RGBA8 with every buffer on `device` (None: host)."""
    h, w = bits.shape[:2]
    a = _buf(host.FORMAT_RGBA16F, w, h, pad_px=1, device=device, data=bits)
    a.set_cs_role("scene_linear_rec2020")
    a.add_step("color_matrix")
    mid = _buf(host.FORMAT_RGBA16F, w, h, device=device)
    wav = _buf(host.FORMAT_RGBA16F, w, h, pad_px=5, device=device)
    rgba8 = _buf(host.FORMAT_RGBA8, w, h, pad_px=3, device=device)
    rc = host.run_node(host.find_node(), [a], [mid], {"mode": "bm3d", "sigma": sigma})
    assert rc == host.CPIPE_OK, host.host_log()
    assert host.run_node(host.find_node(WAVELET), [mid], [wav], {}) == host.CPIPE_OK, host.host_log()
    rc = host.run_node(host.find_node(CONVERT), [wav], [rgba8], {"target_format": "R8G8B8A8_UNORM"})
    assert rc == host.CPIPE_OK, host.host_log()
    return mid, wav, rgba8


def test_chain_on_device_matches_oracle_and_host_path(cuda, oracle_lib):
    bits, sigma = _noisy_rgba("rgb69x51", 8)  # odd sizes: the wavelet's repeated edge
    print("chain bm3d -> wavelet -> convert, 69x51, device buffers", flush=True)
    mid, wav, rgba8 = _chain(bits, sigma, device=0)
    want_mid = _bm3d_want(oracle_lib, bits, sigma)
    want_wav = oracle_lib.wavelet_rgba16f(want_mid, 1.0)
    want_8 = oracle_lib.convert(want_wav, "R8G8B8A8_UNORM")
    np.testing.assert_array_equal(mid.read(), want_mid)
    np.testing.assert_array_equal(wav.read(), want_wav)
    np.testing.assert_array_equal(rgba8.read(), want_8)
    assert rgba8.applied_steps == ["color_matrix", "denoise.bm3d", "denoise.wavelet_bayes_shrink",
                                   "precision_convert"]
    assert rgba8.cs_role == "scene_linear_rec2020"
    print("same chain, host buffers", flush=True)
    hmid, hwav, h8 = _chain(bits, sigma, device=None)
    for dv, hv in ((mid, hmid), (wav, hwav), (rgba8, h8)):
        np.testing.assert_array_equal(dv.read(), hv.read())


def test_mixed_residency_and_host_without_the_suite(cuda, oracle_lib):
    """Input on the device and output on the host (and the reverse), and all
    device buffers behind a host that does not offer "hip_buffer": the nodes go
    through lock_cpu (the device buffers stage through their host copies)."""
    bits, sigma = _noisy_rgba("rgb40x36", 12)
    want = _bm3d_want(oracle_lib, bits, sigma)
    d = host.find_node()
    cases = [(0, None), (None, 0), (0, 0)]
    for k, (din, dout) in enumerate(cases):
        host.offer_hip_buffer(k != 2)
        try:
            inp = _buf(host.FORMAT_RGBA16F, 40, 36, pad_px=2, device=din, data=bits)
            out = _buf(host.FORMAT_RGBA16F, 40, 36, device=dout)
            print(f"bm3d 40x36 input device={din} output device={dout} suite={k != 2}", flush=True)
            rc = host.run_node(d, [inp], [out], {"mode": "bm3d", "sigma": sigma})
            assert rc == host.CPIPE_OK, host.host_log()
            np.testing.assert_array_equal(out.read(), want)
            w8 = _buf(host.FORMAT_RGBA8, 40, 36, device=dout)
            rc = host.run_node(host.find_node(CONVERT), [out], [w8], {"target_format": "R8G8B8A8_UNORM"})
            assert rc == host.CPIPE_OK, host.host_log()
            np.testing.assert_array_equal(w8.read(), oracle_lib.convert(want, "R8G8B8A8_UNORM"))
        finally:
            host.offer_hip_buffer(True)


def test_mixed_precision_on_device_equals_host_path(cuda):
    """precision fp16 (MFMA kernels): the device path runs the same plan as the
    host path, so the bits agree."""
    bits, sigma = _noisy_rgba("rgb64x48", 31)
    d = host.find_node()
    outs = []
    for dev in (0, None):
        inp = _buf(host.FORMAT_RGBA16F, 64, 48, device=dev, data=bits)
        out = _buf(host.FORMAT_RGBA16F, 64, 48, device=dev)
        rc = host.run_node(d, [inp], [out], {"mode": "bm3d", "sigma": sigma, "precision": "fp16"})
        assert rc == host.CPIPE_OK, host.host_log()
        outs.append(out.read())
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("src_fmt,target,dst_fmt", [
    (host.FORMAT_R16U, "R32_SFLOAT", host.FORMAT_R32F),
    (host.FORMAT_R32F, "R16G16B16A16_SFLOAT", host.FORMAT_RGBA16F),
    (host.FORMAT_RGBA16F, "R16G16B16A16_UNORM", host.FORMAT_RGBA16U),
])
def test_convert_on_device_matches_oracle(cuda, oracle_lib, src_fmt, target, dst_fmt):
    rng = np.random.default_rng(src_fmt)
    h, w = 9, 23
    if src_fmt == host.FORMAT_R16U:
        arr = rng.integers(0, 65536, (h, w), dtype=np.uint16)
    elif src_fmt == host.FORMAT_R32F:
        arr = rng.uniform(-1, 2, (h, w)).astype(np.float32)
    else:
        arr = _f16(rng.uniform(-0.5, 1.5, (h, w, 4)))
    inp = _buf(src_fmt, w, h, pad_px=3, data=arr)
    out = _buf(dst_fmt, w, h, pad_px=1)
    rc = host.run_node(host.find_node(CONVERT), [inp], [out], {"target_format": target})
    assert rc == host.CPIPE_OK, host.host_log()
    want = oracle_lib.convert(arr, target)
    np.testing.assert_array_equal(out.read().reshape(want.shape), want)
