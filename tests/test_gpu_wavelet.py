"""The wavelet BayesShrink node (com.cpipe.denoise.wavelet_bayes_shrink) on the GPU.

Bar: bit-identical binary16 output to the CPU oracle (oracle/wavelet_oracle.c,
the reference Halide generator's arithmetic) for every size, stride and
strength tried, a full 4080x3072 frame included; the reference's own unit test
(tests/unit/test_node_denoise_wavelet_bayes_shrink.cpp:38-87) through the plugin
ABI with the minimal host; the parameter clamp of wavelet_bayes_shrink.cpp:37-38.
"""
import numpy as np
import pytest

import os

from cpipe_amd import engine, exr, host, synth

pytestmark = pytest.mark.gpu

NODE = "com.cpipe.denoise.wavelet_bayes_shrink"


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def _from_f16(b):
    return np.asarray(b, np.uint16).view(np.float16).astype(np.float64)


def _random_bits(rng, h, w):
    img = rng.uniform(-0.2, 1.6, size=(h, w, 4)).astype(np.float32)
    img[rng.random((h, w, 4)) < 0.02] = 0.0
    img[rng.random((h, w, 4)) < 0.02] = 3.0e-6  # binary16 subnormals
    return _f16(img)


def _gpu(torch, bits, strength, pad=0):
    h, w, _ = bits.shape
    src_full = torch.zeros((h, w + pad, 4), dtype=torch.int16, device="cuda")
    src = src_full[:, :w]
    src.copy_(torch.from_numpy(bits.view(np.int16)))
    dst_full = torch.full((h, w + pad + 1, 4), -1, dtype=torch.int16, device="cuda")
    dst = dst_full[:, :w]
    engine.wavelet_rgba16f(src, dst, strength)
    torch.cuda.synchronize()
    out = dst_full.cpu().numpy().view(np.uint16)
    assert np.all(out[:, w:] == 0xFFFF), "wrote outside the image"
    return out[:, :w]


@pytest.mark.parametrize("h,w", [(2, 2), (5, 7), (1, 1), (48, 64), (257, 333)])
@pytest.mark.parametrize("strength", [0.0, 0.5, 1.0, 2.0])
def test_matches_oracle_bit_exact(cuda, oracle_lib, h, w, strength):
    rng = np.random.default_rng(h * 1000 + w)
    bits = _random_bits(rng, h, w)
    print(f"wavelet {w}x{h} strength {strength}", flush=True)
    want = oracle_lib.wavelet_rgba16f(bits, strength)
    got = _gpu(cuda, bits, strength, pad=3)
    assert np.array_equal(got, want)


def test_full_frame_bit_exact(cuda, oracle_lib):
    """4080 x 3072: the BM3D workload's scene as RGBA16F (alpha 1)."""
    clean, noisy, _ = synth.workload("rgb12mp", seed=0)
    h, w = noisy.shape[1:]
    rgba = np.empty((h, w, 4), np.float32)
    rgba[..., :3] = np.moveaxis(noisy, 0, -1)
    rgba[..., 3] = 1.0
    bits = _f16(rgba)
    print("wavelet 4080x3072 strength 1", flush=True)
    got = _gpu(cuda, bits, 1.0)
    want = oracle_lib.wavelet_rgba16f(bits, 1.0)
    assert np.array_equal(got, want)


def _node_input(pixels, steps=("color_matrix", "denoise.guided_filter")):
    h, w = pixels.shape[:2]
    b = host.Buffer(host.FORMAT_RGBA16F, w, h)
    b.write(_f16(pixels))
    b.set_cs_role("scene_linear_rec2020")
    for s in steps:
        b.add_step(s)
    return b


def test_reference_unit_test_through_the_node(cuda, oracle_lib):
    """tests/unit/test_node_denoise_wavelet_bayes_shrink.cpp:38-87."""
    px = np.array([[[0.38, 0.30, 0.22, 1.0], [0.22, 0.30, 0.38, 0.8]],
                   [[0.37, 0.30, 0.23, 0.6], [0.23, 0.30, 0.37, 0.4]]], np.float32)
    d = host.find_node(NODE)
    inp = _node_input(px)
    out = host.Buffer(host.FORMAT_RGBA16F, 2, 2)
    assert host.run_node(d, [inp], [out], {}) == host.CPIPE_OK, host.host_log()
    a, b = _from_f16(inp.read()), _from_f16(out.read())
    span = lambda p: float(np.ptp(p[..., 0] - p[..., 2]))
    assert span(b) < span(a)
    luma = lambda p: (p[..., 0] + 2.0 * p[..., 1] + p[..., 2]) * 0.25
    assert np.all(np.abs(luma(b) - luma(a)) <= 0.002)
    assert np.all(np.abs(b[..., 3] - a[..., 3]) <= 0.001)
    assert out.cs_role == "scene_linear_rec2020"
    assert out.applied_steps == ["color_matrix", "denoise.guided_filter", "denoise.wavelet_bayes_shrink"]
    assert np.array_equal(out.read(), oracle_lib.wavelet_rgba16f(inp.read(), 1.0))


def test_node_param_clamp_and_row_padding(cuda, oracle_lib):
    """chroma_strength: absent -> 1, clamped to [0, 2] (wavelet_bayes_shrink.cpp:37-38);
    host buffers with padded rows."""
    rng = np.random.default_rng(11)
    h, w = 9, 13
    bits = _random_bits(rng, h, w)
    d = host.find_node(NODE)
    for params, eff in (({}, 1.0), ({"chroma_strength": 5.0}, 2.0), ({"chroma_strength": -1.0}, 0.0),
                        ({"chroma_strength": 0.7}, 0.7)):
        inp = host.Buffer(host.FORMAT_RGBA16F, w, h, row_bytes=(w + 5) * 8)
        inp.write(bits)
        inp.add_step("color_matrix")
        out = host.Buffer(host.FORMAT_RGBA16F, w, h, row_bytes=(w + 2) * 8)
        assert host.run_node(d, [inp], [out], params) == host.CPIPE_OK, host.host_log()
        want = oracle_lib.wavelet_rgba16f(bits, eff)
        assert np.array_equal(out.read(), want), params
        assert out.applied_steps == ["color_matrix", "denoise.wavelet_bayes_shrink"]


def test_golden_gate(cuda, oracle_lib):
    """assert_denoise_golden for this node (test_golden_isp_nodes.cpp:628-647,795-796)
    on the regenerated pair tests/golden/denoise.wavelet_bayes_shrink/{in,out}.exr:
    PSNR >= 40 dB is the reference's bar; the output here is bit-identical."""
    gdir = os.path.join(os.path.dirname(__file__), "golden", "denoise.wavelet_bayes_shrink")
    inp = _node_input(exr.read(os.path.join(gdir, "in.exr")), steps=("denoise.bm3d",))
    out = host.Buffer(host.FORMAT_RGBA16F, 8, 8)
    assert host.run_node(host.find_node(NODE), [inp], [out], {}) == host.CPIPE_OK, host.host_log()
    want = _f16(exr.read(os.path.join(gdir, "out.exr")))
    np.testing.assert_array_equal(out.read(), want)
    assert synth.psnr(_from_f16(out.read())[..., :3], _from_f16(want)[..., :3]) >= 40.0


def test_host_entry_point(cuda, oracle_lib):
    rng = np.random.default_rng(13)
    bits = _random_bits(rng, 31, 17)
    assert np.array_equal(engine.wavelet_host_rgba16f(bits, 1.3), oracle_lib.wavelet_rgba16f(bits, 1.3))


def test_bm3d_then_wavelet_chain(cuda, oracle_lib):
    """The classic pipeline's edge bm3d.rgb -> wavelet.rgb
    (examples/pipelines/full-classic-pipeline.cpipe.json:127-129): both nodes of
    the library in sequence, the applied-steps ledger in order."""
    rng = np.random.default_rng(17)
    h, w = 24, 32
    px = np.clip(0.4 + 0.05 * rng.standard_normal((h, w, 4)), 0, 1).astype(np.float32)
    px[..., 3] = 1.0
    a = _node_input(px, steps=("color_matrix",))
    a.set_noise_profile([(0.0009, 0.0)] * 3)
    mid = host.Buffer(host.FORMAT_RGBA16F, w, h)
    assert host.run_node(host.find_node(), [a], [mid], {"mode": "bm3d"}) == host.CPIPE_OK, host.host_log()
    out = host.Buffer(host.FORMAT_RGBA16F, w, h)
    assert host.run_node(host.find_node(NODE), [mid], [out], {}) == host.CPIPE_OK, host.host_log()
    assert out.applied_steps == ["color_matrix", "denoise.bm3d", "denoise.wavelet_bayes_shrink"]
    assert np.array_equal(out.read(), oracle_lib.wavelet_rgba16f(mid.read(), 1.0))
