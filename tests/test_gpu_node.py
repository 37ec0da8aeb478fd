"""The node through the plugin ABI on the GPU, replaying the reference's own tests.

tests/unit/test_node_denoise_bm3d.cpp:129-181 and
tests/unit/test_node_denoise_bm3d_params.cpp:13-32, plus the golden gate of
tests/unit/test_golden_isp_nodes.cpp:628-647, with the minimal host
(cpipe_amd/csrc/minihost.cpp) in place of the reference HostContext.
"""
import os

import numpy as np
import pytest

from cpipe_amd import exr, host, synth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def _from_f16(b):
    return np.asarray(b, np.uint16).view(np.float16).astype(np.float32)


def _rmse_rgb(a, b):
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    return float(np.sqrt(np.mean(d * d)))


def _ref_images():
    f = np.float32
    clean = np.zeros((16, 16, 4), np.float32)
    noisy = np.zeros((16, 16, 4), np.float32)
    for y in range(16):
        for x in range(16):
            base = f(0.18) + f(0.006) * f(x) + f(0.004) * f(y)
            c = [base, base + f(0.015), base - f(0.01), f(1.0)]
            n = f(0.055) if ((x + y) & 1) == 0 else f(-0.055)
            clean[y, x] = c
            noisy[y, x] = [c[0] + n, c[1] + n * f(0.75), c[2] + n, c[3]]
    return clean, noisy


def _rgba_input(pixels, noise_profile=None):
    h, w = pixels.shape[:2]
    b = host.Buffer(host.FORMAT_RGBA16F, w, h)
    b.write(_f16(pixels))
    b.set_cs_role("scene_linear_rec2020")
    b.add_step("color_matrix")
    if noise_profile is not None:
        b.set_noise_profile(noise_profile)
    return b


def test_reference_unit_test_compat_mode(cuda):
    d = host.find_node()
    clean, noisy = _ref_images()
    inp = _rgba_input(noisy, [(0.0009, 0.0)] * 3)  # sigma = 0.03 from metadata
    out = host.Buffer(host.FORMAT_RGBA16F, 16, 16)
    assert host.run_node(d, [inp], [out], {}) == host.CPIPE_OK, host.host_log()
    den = _from_f16(out.read())
    noisy_rmse, den_rmse = _rmse_rgb(noisy, clean), _rmse_rgb(den, clean)
    assert den_rmse <= noisy_rmse * 0.20
    assert abs(den[0, 0, 3] - 1.0) <= 0.001
    assert out.cs_role == "scene_linear_rec2020"
    assert out.applied_steps == ["color_matrix", "denoise.bm3d"]
    out2 = host.Buffer(host.FORMAT_RGBA16F, 16, 16)
    assert host.run_node(d, [inp], [out2], {"sigma_override": 0.09}) == host.CPIPE_OK
    assert _rmse_rgb(_from_f16(out2.read()), clean) <= den_rmse
    # the exact known answers of the restated reference arithmetic
    assert den_rmse == pytest.approx(0.003606, abs=5e-7)


def test_reference_params_test(cuda):
    d = host.find_node()
    img = np.array([[0.1, 0.1, 0.1, 1], [0.9, 0.1, 0.1, 1], [0.1, 0.1, 0.1, 1], [0.9, 0.1, 0.1, 1]],
                   np.float32).reshape(2, 2, 4)
    outs = []
    for sigma in (0.0, 0.2):
        inp = _rgba_input(img)
        out = host.Buffer(host.FORMAT_RGBA16F, 2, 2)
        assert host.run_node(d, [inp], [out], {"sigma": sigma}) == host.CPIPE_OK
        outs.append(_from_f16(out.read()))
    assert float(np.max(np.abs(outs[0] - outs[1]))) > 0.1


def test_golden_gate(cuda, oracle_lib):
    """assert_denoise_golden (test_golden_isp_nodes.cpp:628-647,788) on the
    regenerated pair tests/golden/denoise.bm3d/{in,out}.exr."""
    d = host.find_node()
    inp = host.Buffer(host.FORMAT_RGBA16F, 8, 8)
    inp.write(_f16(exr.read(os.path.join(GOLDEN, "denoise.bm3d", "in.exr"))))
    inp.set_cs_role("scene_linear_rec2020")
    inp.add_step("color_matrix")
    out = host.Buffer(host.FORMAT_RGBA16F, 8, 8)
    assert host.run_node(d, [inp], [out]) == host.CPIPE_OK
    got = out.read()
    want = _f16(exr.read(os.path.join(GOLDEN, "denoise.bm3d", "out.exr")))
    np.testing.assert_array_equal(got, want)
    assert synth.psnr(_from_f16(got)[..., :3], _from_f16(want)[..., :3]) >= 40.0


def test_bm3d_mode_rgba16f_and_metadata_sigma(cuda, oracle_lib):
    d = host.find_node()
    clean, noisy, sigma = synth.workload("rgb64x48", seed=21)
    rgba = np.ones((48, 64, 4), np.float32)
    rgba[..., :3] = np.moveaxis(noisy, 0, -1)
    inp = _rgba_input(rgba, [(0.002 ** 2, 0.01)] * 3)
    out = host.Buffer(host.FORMAT_RGBA16F, 64, 48)
    assert host.run_node(d, [inp], [out], {"mode": "bm3d"}) == host.CPIPE_OK, host.host_log()
    got = out.read()
    planes = np.moveaxis(_from_f16(_f16(rgba))[..., :3], -1, 0).copy()
    want = oracle_lib.bm3d_f32(planes, np.float32(sigma))
    np.testing.assert_array_equal(got[..., :3], _f16(np.moveaxis(want, 0, -1)))
    np.testing.assert_array_equal(got[..., 3], _f16(rgba[..., 3]))
    assert out.applied_steps == ["color_matrix", "denoise.bm3d"]


@pytest.mark.parametrize("channels", [1, 3])
def test_bm3d_mode_fp32_planes_with_row_padding(cuda, oracle_lib, channels):
    d = host.find_node()
    clean, noisy, sigma = synth.workload("gray70x52" if channels == 1 else "rgb70x52", seed=5)
    ins, outs = [], []
    for c in range(channels):
        b = host.Buffer(host.FORMAT_R32F, 70, 52, row_bytes=80 * 4)
        b.write(noisy[c])
        b.add_step("color_matrix")
        ins.append(b)
        outs.append(host.Buffer(host.FORMAT_R32F, 70, 52, row_bytes=72 * 4))
    assert host.run_node(d, ins, outs, {"mode": "bm3d", "sigma": float(sigma)}) == host.CPIPE_OK, host.host_log()
    got = np.stack([o.read() for o in outs])
    want = oracle_lib.bm3d_f32(noisy, np.float32(np.clip(sigma, 0, 0.2)))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_node_plan_cache_reuse_and_eviction(cuda, oracle_lib):
    """Plans and buffers are cached across node calls (the host creates and
    destroys the node on every run): repeated shapes, more shapes than cache
    slots (eviction), and a shape coming back after eviction all stay
    bit-identical to the CPU FP32 restatement."""
    d = host.find_node()
    shapes = [(40, 36), (33, 41), (48, 32), (40, 36), (29, 30), (52, 44), (31, 37), (33, 41)]
    for i, (w, h) in enumerate(shapes):
        clean, noisy, sigma = synth.workload(f"gray{w}x{h}", seed=40 + i)
        b = host.Buffer(host.FORMAT_R32F, w, h)
        b.write(noisy[0])
        b.add_step("color_matrix")
        o = host.Buffer(host.FORMAT_R32F, w, h)
        print(f"call {i}: {w}x{h}", flush=True)
        assert host.run_node(d, [b], [o], {"mode": "bm3d", "sigma": float(sigma)}) == host.CPIPE_OK, host.host_log()
        want = oracle_lib.bm3d_f32(noisy, np.float32(sigma))
        np.testing.assert_array_equal(o.read().view(np.uint32), want[0].view(np.uint32))


def test_node_learned_pilot_param(cuda):
    """pilot=learned through the plugin ABI equals an engine plan with the
    packaged weights (the node's copy is embedded in the library)."""
    import torch
    from cpipe_amd import engine, pilot
    d = host.find_node()
    w, h = 96, 80
    clean, noisy, sigma = synth.workload(f"rgb{w}x{h}", seed=3)
    ins, outs = [], []
    for c in range(3):
        b = host.Buffer(host.FORMAT_R32F, w, h)
        b.write(noisy[c])
        b.add_step("color_matrix")
        ins.append(b)
        outs.append(host.Buffer(host.FORMAT_R32F, w, h))
    params = {"mode": "bm3d", "precision": "fp16", "pilot": "learned", "sigma": float(sigma)}
    assert host.run_node(d, ins, outs, params) == host.CPIPE_OK, host.host_log()
    got = np.stack([o.read() for o in outs])
    plan = engine.Plan(w, h, 3, "fp16", 0)
    plan.set_pilot(pilot.Pilot(seed=None, weights=pilot.DEFAULT_WEIGHTS))
    src = torch.from_numpy(noisy).cuda()
    dst = torch.empty_like(src)
    plan.run(src, dst, sigma)
    torch.cuda.synchronize()
    plan.close()
    np.testing.assert_array_equal(got.view(np.uint32), dst.cpu().numpy().view(np.uint32))
    # grey input with the learned pilot is refused; compat mode with it is a bad parameter
    g = host.Buffer(host.FORMAT_R32F, w, h)
    g.write(noisy[0])
    g.add_step("color_matrix")
    go = host.Buffer(host.FORMAT_R32F, w, h)
    assert host.run_node(d, [g], [go], params) == host.CPIPE_UNSUPPORTED
    assert host.run_node(d, ins, outs, {"pilot": "learned"}) == host.CPIPE_BAD_INDEX
