"""The noise-profile sigma map on the GPU (cpipe_bm3d_plan_set_noise_profile,
the node's noise_map=profile).

Bar: exact FP32 plans bit-identical to the CPU FP32 restatement
(oracle_bm3d_f32_profile) for grey and RGB, per-channel and shared profiles,
odd sizes and the clamp, and per frame of a batch (eager and graph replays); mixed-precision plans within the mixed-precision bar
of the FP64 oracle with the same map; switching the map off returns the
scalar path bit for bit; the node reads the profile from the input's
calibration metadata (the pairs of bm3d.cpp:29-52) on host and device buffers.
"""
import numpy as np
import pytest

from cpipe_amd import engine, host, synth

pytestmark = pytest.mark.gpu


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def _from_f16(b):
    return np.asarray(b, np.uint16).view(np.float16).astype(np.float32)


def _run_plan(torch, noisy, precision="fp32", profile=None, sigma=0.05):
    c, h, w = noisy.shape
    plan = engine.Plan(w, h, c, precision=precision)
    if profile is not None:
        plan.set_noise_profile(*profile)
    src = torch.from_numpy(np.ascontiguousarray(noisy)).cuda()
    dst = torch.empty_like(src)
    plan.run(src, dst, sigma)
    torch.cuda.synchronize()
    plan.check()
    return plan, dst.cpu().numpy()


@pytest.mark.parametrize("name,beta,lam,shift", [
    ("rgb96x80", 0.002 ** 2, 0.01, 0.0),
    ("gray71x53", 1e-5, 0.02, 0.0),
    ("rgb67x45", [1e-4, 4e-6, 2e-5], [0.0, 0.03, 0.01], 0.3),
])
def test_exact_fp32_matches_oracle(cuda, oracle_lib, name, beta, lam, shift):
    _, noisy, _ = synth.workload(name, seed=5)
    noisy = (noisy - np.float32(shift)).astype(np.float32)
    print(f"profile {name} beta={beta} lambda={lam}", flush=True)
    _, got = _run_plan(cuda, noisy, profile=(beta, lam))
    want = oracle_lib.bm3d_f32_profile(noisy, beta, lam)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_map_off_returns_scalar_path(cuda, oracle_lib):
    _, noisy, sigma = synth.workload("rgb64x48", seed=6)
    plan, got = _run_plan(cuda, noisy, profile=(0.002 ** 2, 0.01), sigma=sigma)
    plan.set_noise_profile(None)
    src = cuda.from_numpy(noisy).cuda()
    dst = cuda.empty_like(src)
    plan.run(src, dst, sigma)
    cuda.cuda.synchronize()
    want = oracle_lib.bm3d_f32(noisy, np.float32(sigma))
    np.testing.assert_array_equal(dst.cpu().numpy().view(np.uint32), want.view(np.uint32))
    plan.set_noise_profile(0.002 ** 2, 0.01)
    plan.run(src, dst, sigma)
    cuda.cuda.synchronize()
    np.testing.assert_array_equal(dst.cpu().numpy().view(np.uint32), got.view(np.uint32))


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_mixed_precision_against_fp64(cuda, oracle_lib, precision):
    clean, noisy, _ = synth.workload("rgb128x96", seed=8)
    beta, lam = 0.002 ** 2, 0.01
    _, got = _run_plan(cuda, noisy, precision=precision, profile=(beta, lam))
    o64 = oracle_lib.bm3d_f64_profile(noisy, beta, lam)
    o32 = oracle_lib.bm3d_f32_profile(noisy, beta, lam)
    assert synth.psnr(got, o64) > 40.0
    assert abs(synth.psnr(got, clean) - synth.psnr(o32, clean)) <= 0.1


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_per_frame_profiles(cuda, oracle_lib, use_graph):
    """Uneven frames (one smaller than a patch) with their own profiles: every
    frame bit-identical to the CPU restatement with its profile, eager and over
    graph replays with new profiles; then back to the scalar sigmas."""
    torch = cuda
    sizes = [(64, 48), (40, 36), (6, 7), (37, 29)]
    frames = [synth.workload(f"rgb{w}x{h}", seed=20 + i)[1] if min(w, h) >= 8 else
              np.random.default_rng(i).uniform(0, 1, (3, h, w)).astype(np.float32)
              for i, (w, h) in enumerate(sizes)]
    srcs = [torch.from_numpy(f).cuda() for f in frames]
    b = engine.Batch(sizes, 3, "fp32", 0, use_graph=use_graph)
    for rep in range(3):
        betas = np.array([[1e-5 * (1 + f + rep)] * 3 for f in range(len(sizes))], np.float32)
        lams = np.array([[0.01, 0.012, 0.008]] * len(sizes), np.float32) * (1 + 0.5 * rep)
        b.set_noise_profiles(betas, lams)
        outs = [torch.empty_like(t) for t in srcs]
        print(f"batch profiles rep {rep} graph {use_graph}", flush=True)
        b.run(srcs, outs, [0.05] * len(sizes))
        torch.cuda.synchronize()
        b.check()
        for f, (o, fr) in enumerate(zip(outs, frames)):
            want = oracle_lib.bm3d_f32_profile(fr, betas[f], lams[f])
            assert np.array_equal(o.cpu().numpy().view(np.uint32), want.view(np.uint32)), (rep, f)
        _ = (srcs[0] * 2.0).sum().item()  # eager work between replays
    b.set_noise_profiles(None)
    outs = [torch.empty_like(t) for t in srcs]
    b.run(srcs, outs, [0.05] * len(sizes))
    torch.cuda.synchronize()
    for o, fr in zip(outs, frames):
        want = oracle_lib.bm3d_f32(fr, np.float32(0.05))
        assert np.array_equal(o.cpu().numpy().view(np.uint32), want.view(np.uint32))
    b.close()


def test_pilot_and_bad_profiles_are_refused(cuda):
    from cpipe_amd import pilot
    plan = engine.Plan(64, 48, 3)
    with pytest.raises(engine.EngineError) as e:
        plan.set_noise_profile([1e-4, 1e-4], [0.01, 0.01])  # two pairs for three channels
    assert e.value.status == host.CPIPE_BAD_INDEX
    with pytest.raises(engine.EngineError) as e:
        plan.set_noise_profile(float("nan"), 0.01)
    assert e.value.status == host.CPIPE_BAD_INDEX
    net = pilot.Pilot(seed=1)
    plan.set_noise_profile(1e-4, 0.01)
    with pytest.raises(engine.EngineError) as e:
        plan.set_pilot(net)
    assert e.value.status == host.CPIPE_UNSUPPORTED
    plan.set_noise_profile(None)
    plan.set_pilot(net)
    with pytest.raises(engine.EngineError) as e:
        plan.set_noise_profile(1e-4, 0.01)
    assert e.value.status == host.CPIPE_UNSUPPORTED


def _node_rgba(bits, pairs, device=None):
    h, w = bits.shape[:2]
    b = host.Buffer(host.FORMAT_RGBA16F, w, h, device=device)
    b.write(bits)
    b.add_step("color_matrix")
    b.set_noise_profile(pairs)
    return b


def test_node_noise_map_profile(cuda, oracle_lib):
    _, noisy, sigma = synth.workload("rgb64x48", seed=12)
    rgba = np.ones((48, 64, 4), np.float32)
    rgba[..., :3] = np.moveaxis(noisy, 0, -1)
    bits = _f16(rgba)
    pairs = [(0.002 ** 2, 0.01), (0.0015 ** 2, 0.008), (0.0025 ** 2, 0.012)]
    planes = np.moveaxis(_from_f16(bits)[..., :3], -1, 0).copy()
    want = oracle_lib.bm3d_f32_profile(planes, [p[0] for p in pairs], [p[1] for p in pairs])
    d = host.find_node()
    for dev in (None, 0):
        inp = _node_rgba(bits, pairs, device=dev)
        out = host.Buffer(host.FORMAT_RGBA16F, 64, 48, device=dev)
        print(f"node noise_map=profile device={dev}", flush=True)
        rc = host.run_node(d, [inp], [out], {"mode": "bm3d", "noise_map": "profile"})
        assert rc == host.CPIPE_OK, host.host_log()
        got = out.read()
        np.testing.assert_array_equal(got[..., :3], _f16(np.moveaxis(want, 0, -1)))
        np.testing.assert_array_equal(got[..., 3], bits[..., 3])
        assert out.applied_steps == ["color_matrix", "denoise.bm3d"]
    # an explicit sigma wins over the map
    inp = _node_rgba(bits, pairs)
    out = host.Buffer(host.FORMAT_RGBA16F, 64, 48)
    rc = host.run_node(d, [inp], [out], {"mode": "bm3d", "noise_map": "profile", "sigma": 0.04})
    assert rc == host.CPIPE_OK, host.host_log()
    scal = oracle_lib.bm3d_f32(planes, np.float32(0.04))
    np.testing.assert_array_equal(out.read()[..., :3], _f16(np.moveaxis(scal, 0, -1)))


def test_node_grey_planes_with_one_pair_per_cfa_channel(cuda, oracle_lib):
    """A grey plane with a 4-pair (CFA) profile: the node uses the pairs' mean."""
    _, noisy, _ = synth.workload("gray70x52", seed=13)
    pairs = [(1e-5, 0.01), (2e-5, 0.02), (1e-5, 0.015), (3e-5, 0.01)]
    inp = host.Buffer(host.FORMAT_R32F, 70, 52, row_bytes=76 * 4)
    inp.write(noisy[0])
    inp.add_step("color_matrix")
    inp.set_noise_profile(pairs)
    out = host.Buffer(host.FORMAT_R32F, 70, 52)
    rc = host.run_node(host.find_node(), [inp], [out], {"mode": "bm3d", "noise_map": "profile"})
    assert rc == host.CPIPE_OK, host.host_log()
    b = np.float32(np.float32(sum(np.float32(p[0]) for p in pairs)) / np.float32(4))
    lam = np.float32(np.float32(sum(np.float32(p[1]) for p in pairs)) / np.float32(4))
    want = oracle_lib.bm3d_f32_profile(noisy, b, lam)
    np.testing.assert_array_equal(out.read().view(np.uint32), want[0].view(np.uint32))


def test_node_reads_a_dng_noise_profile(cuda, oracle_lib):
    """DNG ingest end to end: the NoiseProfile of a committed DNG fixture
    (three per-plane [S, O] pairs) parsed by cpipe_dng_noise_profile, handed to
    the host's calibration view as (beta, lambda), drives both the node's
    metadata sigma (bm3d.cpp:41-54, no sigma param) and its noise_map=profile run."""
    import os
    from cpipe_amd import dng
    pairs = dng.noise_profile(os.path.join(os.path.dirname(__file__), "golden", "dng", "le_ifd0.dng"))
    assert len(pairs) == 3
    _, noisy, _ = synth.workload("rgb64x48", seed=14)
    rgba = np.ones((48, 64, 4), np.float32)
    rgba[..., :3] = np.moveaxis(noisy, 0, -1)
    bits = _f16(rgba)
    planes = np.moveaxis(_from_f16(bits)[..., :3], -1, 0).copy()
    d = host.find_node()
    print("node with DNG pairs: metadata sigma, then noise_map=profile", flush=True)
    inp = _node_rgba(bits, pairs)
    out = host.Buffer(host.FORMAT_RGBA16F, 64, 48)
    assert host.run_node(d, [inp], [out], {"mode": "bm3d"}) == host.CPIPE_OK, host.host_log()
    sig = np.float32(synth.sigma_from_noise_profile(pairs))
    want = oracle_lib.bm3d_f32(planes, sig)
    np.testing.assert_array_equal(out.read()[..., :3], _f16(np.moveaxis(want, 0, -1)))
    inp = _node_rgba(bits, pairs)
    out = host.Buffer(host.FORMAT_RGBA16F, 64, 48)
    rc = host.run_node(d, [inp], [out], {"mode": "bm3d", "noise_map": "profile"})
    assert rc == host.CPIPE_OK, host.host_log()
    want = oracle_lib.bm3d_f32_profile(planes, [p[0] for p in pairs], [p[1] for p in pairs])
    np.testing.assert_array_equal(out.read()[..., :3], _f16(np.moveaxis(want, 0, -1)))
