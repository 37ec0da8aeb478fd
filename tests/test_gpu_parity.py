"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar: integer/index work (match lists) and the reference-compatible f16 filter
are bit-exact; the FP32 BM3D output is required to be within max-abs 1e-5 of
the CPU FP32 restatement (it follows the same operation order, so it is
expected to be bit-identical) and PSNR >= 50 dB / max-abs <= 1e-3 against it,
plus close to the FP64 textbook oracle.
"""
import numpy as np
import pytest

from cpipe_amd import engine, synth

pytestmark = pytest.mark.gpu


def _run_gpu(torch, noisy, sigma, want_basic=False, want_matches=False):
    c, h, w = noisy.shape
    src = torch.from_numpy(np.ascontiguousarray(noisy)).to("cuda:0")
    out = torch.empty_like(src)
    plan = engine.Plan(w, h, c, "fp32", 0)
    plan.run(src, out, sigma)
    res = {"out": None}
    if want_basic:
        b = torch.empty_like(src)
        plan.basic(b)
        res["basic"] = b
    if want_matches:
        res["m1"] = plan.matches(1)
        res["m2"] = plan.matches(2)
    torch.cuda.synchronize()
    plan.check()
    res["out"] = out.cpu().numpy()
    if want_basic:
        res["basic"] = res["basic"].cpu().numpy()
    if want_matches:
        res["m1"] = tuple(t.cpu().numpy() for t in res["m1"])
        res["m2"] = tuple(t.cpu().numpy() for t in res["m2"])
    plan.close()
    return res


# Block matching pairs ref columns 2k, 2k+1 of a 16-column tile. A tile whose
# forced last ref column (W-8 not a multiple of the stride) is the second of a
# pair, or whose forced last ref row breaks the 3-row spacing, sums every ref's
# patch rows itself; the other tiles share chunks between paired refs and
# between refs 3 rows apart.
CASES = [
    ("gray64x64", 0),   # one tile, forced last column second of its pair, forced last row
    ("gray61x47", 1),   # forced last column first of its pair: shared path
    ("rgb96x80", 2),
    ("rgb130x70", 3),   # three tile columns: two shared, the last per-ref
    ("gray300x260", 4),  # several tiles in both directions
]


@pytest.mark.parametrize("name,seed", CASES)
def test_bm3d_fp32_matches_cpu_fp32(cuda, oracle_lib, name, seed):
    torch = cuda
    clean, noisy, sigma = synth.workload(name, seed=seed)
    print(f"case {name} seed {seed} sigma {sigma}", flush=True)
    g = _run_gpu(torch, noisy, sigma, want_basic=True, want_matches=True)
    out, basic, (p1, c1, p2, c2) = oracle_lib.bm3d_f32(noisy, sigma, nthreads=8, want_basic=True,
                                                           want_matches=True)
    # match lists: bit-exact
    np.testing.assert_array_equal(g["m1"][1], c1)
    np.testing.assert_array_equal(g["m1"][0].view(np.uint32), p1)
    np.testing.assert_array_equal(g["m2"][1], c2)
    np.testing.assert_array_equal(g["m2"][0].view(np.uint32), p2)
    # basic and final estimates
    assert synth.max_abs(g["basic"], basic) <= 1e-5
    err = synth.max_abs(g["out"], out)
    assert err <= 1e-5, err
    assert synth.psnr(g["out"], out) >= 50.0
    # against the FP64 textbook oracle
    o64 = oracle_lib.bm3d_f64(noisy, sigma)
    assert synth.psnr(g["out"], o64) >= 50.0
    assert synth.psnr(g["out"], clean) > synth.psnr(noisy, clean) + 3.0


def test_bm3d_bitwise_and_deterministic(cuda, oracle_lib):
    torch = cuda
    clean, noisy, sigma = synth.workload("rgb96x80", seed=11)
    a = _run_gpu(torch, noisy, sigma)["out"]
    b = _run_gpu(torch, noisy, sigma)["out"]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    ref = oracle_lib.bm3d_f32(noisy, sigma)
    n_diff = int(np.count_nonzero(a.view(np.uint32) != ref.view(np.uint32)))
    print(f"bitwise differences vs CPU FP32: {n_diff} of {a.size}")
    assert n_diff == 0


@pytest.mark.parametrize("w,h", [(96, 90), (1600, 1400)])
def test_near_ties_bit_exact(cuda, oracle_lib, w, h):
    """A flat patch of 1/1024-step values: SSDs near 0.1 lie many to a 2^-12 bucket,
    so most refs keep slots whose rank keys share a bucket and block matching's
    second stage (exact SSDs for those slots) decides the order. Match lists and
    outputs must equal the CPU FP32 restatement bit for bit; 96x90 runs one ref
    per lane (4 workgroup items), 1600x1400 the paired path (1156 items)."""
    torch = cuda
    rng = np.random.default_rng(w)
    noisy = (np.round(rng.uniform(0.45, 0.55, size=(1, h, w)) * 1024) / 1024).astype(np.float32)
    sigma = 0.02
    print(f"near ties {w}x{h} sigma {sigma}", flush=True)
    g = _run_gpu(torch, noisy, sigma, want_matches=True)
    out, _, (p1, c1, p2, c2) = oracle_lib.bm3d_f32(noisy, sigma, nthreads=8, want_basic=True,
                                                       want_matches=True)
    np.testing.assert_array_equal(g["m1"][1], c1)
    np.testing.assert_array_equal(g["m1"][0].view(np.uint32), p1)
    np.testing.assert_array_equal(g["m2"][1], c2)
    np.testing.assert_array_equal(g["m2"][0].view(np.uint32), p2)
    assert np.array_equal(g["out"].view(np.uint32), out.view(np.uint32))
    # the first stage alone would order some lists differently: recompute the
    # step-1 keys of the kept members and check that some list is not key-sorted
    plane = noisy[0].astype(np.float64)
    unsorted = 0
    for r in range(min(len(c1), 400)):
        n = int(c1[r])
        ry, rx = int(p1[r, 0]) >> 16, int(p1[r, 0]) & 0xFFFF
        keys = []
        for m in range(1, n):
            qy, qx = int(p1[r, m]) >> 16, int(p1[r, m]) & 0xFFFF
            d = float(np.sum((plane[ry:ry + 8, rx:rx + 8] - plane[qy:qy + 8, qx:qx + 8]) ** 2))
            bits = int(np.array(d, np.float32).view(np.uint32))
            keys.append((bits & ~0x7FF) | ((qy - ry + 19) * 39 + (qx - rx + 19)))
        unsorted += int(keys != sorted(keys))
    print(f"lists not in key order (of the first 400): {unsorted}")
    assert unsorted > 0


@pytest.mark.parametrize("w,h,c", [(7, 9, 1), (5, 5, 3), (8, 8, 1), (8, 8, 3), (9, 8, 3), (40, 8, 1)])
def test_small_images(cuda, oracle_lib, w, h, c):
    torch = cuda
    rng = np.random.default_rng(w * 100 + h)
    noisy = rng.uniform(0, 1, size=(c, h, w)).astype(np.float32)
    print(f"small image {w}x{h}x{c}", flush=True)
    g = _run_gpu(torch, noisy, 0.05)["out"]
    ref = oracle_lib.bm3d_f32(noisy, 0.05)
    assert synth.max_abs(g, ref) <= 1e-5


def _random_shapes(n, seed):
    """Seeded shapes: 1 or 3 channels, sides 8..700 (some below 40), sigma 0..0.12."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        c = int(rng.choice([1, 3]))
        w = int(rng.integers(8, 700)) if rng.random() >= 0.2 else int(rng.integers(1, 40))
        h = int(rng.integers(8, 500)) if rng.random() >= 0.2 else int(rng.integers(1, 40))
        out.append((i, c, w, h, float(rng.uniform(0.0, 0.12))))
    return out


@pytest.mark.parametrize("i,c,w,h,sigma", _random_shapes(16, 7))
def test_random_shapes_bit_identical(cuda, oracle_lib, i, c, w, h, sigma):
    """Border tiles, clamped last refs and partial tiles at arbitrary frame sizes."""
    print(f"random shape {i}: {c}x{h}x{w} sigma {sigma:.4f}", flush=True)
    clean = synth.clean_scene(w, h, c, seed=i)
    noisy = synth.add_gaussian(clean, sigma, seed=1000 + i).astype(np.float32)
    g = _run_gpu(cuda, noisy, sigma)["out"]
    ref = oracle_lib.bm3d_f32(noisy, sigma, nthreads=8)
    np.testing.assert_array_equal(g.view(np.uint32), ref.view(np.uint32))


def test_sigma_zero_is_identity(cuda):
    torch = cuda
    clean, noisy, _ = synth.workload("rgb64x48", seed=5)
    g = _run_gpu(torch, noisy, 0.0)["out"]
    assert synth.max_abs(g, noisy) <= 2e-5


def test_constant_image_closed_form(cuda):
    """Grey constant RGB: opponent Y = v, U = V = 0; step 2 scales Y by the Wiener
    gain of the group DC b = 32 v against sigma_Y = sigma / sqrt(3)."""
    torch = cuda
    v, sigma = 0.37, 0.05
    img = np.full((3, 40, 56), v, np.float32)
    g = _run_gpu(torch, img, sigma)["out"]
    b2 = (32 * v) ** 2
    want = v * b2 / (b2 + sigma * sigma / 3.0)
    np.testing.assert_allclose(g, want, rtol=1e-5)


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


@pytest.mark.parametrize("w,h", [(16, 16), (2, 2), (37, 23), (1, 1), (1, 5)])
@pytest.mark.parametrize("sigma", [0.0, 0.01, 0.03, 0.09, 0.2])
@pytest.mark.parametrize("stage", [1, 2])
def test_compat_bitwise(cuda, oracle_lib, w, h, sigma, stage):
    torch = cuda
    rng = np.random.default_rng(w * 31 + h)
    img = rng.uniform(-0.2, 1.5, size=(h, w, 4)).astype(np.float32)
    bits = _f16(img)
    want = oracle_lib.compat_rgba16f(bits, sigma, stage)
    src = torch.from_numpy(bits.view(np.int16)).to("cuda:0")
    dst = torch.empty_like(src)
    engine.compat_rgba16f(src, dst, sigma, stage)
    torch.cuda.synchronize()
    got = dst.cpu().numpy().view(np.uint16)
    np.testing.assert_array_equal(got, want)


def test_bm3d_rgba16f_path(cuda, oracle_lib):
    torch = cuda
    clean, noisy, sigma = synth.workload("rgb72x56", seed=8)
    h, w = noisy.shape[1:]
    rgba = np.ones((h, w, 4), np.float32)
    rgba[..., :3] = np.moveaxis(noisy, 0, -1)
    rgba[..., 3] = 0.5
    bits = _f16(rgba)
    planes = np.moveaxis(bits.view(np.float16).astype(np.float32)[..., :3], -1, 0).copy()
    want = oracle_lib.bm3d_f32(planes, sigma)
    src = torch.from_numpy(bits.view(np.int16)).to("cuda:0")
    dst = torch.empty_like(src)
    plan = engine.Plan(w, h, 3)
    plan.run_rgba16f(src, dst, sigma)
    torch.cuda.synchronize()
    got = dst.cpu().numpy().view(np.uint16)
    np.testing.assert_array_equal(got[..., 3], bits[..., 3])
    np.testing.assert_array_equal(got[..., :3], _f16(np.moveaxis(want, 0, -1)))
    plan.close()


@pytest.mark.parametrize("profile", [False, True])
def test_full_frame_bit_exact(cuda, oracle_lib, profile):
    """The headline workload itself (4080x3072x3, Poisson-Gaussian gain 0.01 read
    0.002): the exact FP32 output is bit-identical to the CPU FP32 restatement,
    with the one sigma and with the noise-profile map."""
    import os
    torch = cuda
    clean, noisy, sigma = synth.workload("rgb12mp", seed=0)
    c, h, w = noisy.shape
    print(f"rgb12mp full frame, noise map {profile}", flush=True)
    plan = engine.Plan(w, h, c, "fp32", 0)
    if profile:
        plan.set_noise_profile(0.002 ** 2, 0.01)
    src = torch.from_numpy(noisy).cuda()
    dst = torch.empty_like(src)
    plan.run(src, dst, sigma)
    torch.cuda.synchronize()
    plan.check()
    got = dst.cpu().numpy()
    plan.close()
    del src, dst
    nt = min(16, os.cpu_count() or 1)
    if profile:
        want = oracle_lib.bm3d_f32_profile(noisy, 0.002 ** 2, 0.01, nthreads=nt)
    else:
        want = oracle_lib.bm3d_f32(noisy, np.float32(sigma), nthreads=nt)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    assert synth.psnr(got, clean) > synth.psnr(noisy, clean) + 6.0


def test_error_flag_is_cleared_by_every_run(cuda):
    """A run marked as failed makes plan.check() raise; the next run clears the
    flag, so the same plan (e.g. a cached node plan) reports OK again."""
    torch = cuda
    clean, noisy, sigma = synth.workload("rgb64x48", seed=6)
    src = torch.from_numpy(noisy).cuda()
    out = torch.empty_like(src)
    plan = engine.Plan(64, 48, 3, "fp32", 0)
    plan.check()  # before any run
    plan.run(src, out, sigma)
    plan.check()
    plan.debug_flag()
    with pytest.raises(engine.EngineError) as e:
        plan.check()
    assert e.value.status == 7  # CPIPE_INTERNAL_ERROR
    plan.run(src, out, sigma)
    plan.check()
    plan.close()


@pytest.mark.parametrize("case", ["single", "two", "dup16x8", "dup40x8", "dup40x40", "rgb"])
def test_closed_form_images(cuda, oracle_lib, case):
    """The oracle's closed-form pin images (tests/test_oracle_shrinkage.py) run
    through the GPU path: bit-identical to the CPU FP32 restatement, which the
    closed forms pin to FP32 rounding."""
    import test_oracle_shrinkage as cf
    torch = cuda
    if case == "single":
        rng = np.random.default_rng(11)
        img = cf._coeff_patch({(0, 0): 3.0, (0, 1): 0.130, (2, 3): -0.131, (1, 0): 0.3,
                               (3, 3): -0.2, (5, 6): 0.14}, rng)[None].astype(np.float32)
        sigma = cf.SIG1
    elif case == "two":
        img, sigma = cf._two_group_image()[None], cf.SIG2
    elif case == "rgb":
        img = np.random.default_rng(21).uniform(0.1, 0.9, (3, 8, 8)).astype(np.float32)
        sigma = 0.06
    else:
        w, h = (int(v) for v in case[3:].split("x"))
        seed = {"dup16x8": 1, "dup40x8": 2, "dup40x40": 3}[case]
        img, sigma = cf._duplicate_image(w, h, seed, cf.LAMBDA * cf.SIG3)[None], cf.SIG3
    print(f"closed-form image {case} {img.shape}", flush=True)
    g = _run_gpu(torch, img, sigma, want_basic=True)
    out, basic = oracle_lib.bm3d_f32(img, sigma, want_basic=True)
    np.testing.assert_array_equal(g["basic"].view(np.uint32), basic.view(np.uint32))
    np.testing.assert_array_equal(g["out"].view(np.uint32), out.view(np.uint32))
