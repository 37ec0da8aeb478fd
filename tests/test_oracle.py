"""Pins for the CPU oracle (runs without a GPU).

Each check is chosen so a plausible mistake fails it: known answers from the
reference's own tests, library routines that compute the same thing
independently (numpy's Kaiser window and float16 casts), closed forms
(orthonormality), brute force on tiny inputs, and special cases.
"""
import os

import numpy as np
import pytest

from cpipe_amd import exr, synth

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def _from_f16(b):
    return np.asarray(b, np.uint16).view(np.float16).astype(np.float32)


# ---------------------------------------------------------------- f16 codecs
def test_f16_decode_exhaustive(oracle_lib):
    bits = np.arange(65536, dtype=np.uint16)
    got = oracle_lib.f16_bits_to_f32(bits)
    want = bits.view(np.float16).astype(np.float32)
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    assert same.all()


def test_f16_encode_matches_numpy_rne(oracle_lib):
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.uniform(-70000, 70000, 4000), rng.uniform(-1, 1, 4000) * 2.0 ** rng.integers(-30, 5, 4000),
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                  np.inf, -np.inf, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11])]).astype(np.float32)
    got = oracle_lib.f32_to_f16_bits(x)
    np.testing.assert_array_equal(got, _f16(x))


# ------------------------------------------------- reference-compatible step
def _ref_test_images():
    """tests/unit/test_node_denoise_bm3d.cpp:38-50 (16x16 ramp + checkerboard)."""
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


def _rmse_rgb(a, b):
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    return float(np.sqrt(np.mean(d * d)))


def test_compat_reference_unit_test_known_answers(oracle_lib):
    clean, noisy = _ref_test_images()
    bits = _f16(noisy)
    out = _from_f16(oracle_lib.compat_rgba16f(bits, 0.03))
    noisy_rmse = _rmse_rgb(noisy, clean)
    den = _rmse_rgb(out, clean)
    assert noisy_rmse == pytest.approx(0.050832, abs=5e-7)
    assert den == pytest.approx(0.003606, abs=5e-7)
    assert den <= noisy_rmse * 0.20  # test_node_denoise_bm3d.cpp:165
    assert abs(out[0, 0, 3] - 1.0) <= 0.001  # :166
    over = _from_f16(oracle_lib.compat_rgba16f(bits, 0.09))
    assert _rmse_rgb(over, clean) <= den  # :180


def test_compat_params_known_answer(oracle_lib):
    """test_node_denoise_bm3d_params.cpp:22-31 -- 2x2 image, sigma 0 vs 0.2."""
    img = np.array([[0.1, 0.1, 0.1, 1], [0.9, 0.1, 0.1, 1], [0.1, 0.1, 0.1, 1], [0.9, 0.1, 0.1, 1]],
                   np.float32).reshape(2, 2, 4)
    bits = _f16(img)
    off = _from_f16(oracle_lib.compat_rgba16f(bits, 0.0))
    strong = _from_f16(oracle_lib.compat_rgba16f(bits, 0.2))
    delta = float(np.max(np.abs(off - strong)))
    assert delta > 0.1
    assert delta == pytest.approx(0.40002, abs=1e-4)


def test_compat_golden_fixture(oracle_lib):
    """Regenerated golden pair tests/golden/denoise.bm3d/{in,out}.exr (the reference's
    EXRs are LFS stubs; FLOAT pixels, written into an RGBA16F buffer as the gate does)."""
    inp = _f16(exr.read(os.path.join(GOLDEN, "denoise.bm3d", "in.exr")))
    want = _f16(exr.read(os.path.join(GOLDEN, "denoise.bm3d", "out.exr")))
    got = oracle_lib.compat_rgba16f(inp, 0.03)
    np.testing.assert_array_equal(got, want)
    px = _from_f16(got[0, 0])
    np.testing.assert_allclose(px, [0.093018, 0.094482, 0.093018, 1.0], atol=1e-6)
    # the reference gate: PSNR >= 40 dB against its self-reference (here: exact)
    assert synth.psnr(_from_f16(got)[..., :3], _from_f16(want)[..., :3]) >= 40.0


def test_compat_stage2_identity(oracle_lib):
    rng = np.random.default_rng(3)
    bits = _f16(rng.uniform(0, 2, size=(5, 7, 4)))
    np.testing.assert_array_equal(oracle_lib.compat_rgba16f(bits, 0.1, stage=2), bits)


def test_compat_is_independent_numpy_restatement(oracle_lib):
    """A vectorised numpy restatement of bm3d_step1_generator.cpp:42-56."""
    rng = np.random.default_rng(9)
    for (h, w) in [(6, 9), (3, 1), (11, 4)]:
        img = rng.uniform(-0.3, 1.2, size=(h, w, 4)).astype(np.float32)
        bits = _f16(img)
        v = _from_f16(bits)
        for sigma in (0.0, 0.015, 0.03, 0.5):
            s = np.float32(min(1.0, max(0.0, np.float32(sigma) / np.float32(0.03))))
            ys = np.arange(h)
            xs = np.arange(w)
            by, bx = (ys // 2) * 2, (xs // 2) * 2
            y1, x1 = np.minimum(by + 1, h - 1), np.minimum(bx + 1, w - 1)
            basic = (((v[by][:, bx] + v[by][:, x1]) + v[y1][:, bx]) + v[y1][:, x1]) * np.float32(0.25)
            den = v * (np.float32(1) - s) + basic * s
            want = np.maximum(np.float32(0), den)
            want[..., 3] = v[..., 3]
            np.testing.assert_array_equal(oracle_lib.compat_rgba16f(bits, sigma), _f16(want))


# ------------------------------------------------------------ BM3D pieces
def test_dct_matrix_orthonormal(oracle_lib):
    c = oracle_lib.dct_matrix()
    np.testing.assert_allclose(c @ c.T, np.eye(8), atol=1e-12)
    # DCT-II: row u is a cosine of frequency u; row 0 is constant
    np.testing.assert_allclose(c[0], np.full(8, np.sqrt(1 / 8)), atol=1e-15)
    x = np.random.default_rng(1).normal(size=8)
    # matches scipy's orthonormal DCT-II
    from scipy.fft import dct
    np.testing.assert_allclose(c @ x, dct(x, norm="ortho"), atol=1e-12)


def test_kaiser_matches_numpy(oracle_lib):
    np.testing.assert_allclose(oracle_lib.kaiser(), np.kaiser(8, 2.0), rtol=1e-13, atol=1e-15)


def test_reference_positions(oracle_lib):
    assert list(oracle_lib.ref_positions(8)) == [0]
    assert list(oracle_lib.ref_positions(11)) == [0, 3]
    assert list(oracle_lib.ref_positions(12)) == [0, 3, 4]
    assert list(oracle_lib.ref_positions(20))[-2:] == [9, 12]
    p = oracle_lib.ref_positions(4080)
    assert len(p) == 1359 and p[-1] == 4072 and np.all(np.diff(p) <= 3)
    assert len(oracle_lib.ref_positions(3072)) == 1023


def _brute_force_bm(plane, tau_ssd):
    """Float64 brute force: all candidates, keep SSD < tau, ref first, sort by (ssd, raster)."""
    h, w = plane.shape
    from oracle.oracle import ref_positions
    res_pos, res_cnt = [], []
    for ry in ref_positions(h):
        for rx in ref_positions(w):
            refp = plane[ry:ry + 8, rx:rx + 8].astype(np.float64)
            cands = []
            for qy in range(max(0, ry - 19), min(h - 8, ry + 19) + 1):
                for qx in range(max(0, rx - 19), min(w - 8, rx + 19) + 1):
                    if qx == rx and qy == ry:
                        continue
                    d = float(np.sum((refp - plane[qy:qy + 8, qx:qx + 8]) ** 2))
                    if d < tau_ssd:
                        cands.append((d, qy, qx))
            cands.sort()
            sel = [(ry, rx)] + [(qy, qx) for _, qy, qx in cands[:15]]
            res_cnt.append(len(sel))
            res_pos.append([(qy << 16) | qx for qy, qx in sel] + [0] * (16 - len(sel)))
    return np.array(res_pos, np.uint32), np.array(res_cnt, np.uint8)


@pytest.mark.parametrize("shape,seed", [((20, 23), 0), ((31, 17), 1)])
def test_block_matching_brute_force(oracle_lib, shape, seed):
    rng = np.random.default_rng(seed)
    plane = rng.uniform(0, 1, size=shape).astype(np.float32)
    # quantise so that exact ties occur and the tie rule is exercised
    plane = (np.round(plane * 4) / 4).astype(np.float32)
    tau = 64.0 * 0.2
    pos_bf, cnt_bf = _brute_force_bm(plane, tau)
    pos64, cnt64 = oracle_lib.block_match_f64(plane, tau)
    np.testing.assert_array_equal(cnt64, cnt_bf)
    np.testing.assert_array_equal(pos64, pos_bf)
    # FP32 mirror: on quarter-step values every SSD is exact in FP32 too
    pos32, cnt32 = oracle_lib.block_match_f32(plane, np.float32(tau))
    np.testing.assert_array_equal(cnt32, cnt_bf)
    np.testing.assert_array_equal(pos32, pos_bf)


def _brute_force_bm_keys(plane, tau_ssd, keyed=19):
    """Brute force of the FP32 rank rule: keep SSD < tau and the reference first;
    of the other candidates keep the `keyed` smallest keys (key = the FP32 SSD's
    bits with the low 11 replaced by (dy + 19) * 39 + (dx + 19): the SSD rounded
    down to 12 mantissa bits, then raster position), and of those the first 15 by
    exact (SSD, raster position). keyed=15 with no second stage gives the key
    order alone. The SSDs are computed exactly in float64."""
    h, w = plane.shape
    from oracle.oracle import ref_positions
    res_pos, res_cnt = [], []
    for ry in ref_positions(h):
        for rx in ref_positions(w):
            refp = plane[ry:ry + 8, rx:rx + 8].astype(np.float64)
            cands = []
            for qy in range(max(0, ry - 19), min(h - 8, ry + 19) + 1):
                for qx in range(max(0, rx - 19), min(w - 8, rx + 19) + 1):
                    if qx == rx and qy == ry:
                        continue
                    d = float(np.sum((refp - plane[qy:qy + 8, qx:qx + 8]) ** 2))
                    if d < tau_ssd:
                        bits = int(np.array(d, np.float32).view(np.uint32))
                        key = (bits & ~0x7FF) | ((qy - ry + 19) * 39 + (qx - rx + 19))
                        cands.append((key, d, qy, qx))
            cands.sort()
            kept = sorted(cands[:keyed], key=lambda c: (c[1], c[2], c[3]))
            sel = [(ry, rx)] + [(qy, qx) for _, _, qy, qx in kept[:15]]
            res_cnt.append(len(sel))
            res_pos.append([(qy << 16) | qx for qy, qx in sel] + [0] * (16 - len(sel)))
    return np.array(res_pos, np.uint32), np.array(res_cnt, np.uint8)


@pytest.mark.parametrize("shape,seed", [((40, 44), 3), ((52, 46), 4)])
def test_block_matching_rank_keys(oracle_lib, shape, seed):
    """On a flat patch of 1/1024-step values every SSD is exact in FP32 whatever the
    summation order (multiples of 2^-20 below 1), and the SSDs near 0.1 lie 16 to
    a 12-mantissa-bit bucket: the key order alone differs from the exact order on
    many refs, so the second stage (exact re-rank of the 19 kept candidates) is
    exercised, and the FP32 lists must equal the two-stage brute force. The
    two-stage result differs from a plain (SSD, raster) ranking only where five
    later candidates with larger SSDs share a member's bucket, here on a few refs
    at most."""
    rng = np.random.default_rng(seed)
    plane = (np.round(rng.uniform(0.45, 0.55, size=shape) * 1024) / 1024).astype(np.float32)
    tau = 64.0 * 0.05
    pos_k, cnt_k = _brute_force_bm_keys(plane, tau)
    pos32, cnt32 = oracle_lib.block_match_f32(plane, np.float32(tau))
    np.testing.assert_array_equal(cnt32, cnt_k)
    np.testing.assert_array_equal(pos32, pos_k)
    pos_keys_only, _ = _brute_force_bm_keys(plane, tau, keyed=15)
    assert np.any(np.any(pos_keys_only != pos_k, axis=1))
    pos_exact, _ = _brute_force_bm(plane, tau)
    differ = np.any(pos_exact != pos_k, axis=1)
    print(f"refs whose list differs from the plain (SSD, raster) ranking: {differ.sum()} of {len(differ)}")
    assert differ.mean() <= 0.05


def test_sigma_zero_is_identity(oracle_lib):
    clean, noisy, _ = synth.workload("rgb48x40", seed=2)
    out = oracle_lib.bm3d_f32(noisy, 0.0)
    assert synth.max_abs(out, noisy) <= 2e-5
    out64 = oracle_lib.bm3d_f64(noisy.astype(np.float64), 0.0)
    assert synth.max_abs(out64, noisy) <= 1e-12


def test_constant_image_closed_form(oracle_lib):
    """Constant image: every group has 16 identical members, so step 1 (hard threshold)
    returns the constant and step 2 scales it by the Wiener gain of the 3D DC
    coefficient b = v * 8 * sqrt(16): out = v * b^2 / (b^2 + sigma^2)."""
    v, sigma = 0.25, 0.1
    img = np.full((1, 33, 29), v, np.float32)
    b = v * 8 * 4
    want = v * b * b / (b * b + sigma * sigma)
    out64, basic64 = oracle_lib.bm3d_f64(img.astype(np.float64), sigma, want_basic=True)
    assert synth.max_abs(basic64, img) <= 1e-12
    np.testing.assert_allclose(out64, want, rtol=1e-12)
    out32, basic32 = oracle_lib.bm3d_f32(img, sigma, want_basic=True)
    assert synth.max_abs(basic32, img) <= 1e-6
    np.testing.assert_allclose(out32, want, rtol=1e-5)


def test_small_image_is_copied(oracle_lib):
    img = np.random.default_rng(0).uniform(size=(3, 5, 7)).astype(np.float32)
    np.testing.assert_array_equal(oracle_lib.bm3d_f32(img, 0.1), img)


def test_single_group_output_within_input_range(oracle_lib):
    # an 8x8 image has exactly one group (the reference alone): output stays in the input's range
    img = np.random.default_rng(4).uniform(0.2, 0.6, size=(1, 8, 8)).astype(np.float32)
    out = oracle_lib.bm3d_f64(img.astype(np.float64), 0.05)
    assert out.min() >= img.min() - 0.3 and out.max() <= img.max() + 0.3
    assert np.isfinite(out).all()


def test_fp32_mirror_agrees_with_fp64(oracle_lib):
    clean, noisy, sigma = synth.workload("gray96x96", seed=7)
    o32 = oracle_lib.bm3d_f32(noisy, sigma)
    o64 = oracle_lib.bm3d_f64(noisy, sigma)
    assert synth.psnr(o32, o64) > 60.0
    assert synth.psnr(o64, clean) > synth.psnr(noisy, clean) + 5.0


def test_fp32_mirror_is_thread_count_invariant(oracle_lib):
    clean, noisy, sigma = synth.workload("rgb64x64", seed=4)
    a = oracle_lib.bm3d_f32(noisy, sigma, nthreads=1)
    b = oracle_lib.bm3d_f32(noisy, sigma, nthreads=7)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_wiener_with_clean_pilot_beats_hard_threshold(oracle_lib):
    """Step 2 (Wiener on the basic estimate) improves on step 1 for Gaussian noise."""
    clean, noisy, sigma = synth.workload("gray128x128", seed=0)
    out, basic = oracle_lib.bm3d_f64(noisy.astype(np.float64), sigma, want_basic=True)
    assert synth.psnr(out, clean) > synth.psnr(basic, clean)


def test_dct8_f32_butterfly_matches_the_dense_transform(oracle_lib):
    """The FP32 butterfly (the order the FP32 restatement and the GPU exact
    kernels share) against the FP64 dense DCT-II matrix: forward, inverse and
    round trip, to FP32 rounding; a dropped term, a wrong sign or a swapped
    index in any output is off by O(1)."""
    c = oracle_lib.dct_matrix()
    rng = np.random.default_rng(0)
    for _ in range(200):
        x = rng.uniform(-1, 1, 8).astype(np.float32)
        X = oracle_lib.dct8_f32(x)
        np.testing.assert_allclose(X, c @ x.astype(np.float64), atol=2e-6, rtol=0)
        xi = oracle_lib.dct8_f32(X, inverse=True)
        np.testing.assert_allclose(xi, c.T @ X.astype(np.float64), atol=2e-6, rtol=0)
        np.testing.assert_allclose(xi, x, atol=4e-6, rtol=0)
    # unit vectors: each output column of the matrix separately
    for n in range(8):
        e = np.zeros(8, np.float32)
        e[n] = 1.0
        np.testing.assert_allclose(oracle_lib.dct8_f32(e), c[:, n], atol=1e-7, rtol=0)
        np.testing.assert_allclose(oracle_lib.dct8_f32(e, inverse=True), c[n, :], atol=1e-7, rtol=0)
    # constant input: only the DC survives (exact zeros by the butterfly's symmetry)
    X = oracle_lib.dct8_f32(np.full(8, 0.25, np.float32))
    assert X[0] == np.float32(np.float32(2.0) * np.float32(1 / np.sqrt(8)))  # (e0 + e1) Q0, e = 1 + 1
    assert np.all(X[1:] == 0.0)
