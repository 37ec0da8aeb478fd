"""Closed-form pins for the BM3D oracle's shrinkage and aggregation (no GPU).

The reference has no BM3D arithmetic; its design prose
(docs/research/07-classic-isp-algorithms.md:548-551 in xndcn/cpipe) fixes the
hard threshold at 2.7 sigma and weights the aggregation by the count of
non-zero coefficients, and the published algorithm (Dabov et al. 2007, sec. 2;
Lebrun, IPOL 2012, eqs. 2-4) fixes the rest: orthonormal 2D DCT-II, orthonormal
Haar along the stack, weights 1/(sigma^2 N_nonzero) after the hard threshold and
1/(sigma^2 sum W^2) after the Wiener step with W = B^2 / (B^2 + sigma^2), an 8x8
Kaiser window (beta 2). The prose's "proportional to count" is read as the
paper's inverse count; test_inverse_count_weight_beats_count shows why.

Every expected value here is computed in float64 with scipy's dctn, numpy's
Kaiser window and a Haar matrix built from its definition, never with the
oracle's own butterflies or loops, on images whose groups are known in advance:
  * one 8x8 patch (one group of one member): step 1 = idctn(HT(dctn(x))), step 2
    = idctn(W * dctn(x));
  * two overlapping one-member groups: the weighted mean of their estimates;
  * exact duplicate patches (periodic images): a group of n identical members
    has the 3D DC sqrt(n) * X and zero detail, so each member comes back as
    idctn(HT(sqrt(n) X) / sqrt(n)) with weight 1/N_nonzero(sqrt(n) X);
  * an RGB patch: the opponent transform with sigma_c = sigma * |row_c|.
Then the same checks run against oracle builds with each of four planted
mistakes (threshold 2.5 sigma, weight N instead of 1/N, unnormalised Haar,
Wiener weight sum W^2 instead of its inverse) and must fail on every one.
"""
import numpy as np
import pytest
from scipy.fft import dctn, idctn

from cpipe_amd import synth

LAMBDA = 2.7
TAU_HT = 2500.0 / 255.0 ** 2   # per pixel, on SSD / 64
TAU_WIE = 400.0 / 255.0 ** 2
KAISER2 = np.outer(np.kaiser(8, 2.0), np.kaiser(8, 2.0))
OPP = np.array([[1 / 3, 1 / 3, 1 / 3], [1 / 2, 0.0, -1 / 2], [1 / 4, -1 / 2, 1 / 4]])


def dct2(x):
    return dctn(x, norm="ortho")


def idct2(x):
    return idctn(x, norm="ortho")


def haar_matrix(n):
    """Orthonormal Haar matrix of size n (power of two), rows = basis vectors:
    H_1 = [1]; H_2m = [H_m (x) (1, 1); I_m (x) (1, -1)] / sqrt(2)."""
    h = np.ones((1, 1))
    while h.shape[0] < n:
        m = h.shape[0]
        h = np.vstack([np.kron(h, [1.0, 1.0]), np.kron(np.eye(m), [1.0, -1.0])]) / np.sqrt(2.0)
    return h


def ref_positions(n):
    """Reference patch positions along an axis: stride 3, last one forced to n - 8."""
    last = n - 8
    p = list(range(0, last + 1, 3))
    if p[-1] != last:
        p.append(last)
    return p


def groups_from_lists(pos, cnt, w, h):
    """Member positions (x, y) of every reference's group: the first floor_pow2(count)
    entries of its match list."""
    out = []
    nx, ny = len(ref_positions(w)), len(ref_positions(h))
    for r in range(nx * ny):
        c = int(cnt[r])
        n = 1 << (c.bit_length() - 1)
        out.append([(int(v) & 0xFFFF, int(v) >> 16) for v in pos[r][:n]])
    return out


def textbook_step(noisy, basic, groups, sig):
    """One BM3D step on one plane from the published algorithm: 3D transform =
    dctn per member then the orthonormal Haar matrix along the stack; step 1 hard
    threshold at 2.7 sigma with weight 1/N_nonzero (the 1/sigma^2 factor is common
    to all groups of the plane and cancels), step 2 Wiener with the basic
    estimate's spectrum and weight 1/sum W^2; Kaiser-weighted aggregation."""
    h, w = noisy.shape
    num = np.zeros((h, w))
    den = np.zeros((h, w))
    for g in groups:
        n = len(g)
        H = haar_matrix(n)
        Z = np.stack([dct2(noisy[y:y + 8, x:x + 8]) for x, y in g])          # n x 8 x 8
        T = np.tensordot(H, Z, axes=1)
        if basic is None:
            keep = np.abs(T) > LAMBDA * sig
            nnz = int(keep.sum())
            T = np.where(keep, T, 0.0)
            wgt = 1.0 / nnz if nnz > 0 else 1.0
        else:
            B = np.tensordot(H, np.stack([dct2(basic[y:y + 8, x:x + 8]) for x, y in g]), axes=1)
            W = B * B / (B * B + sig * sig)
            T = W * T
            sw = float((W * W).sum())
            wgt = 1.0 / sw if sw > 0 else 1.0
        est = np.tensordot(H.T, T, axes=1)
        for (x, y), e in zip(g, est):
            num[y:y + 8, x:x + 8] += wgt * KAISER2 * idct2(e)
            den[y:y + 8, x:x + 8] += wgt * KAISER2
    return num / den


def min_threshold_margin(patches, sig, scale=1.0):
    """Smallest distance of any |scale * dctn(patch)| coefficient to the threshold."""
    return min(float(np.min(np.abs(np.abs(scale * dct2(p)) - LAMBDA * sig))) for p in patches)


def shift_ssd_min(img, positions, tau):
    """Smallest SSD/64 between two distinct patches of one row of refs that are
    not exact duplicates (must exceed tau so that only duplicates match)."""
    best = np.inf
    for a in positions:
        for b in positions:
            if a == b:
                continue
            pa, pb = img[a[1]:a[1] + 8, a[0]:a[0] + 8], img[b[1]:b[1] + 8, b[0]:b[0] + 8]
            d = float(np.sum((pa - pb) ** 2)) / 64.0
            if d > 0.0:
                best = min(best, d)
    return best


# --------------------------------------------------------------------------- cases
# Each case returns (max |oracle - closed form| in FP64, same in FP32).

def _coeff_patch(spec, rng, small=0.08):
    """An 8x8 patch given by its DCT coefficients: spec {(u, v): value}, every
    other coefficient uniform in +-small."""
    X = rng.uniform(-small, small, (8, 8))
    for (u, v), val in spec.items():
        X[u, v] = val
    return idct2(X)


SIG1 = 0.05


def case_single_patch(o):
    """8x8 image = one reference, one member. Coefficients straddle 2.7 sigma:
    0.130 and -0.131 lie between 2.5 and 2.7 sigma (zeroed), 0.14 and more are
    kept, the rest (|c| < 0.08) go."""
    rng = np.random.default_rng(11)
    lt = LAMBDA * SIG1
    x = _coeff_patch({(0, 0): 3.0, (0, 1): 0.130, (2, 3): -0.131, (1, 0): 0.3,
                      (3, 3): -0.2, (5, 6): 0.14}, rng)
    img = x[None].astype(np.float32)
    x64 = img[0].astype(np.float64)  # the FP32-rounded input both oracles see
    assert min_threshold_margin([x64], SIG1) > 1e-3
    o64, b64 = o.bm3d_f64(x64[None], SIG1, want_basic=True)
    o32, b32 = o.bm3d_f32(img, SIG1, want_basic=True)
    X = dct2(x64)
    basic = idct2(np.where(np.abs(X) > lt, X, 0.0))
    B = dct2(basic)
    out = idct2(B * B / (B * B + SIG1 ** 2) * X)
    e64 = max(np.abs(b64[0] - basic).max(), np.abs(o64[0] - out).max())
    e32 = max(np.abs(b32[0] - basic).max(), np.abs(o32[0] - out).max())
    return e64, e32


def case_dc_below_threshold(o):
    """DC not exempt: a mean of 0.01 (DC 0.08 < 2.7 sigma) is removed while one AC
    coefficient above the threshold survives; with every coefficient below the
    threshold the estimate is 0 (N_nonzero = 0 -> weight 1, zero numerator)."""
    rng = np.random.default_rng(12)
    errs64, errs32 = [], []
    for spec in ({(0, 0): 0.08, (2, 1): 0.5}, {(0, 0): 0.08}):
        x = _coeff_patch(spec, rng, small=0.05).astype(np.float32)
        xd = x.astype(np.float64)
        X = dct2(xd)
        basic = idct2(np.where(np.abs(X) > LAMBDA * SIG1, X, 0.0))
        assert min_threshold_margin([xd], SIG1) > 1e-3
        _, b64 = o.bm3d_f64(xd[None], SIG1, want_basic=True)
        _, b32 = o.bm3d_f32(x[None], SIG1, want_basic=True)
        errs64.append(np.abs(b64[0] - basic).max())
        errs32.append(np.abs(b32[0] - basic).max())
        if len(spec) == 1:
            assert np.all(b64 == 0.0) and np.all(b32 == 0.0)
    return max(errs64), max(errs32)


SIG2 = 0.1


def _two_group_image():
    rng = np.random.default_rng(5)
    return rng.uniform(0.0, 1.0, (8, 11)).astype(np.float32)


def case_two_groups(o):
    """8x11 image: references at x = 0 and 3, each its own only match (shifted
    patches are far apart), so each pixel of columns 3..7 is the weighted mean of
    two estimates with weights 1/N_nonzero (step 1) and 1/sum W^2 (step 2)."""
    img = _two_group_image()
    x = img.astype(np.float64)
    cands = [(i, 0) for i in range(4)]
    assert shift_ssd_min(x, cands, TAU_HT) > 3 * TAU_HT
    groups = [[(0, 0)], [(3, 0)]]
    assert min_threshold_margin([x[:, 0:8], x[:, 3:11]], SIG2) > 1e-4
    nnz = [int((np.abs(dct2(x[:, a:a + 8])) > LAMBDA * SIG2).sum()) for a in (0, 3)]
    assert nnz[0] != nnz[1]  # otherwise the weights would not matter
    basic = textbook_step(x, None, groups, SIG2)
    assert shift_ssd_min(basic, cands, TAU_WIE) > 3 * TAU_WIE
    out = textbook_step(x, basic, groups, SIG2)
    o64, b64 = o.bm3d_f64(x[None], SIG2, want_basic=True)
    o32, b32 = o.bm3d_f32(img[None], SIG2, want_basic=True)
    e64 = (np.abs(b64[0] - basic).max(), np.abs(o64[0] - out).max())
    e32 = (np.abs(b32[0] - basic).max(), np.abs(o32[0] - out).max())
    return e64, e32


def _duplicate_image(w, h, seed, lam_sig):
    """A periodic image (period 8 in x, and in y when h > 8) whose 8x8 cell has
    coefficients at 0.6 and 0.8 x the threshold / sqrt(2) bands plus large ones
    that keep shifted patches far apart."""
    rng = np.random.default_rng(seed)
    X = rng.uniform(-0.06, 0.06, (8, 8))
    big = rng.choice(64, 10, replace=False)
    X.flat[big] = rng.choice([-1, 1], 10) * rng.uniform(0.6, 1.2, 10)
    X[0, 0] = 4.0
    # between lam*sigma/sqrt(2) and lam*sigma (kept with the sqrt(2) Haar scale) and
    # between lam*sigma/2 and lam*sigma/sqrt(2) (kept only by an unnormalised Haar)
    for (u, v), f in (((1, 2), 0.8), ((4, 1), -0.82), ((2, 5), 0.6), ((6, 6), -0.62)):
        X[u, v] = f * lam_sig
    cell = idct2(X)
    reps_y = 1 if h == 8 else h // 8 + 1
    img = np.tile(cell, (reps_y, w // 8 + 1))[:h, :w]
    return img.astype(np.float32)


SIG3 = 0.08


def case_duplicates(o, w, h, seed):
    """Groups of exact duplicates. The group structure comes from the oracle's
    block matching (pinned by brute force elsewhere); the test checks that every
    member is an exact duplicate of its reference patch, then compares step 1
    with the closed form idctn(HT(sqrt(n) X) / sqrt(n)), weight 1/N_nonzero."""
    img = _duplicate_image(w, h, seed, LAMBDA * SIG3)
    x = img.astype(np.float64)
    pos, cnt = o.block_match_f64(x)
    groups = groups_from_lists(pos, cnt, w, h)
    ns = set()
    num = np.zeros((h, w))
    den = np.zeros((h, w))
    for g in groups:
        rx, ry = g[0]
        ref = x[ry:ry + 8, rx:rx + 8]
        for (mx, my) in g:
            assert np.array_equal(x[my:my + 8, mx:mx + 8], ref), "member is not a duplicate"
        n = len(g)
        ns.add(n)
        Xs = np.sqrt(n) * dct2(ref)
        assert float(np.min(np.abs(np.abs(Xs) - LAMBDA * SIG3))) > 1e-4
        keep = np.abs(Xs) > LAMBDA * SIG3
        est = idct2(np.where(keep, Xs, 0.0) / np.sqrt(n))
        wgt = 1.0 / keep.sum()
        for (mx, my) in g:
            num[my:my + 8, mx:mx + 8] += wgt * KAISER2 * est
            den[my:my + 8, mx:mx + 8] += wgt * KAISER2
    basic = num / den
    # the general textbook step (Haar matrix) agrees with the duplicate closed form
    assert np.abs(textbook_step(x, None, groups, SIG3) - basic).max() < 1e-12
    _, b64 = o.bm3d_f64(x[None], SIG3, want_basic=True)
    _, b32 = o.bm3d_f32(img[None], SIG3, want_basic=True)
    return np.abs(b64[0] - basic).max(), np.abs(b32[0] - basic).max(), ns


def case_rgb_patch(o):
    """RGB 8x8: opponent transform z = A rgb, per-channel sigma sigma * |A_c|, one
    group; rgb out = A^-1 (per-channel closed forms)."""
    rng = np.random.default_rng(21)
    sig = 0.06
    rgb = rng.uniform(0.1, 0.9, (3, 8, 8)).astype(np.float32)
    x = rgb.astype(np.float64)
    z = np.tensordot(OPP, x, axes=1)
    sc = sig * np.linalg.norm(OPP, axis=1)
    basic_z, out_z = [], []
    for c in range(3):
        Z = dct2(z[c])
        assert float(np.min(np.abs(np.abs(Z) - LAMBDA * sc[c]))) > 1e-4
        b = idct2(np.where(np.abs(Z) > LAMBDA * sc[c], Z, 0.0))
        B = dct2(b)
        basic_z.append(b)
        out_z.append(idct2(B * B / (B * B + sc[c] ** 2) * Z))
    inv = np.linalg.inv(OPP)
    basic = np.tensordot(inv, np.stack(basic_z), axes=1)
    out = np.tensordot(inv, np.stack(out_z), axes=1)
    o64, b64 = o.bm3d_f64(x, sig, want_basic=True)
    o32, b32 = o.bm3d_f32(rgb, sig, want_basic=True)
    return (max(np.abs(b64 - basic).max(), np.abs(o64 - out).max()),
            max(np.abs(b32 - basic).max(), np.abs(o32 - out).max()))


TOL64, TOL32 = 1e-12, 2e-5


# --------------------------------------------------------------------------- pins

def test_haar_matrix_is_orthonormal():
    for n in (1, 2, 4, 8, 16):
        H = haar_matrix(n)
        np.testing.assert_allclose(H @ H.T, np.eye(n), atol=1e-14)
        # the first row is the DC: constant 1/sqrt(n)
        np.testing.assert_allclose(H[0], np.full(n, 1 / np.sqrt(n)), atol=1e-15)


def test_single_patch_step1_and_step2_closed_form(oracle_lib):
    e64, e32 = case_single_patch(oracle_lib)
    assert e64 < TOL64 and e32 < TOL32


def test_dc_is_thresholded_like_any_coefficient(oracle_lib):
    e64, e32 = case_dc_below_threshold(oracle_lib)
    assert e64 < TOL64 and e32 < TOL32


def test_two_overlapping_groups_weighted_mean(oracle_lib):
    (b64, o64), (b32, o32) = case_two_groups(oracle_lib)
    assert b64 < TOL64 and o64 < TOL64
    assert b32 < TOL32 and o32 < TOL32


@pytest.mark.parametrize("w,h,seed,want_n", [(16, 8, 1, {1, 2}), (40, 8, 2, {2, 4}),
                                             (40, 40, 3, {8, 16})])
def test_duplicate_groups_haar_scale(oracle_lib, w, h, seed, want_n):
    e64, e32, ns = case_duplicates(oracle_lib, w, h, seed)
    assert want_n <= ns, ns  # the group sizes the case is meant to exercise occur
    assert e64 < TOL64 and e32 < TOL32


def test_rgb_patch_opponent_sigmas(oracle_lib):
    e64, e32 = case_rgb_patch(oracle_lib)
    assert e64 < TOL64 and e32 < TOL32


def test_inverse_count_weight_beats_count(oracle_lib, tmp_path):
    """The paper's 1/N_nonzero weight against a literal reading of the prose
    ("proportional to count", planted mistake 2): on Gaussian noise, weighting
    sparse (well-matched) groups up gives the better basic estimate."""
    clean, noisy, sigma = synth.workload("gray96x96", seed=2)
    _, basic = oracle_lib.bm3d_f64(noisy.astype(np.float64), sigma, want_basic=True)
    with oracle_lib.using(oracle_lib.build_planted(2, str(tmp_path))):
        _, basic_n = oracle_lib.bm3d_f64(noisy.astype(np.float64), sigma, want_basic=True)
    assert synth.psnr(basic, clean) > synth.psnr(basic_n, clean) + 0.3


# --------------------------------------------------------------------------- planted mistakes

@pytest.fixture(scope="module")
def planted(tmp_path_factory, oracle_lib):
    d = str(tmp_path_factory.mktemp("planted"))
    return {k: oracle_lib.build_planted(k, d) for k in (1, 2, 3, 4)}


def _fails(err, tol):
    return err > 100 * tol


def test_planted_threshold_2p5_fails(oracle_lib, planted):
    with oracle_lib.using(planted[1]):
        e64, e32 = case_single_patch(oracle_lib)
    assert _fails(e64, TOL64) and _fails(e32, TOL32)


def test_planted_count_weight_fails(oracle_lib, planted):
    with oracle_lib.using(planted[2]):
        (b64, _), (b32, _) = case_two_groups(oracle_lib)
        d64, d32, _ = case_duplicates(oracle_lib, 40, 40, 3)
    assert _fails(b64, TOL64) and _fails(b32, TOL32)
    assert _fails(d64, TOL64) and _fails(d32, TOL32)


def test_planted_unnormalised_haar_fails(oracle_lib, planted):
    with oracle_lib.using(planted[3]):
        for w, h, seed in ((16, 8, 1), (40, 8, 2), (40, 40, 3)):
            e64, e32, _ = case_duplicates(oracle_lib, w, h, seed)
            assert _fails(e64, TOL64) and _fails(e32, TOL32), (w, h)
        # one-member groups have no stack transform: the single patch still passes
        s64, s32 = case_single_patch(oracle_lib)
    assert s64 < TOL64 and s32 < TOL32


def test_planted_wiener_weight_fails(oracle_lib, planted):
    with oracle_lib.using(planted[4]):
        (b64, o64), (b32, o32) = case_two_groups(oracle_lib)
    assert b64 < TOL64 and b32 < TOL32          # step 1 untouched
    assert _fails(o64, TOL64) and _fails(o32, TOL32)
