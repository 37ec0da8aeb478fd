"""Pins the FP64 oracle of the learned pilot: analytic gradients vs central
finite differences (with the exact float32 step), and the residual identity
(zero last layer => output = input). Runs without a GPU."""
import numpy as np
import pytest

CINP = [16, 32, 32, 32, 32, 32]
COUTP = [32, 32, 32, 32, 32, 16]
CINR = [4, 32, 32, 32, 32, 32]
COUTR = [32, 32, 32, 32, 32, 3]


def _layer_slices():
    off, out = 0, []
    for l in range(6):
        nw = 9 * CINP[l] * COUTP[l]
        out.append((off, nw))
        off += nw + COUTP[l]
    return out, off


def _random_params(seed):
    rng = np.random.default_rng(seed)
    sl, n = _layer_slices()
    p = np.zeros(n, np.float32)
    for l, (off, nw) in enumerate(sl):
        w = np.zeros((9, CINP[l], COUTP[l]), np.float32)
        std = np.sqrt(2.0 / (9 * CINR[l])) * (0.3 if l == 5 else 1.0)
        w[:, :CINR[l], :COUTR[l]] = rng.normal(0, std, size=(9, CINR[l], COUTR[l]))
        p[off:off + nw] = w.ravel()
        p[off + nw:off + nw + COUTR[l]] = rng.normal(0, 0.05, size=COUTR[l])
    return p


def test_param_count(oracle_lib):
    assert oracle_lib.pilot_param_count() == _layer_slices()[1] == 46256


def test_residual_identity(oracle_lib):
    p = _random_params(0)
    sl, _ = _layer_slices()
    off, nw = sl[5]
    p[off:off + nw + COUTP[5]] = 0.0
    x = np.random.default_rng(1).normal(size=(2, 3, 5, 7))
    out = oracle_lib.pilot_forward_f64(p, x, np.array([0.02, 0.08]))
    np.testing.assert_allclose(out, x, atol=1e-15)


def test_gradient_matches_finite_differences(oracle_lib):
    rng = np.random.default_rng(2)
    p = _random_params(3)
    x = rng.normal(0.3, 0.2, size=(2, 3, 6, 7))
    t = x + rng.normal(0, 0.05, size=x.shape)
    sig = np.array([0.03, 0.07])
    loss, g = oracle_lib.pilot_grad_f64(p, x, t, sig)
    assert loss == pytest.approx(oracle_lib.pilot_loss_f64(p, x, t, sig), rel=1e-12)
    sl, _ = _layer_slices()
    checked = 0
    for l, (off, nw) in enumerate(sl):
        live = [off + (tap * CINP[l] + ci) * COUTP[l] + co
                for tap in range(9) for ci in range(CINR[l]) for co in range(COUTR[l])]
        live += [off + nw + co for co in range(COUTR[l])]
        for i in rng.choice(live, size=6, replace=False):
            eps = 2e-5 * max(1.0, abs(float(p[i])))
            pp, pm = p.copy(), p.copy()
            pp[i] = np.float32(p[i] + eps)
            pm[i] = np.float32(p[i] - eps)
            step = float(pp[i]) - float(pm[i])
            fd = (oracle_lib.pilot_loss_f64(pp, x, t, sig) - oracle_lib.pilot_loss_f64(pm, x, t, sig)) / step
            assert abs(fd - g[i]) <= 2e-3 * max(abs(fd), 1e-4) + 1e-9, (l, i, fd, g[i])
            checked += 1
    assert checked == 36
    # padded parameters get no gradient
    off, nw = sl[0]
    w0 = g[off:off + nw].reshape(9, 16, 32)
    assert np.all(w0[:, 4:, :] == 0)
