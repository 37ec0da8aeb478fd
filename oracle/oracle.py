"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module, and only as the reference result (or the timed CPU baseline). The
product package ``cpipe_amd`` never imports it.

What each function restates is documented in ``oracle/bm3d_oracle.h``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

K = 8
R = 19
STRIDE = 3
NMAX = 16
TAU_HT_SSD = np.float32(64.0 * 2500.0 / (255.0 * 255.0))
TAU_WIE_SSD = np.float32(64.0 * 400.0 / (255.0 * 255.0))

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def build_planted(plant: int, out_dir: str) -> str:
    """Compile a copy of the oracle with planted mistake `plant` (ORC_PLANT in
    bm3d_oracle.c) into out_dir; returns its path. Used only to show that the
    closed-form pins fail on each planted mistake."""
    path = os.path.join(out_dir, f"liboracle_plant{int(plant)}.so")
    src = [os.path.join(HERE, f) for f in ("bm3d_oracle.c", "pilot_oracle.c", "wavelet_oracle.c",
                                           "convert_oracle.c")]
    cmd = ["gcc", "-O2", "-march=x86-64-v3", "-fPIC", "-std=c11", "-D_GNU_SOURCE",
           "-ffp-contract=off", f"-DORC_PLANT={int(plant)}", "-shared", "-o", path] + src + ["-lm",
                                                                                          "-lpthread"]
    subprocess.run(cmd, check=True)
    return path


class using:
    """Context manager: route this module's functions to another build of the
    oracle (a planted-mistake copy from build_planted) while the block runs."""

    def __init__(self, path: str):
        self.path = path
        self.saved = None

    def __enter__(self):
        global _lib
        self.saved = _lib
        _lib = _load(self.path)
        return self

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


def _load(path: str) -> ctypes.CDLL:
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    I = ctypes.c_int
    F = ctypes.c_float
    D = ctypes.c_double
    S = ctypes.c_size_t
    L.oracle_f32_to_f16.argtypes = [F]
    L.oracle_f32_to_f16.restype = ctypes.c_uint16
    L.oracle_f16_to_f32.argtypes = [ctypes.c_uint16]
    L.oracle_f16_to_f32.restype = F
    L.oracle_compat_rgba16f.argtypes = [P, S, P, S, I, I, F, I]
    L.oracle_wavelet_rgba16f.argtypes = [P, S, P, S, I, I, F]
    L.oracle_convert_r16u_to_f32.argtypes = [P, S, P, S, I, I]
    L.oracle_convert_f32_to_rgba16f.argtypes = [P, S, P, S, I, I]
    L.oracle_convert_rgba16f_to_unorm.argtypes = [P, S, P, S, I, I, I]
    L.oracle_compat_rgba16f.restype = None
    L.oracle_ref_count.argtypes = [I]
    L.oracle_ref_count.restype = I
    L.oracle_ref_pos.argtypes = [I, I]
    L.oracle_ref_pos.restype = I
    L.oracle_dct_matrix_f64.argtypes = [P]
    L.oracle_dct8_f32.argtypes = [P, P, I]
    L.oracle_dct8_f32.restype = None
    L.oracle_kaiser_f64.argtypes = [P]
    L.oracle_bm_f32.argtypes = [P, I, I, F, P, P, I]
    L.oracle_bm_f64.argtypes = [P, I, I, D, P, P]
    L.oracle_bm3d_f32.argtypes = [P, P, P, I, I, I, F, I, P, P, P, P]
    L.oracle_bm3d_f32.restype = I
    L.oracle_bm3d_f64.argtypes = [P, P, P, I, I, I, D]
    L.oracle_bm3d_f64.restype = I
    L.oracle_sigma_map_f32.argtypes = [P, S, I, I, I, P, P, P]
    L.oracle_sigma_map_f32.restype = I
    L.oracle_bm3d_f32_profile.argtypes = [P, P, P, I, I, I, P, P, I]
    L.oracle_bm3d_f32_profile.restype = I
    L.oracle_bm3d_f64_profile.argtypes = [P, P, P, I, I, I, P, P]
    L.oracle_bm3d_f64_profile.restype = I
    L.oracle_pilot_param_count.restype = S
    L.oracle_pilot_forward_f64.argtypes = [P, P, P, I, I, I, P]
    L.oracle_pilot_loss_f64.argtypes = [P, P, P, P, I, I, I]
    L.oracle_pilot_loss_f64.restype = D
    L.oracle_pilot_grad_f64.argtypes = [P, P, P, P, I, I, I, P]
    L.oracle_pilot_grad_f64.restype = D
    return L


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _plane_ptrs(planes):
    arr = (ctypes.c_void_p * 3)(*([_ptr(p) for p in planes] + [0] * (3 - len(planes))))
    return arr


def f32_to_f16_bits(x: np.ndarray) -> np.ndarray:
    L = lib()
    flat = np.asarray(x, dtype=np.float32).ravel()
    return np.array([L.oracle_f32_to_f16(float(v)) for v in flat], dtype=np.uint16).reshape(np.shape(x))


def f16_bits_to_f32(h: np.ndarray) -> np.ndarray:
    L = lib()
    flat = np.asarray(h, dtype=np.uint16).ravel()
    return np.array([L.oracle_f16_to_f32(int(v)) for v in flat], dtype=np.float32).reshape(np.shape(h))


def compat_rgba16f(rgba_bits: np.ndarray, sigma: float, stage: int = 1) -> np.ndarray:
    """Restated reference step (stage 1: 2x2 blend, stage 2: identity) on (H, W, 4) uint16."""
    src = np.ascontiguousarray(rgba_bits, dtype=np.uint16)
    h, w, c = src.shape
    assert c == 4
    out = np.empty_like(src)
    lib().oracle_compat_rgba16f(_ptr(src), w, _ptr(out), w, w, h, float(sigma), int(stage))
    return out


def wavelet_rgba16f(rgba_bits: np.ndarray, chroma_strength: float = 1.0) -> np.ndarray:
    """Restated wavelet BayesShrink chroma denoise on (H, W, 4) uint16 (binary16 bits)."""
    src = np.ascontiguousarray(rgba_bits, dtype=np.uint16)
    h, w, c = src.shape
    assert c == 4
    out = np.empty_like(src)
    lib().oracle_wavelet_rgba16f(_ptr(src), w, _ptr(out), w, w, h, float(chroma_strength))
    return out


def convert(src: np.ndarray, target: str) -> np.ndarray:
    """Restated precision conversions: (H, W) uint16 -> "R32_SFLOAT" (H, W) float32;
    (H, W) float32 -> "R16G16B16A16_SFLOAT" (H, W, 4) uint16 bits; (H, W, 4) uint16
    binary16 bits -> "R8G8B8A8_UNORM" (H, W, 4) uint8 or "R16G16B16A16_UNORM" uint16."""
    src = np.ascontiguousarray(src)
    h, w = src.shape[:2]
    L = lib()
    if target == "R32_SFLOAT":
        assert src.dtype == np.uint16 and src.ndim == 2
        out = np.empty((h, w), np.float32)
        L.oracle_convert_r16u_to_f32(_ptr(src), src.strides[0], _ptr(out), out.strides[0], w, h)
    elif target == "R16G16B16A16_SFLOAT":
        assert src.dtype == np.float32 and src.ndim == 2
        out = np.empty((h, w, 4), np.uint16)
        L.oracle_convert_f32_to_rgba16f(_ptr(src), src.strides[0], _ptr(out), out.strides[0], w, h)
    else:
        assert src.dtype == np.uint16 and src.shape[2] == 4
        bits = 8 if target == "R8G8B8A8_UNORM" else 16
        out = np.empty((h, w, 4), np.uint8 if bits == 8 else np.uint16)
        L.oracle_convert_rgba16f_to_unorm(_ptr(src), src.strides[0], _ptr(out), out.strides[0], w, h, bits)
    return out


def ref_positions(n: int) -> np.ndarray:
    L = lib()
    return np.array([L.oracle_ref_pos(i, n) for i in range(L.oracle_ref_count(n))], dtype=np.int64)


def dct8_f32(x, inverse: bool = False) -> np.ndarray:
    """One 8-point orthonormal DCT-II (or its inverse) with the FP32 butterfly
    order of the FP32 restatement (bm3d_oracle.c dct8_fwd_f32 / dct8_inv_f32)."""
    a = np.ascontiguousarray(x, np.float32).reshape(8)
    out = np.empty(8, np.float32)
    lib().oracle_dct8_f32(_ptr(a), _ptr(out), 1 if inverse else 0)
    return out


def dct_matrix() -> np.ndarray:
    c = np.zeros(64, dtype=np.float64)
    lib().oracle_dct_matrix_f64(_ptr(c))
    return c.reshape(8, 8)


def kaiser() -> np.ndarray:
    k = np.zeros(8, dtype=np.float64)
    lib().oracle_kaiser_f64(_ptr(k))
    return k


def block_match_f32(plane: np.ndarray, tau_ssd=TAU_HT_SSD, nthreads: int = 8):
    p = np.ascontiguousarray(plane, dtype=np.float32)
    h, w = p.shape
    L = lib()
    nref = L.oracle_ref_count(w) * L.oracle_ref_count(h)
    pos = np.zeros((nref, NMAX), dtype=np.uint32)
    cnt = np.zeros(nref, dtype=np.uint8)
    L.oracle_bm_f32(_ptr(p), w, h, float(tau_ssd), _ptr(pos), _ptr(cnt), int(nthreads))
    return pos, cnt


def block_match_f64(plane: np.ndarray, tau_ssd=64.0 * 2500.0 / 255.0**2):
    p = np.ascontiguousarray(plane, dtype=np.float64)
    h, w = p.shape
    L = lib()
    nref = L.oracle_ref_count(w) * L.oracle_ref_count(h)
    pos = np.zeros((nref, NMAX), dtype=np.uint32)
    cnt = np.zeros(nref, dtype=np.uint8)
    L.oracle_bm_f64(_ptr(p), w, h, float(tau_ssd), _ptr(pos), _ptr(cnt))
    return pos, cnt


def bm3d_f32(planes: np.ndarray, sigma: float, nthreads: int = 8, want_basic: bool = False,
             want_matches: bool = False):
    """planes: (C, H, W) float32, C in {1, 3}. Returns out (and basic / matches if asked)."""
    src = np.ascontiguousarray(planes, dtype=np.float32)
    c, h, w = src.shape
    out = np.empty_like(src)
    basic = np.empty_like(src) if want_basic else None
    L = lib()
    nref = max(L.oracle_ref_count(w) * L.oracle_ref_count(h), 1)
    m1 = np.zeros((nref, NMAX), np.uint32) if want_matches else None
    c1 = np.zeros(nref, np.uint8) if want_matches else None
    m2 = np.zeros((nref, NMAX), np.uint32) if want_matches else None
    c2 = np.zeros(nref, np.uint8) if want_matches else None
    rc = L.oracle_bm3d_f32(
        _plane_ptrs(list(src)), _plane_ptrs(list(out)),
        _plane_ptrs(list(basic)) if want_basic else None, w, h, c, float(sigma), int(nthreads),
        _ptr(m1) if want_matches else None, _ptr(c1) if want_matches else None,
        _ptr(m2) if want_matches else None, _ptr(c2) if want_matches else None)
    if rc != 0:
        raise RuntimeError(f"oracle_bm3d_f32 failed: {rc}")
    res = [out]
    if want_basic:
        res.append(basic)
    if want_matches:
        res.append((m1, c1, m2, c2))
    return res[0] if len(res) == 1 else tuple(res)


def bm3d_f64(planes: np.ndarray, sigma: float, want_basic: bool = False):
    src = np.ascontiguousarray(planes, dtype=np.float64)
    c, h, w = src.shape
    out = np.empty_like(src)
    basic = np.empty_like(src) if want_basic else None
    rc = lib().oracle_bm3d_f64(_plane_ptrs(list(src)), _plane_ptrs(list(out)),
                               _plane_ptrs(list(basic)) if want_basic else None, w, h, c,
                               float(sigma))
    if rc != 0:
        raise RuntimeError(f"oracle_bm3d_f64 failed: {rc}")
    return (out, basic) if want_basic else out


def _profile(beta, lam, c, dtype):
    b = np.ascontiguousarray(np.broadcast_to(np.asarray(beta, dtype), (c,)))
    l_ = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, dtype), (c,)))
    return b, l_


def sigma_map_f32(planes: np.ndarray, beta, lam) -> np.ndarray:
    """Per-ref sigma of the noise profile variance beta + lam * x: (C, ny, nx) float32."""
    src = np.ascontiguousarray(planes, dtype=np.float32)
    c, h, w = src.shape
    L = lib()
    nx, ny = L.oracle_ref_count(w), L.oracle_ref_count(h)
    out = np.zeros((c, ny, nx), np.float32)
    b, l_ = _profile(beta, lam, c, np.float32)
    rc = L.oracle_sigma_map_f32(_plane_ptrs(list(src)), w, w, h, c, _ptr(b), _ptr(l_), _ptr(out))
    if rc != 0:
        raise RuntimeError(f"oracle_sigma_map_f32 failed: {rc}")
    return out


def bm3d_f32_profile(planes: np.ndarray, beta, lam, nthreads: int = 8, want_basic: bool = False):
    """BM3D with the per-ref sigma map of a noise profile (per channel beta, lam)."""
    src = np.ascontiguousarray(planes, dtype=np.float32)
    c, h, w = src.shape
    out = np.empty_like(src)
    basic = np.empty_like(src) if want_basic else None
    b, l_ = _profile(beta, lam, c, np.float32)
    rc = lib().oracle_bm3d_f32_profile(_plane_ptrs(list(src)), _plane_ptrs(list(out)),
                                       _plane_ptrs(list(basic)) if want_basic else None, w, h, c,
                                       _ptr(b), _ptr(l_), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle_bm3d_f32_profile failed: {rc}")
    return (out, basic) if want_basic else out


def bm3d_f64_profile(planes: np.ndarray, beta, lam) -> np.ndarray:
    src = np.ascontiguousarray(planes, dtype=np.float64)
    c, h, w = src.shape
    out = np.empty_like(src)
    b, l_ = _profile(beta, lam, c, np.float64)
    rc = lib().oracle_bm3d_f64_profile(_plane_ptrs(list(src)), _plane_ptrs(list(out)), None, w, h, c,
                                       _ptr(b), _ptr(l_))
    if rc != 0:
        raise RuntimeError(f"oracle_bm3d_f64_profile failed: {rc}")
    return out


# ---------------------------------------------------------------- learned pilot
def pilot_param_count() -> int:
    return int(lib().oracle_pilot_param_count())


def pilot_forward_f64(params: np.ndarray, inp: np.ndarray, sigma) -> np.ndarray:
    """inp: (N, 3, H, W) opponent planes; sigma: (N,). Returns (N, 3, H, W) float64."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(inp, dtype=np.float64)
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    n, c, h, w = x.shape
    out = np.empty_like(x)
    lib().oracle_pilot_forward_f64(_ptr(p), _ptr(x), _ptr(s), w, h, n, _ptr(out))
    return out


def pilot_loss_f64(params, inp, target, sigma) -> float:
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(inp, dtype=np.float64)
    t = np.ascontiguousarray(target, dtype=np.float64)
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    n, c, h, w = x.shape
    return float(lib().oracle_pilot_loss_f64(_ptr(p), _ptr(x), _ptr(t), _ptr(s), w, h, n))


def pilot_grad_f64(params, inp, target, sigma):
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(inp, dtype=np.float64)
    t = np.ascontiguousarray(target, dtype=np.float64)
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    n, c, h, w = x.shape
    g = np.zeros(pilot_param_count(), dtype=np.float64)
    loss = lib().oracle_pilot_grad_f64(_ptr(p), _ptr(x), _ptr(t), _ptr(s), w, h, n, _ptr(g))
    return float(loss), g
