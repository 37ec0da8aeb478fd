"""Seeded synthetic scenes and noise for the BM3D node's workloads.

Scene: axis-aligned rectangles and discs at uniform levels over a uniform
background, plus two sinusoidal gratings (period 4-32 px) each confined to a
random rectangle. Everything is generated in float64 with numpy's PCG64 from
one seed, then cast to float32, so the CPU and GPU paths see identical bytes.

Noise models (both unclipped, as BM3D expects):
  * Gaussian:          y = x + sigma * N(0, 1)
  * Poisson-Gaussian:  y = gain * Poisson(x / gain) + read * N(0, 1)
The node's scalar sigma for the Poisson-Gaussian case is the metadata rule of
src/cpipe/nodes/denoise/bm3d.cpp:49-54: sigma^2 = mean_c max(0, beta + 0.25 lambda)
with beta = read^2 and lambda = gain.
"""
from __future__ import annotations

import math

import numpy as np


def clean_scene(width: int, height: int, channels: int = 1, seed: int = 0) -> np.ndarray:
    """Return a (channels, height, width) float32 scene in [0, 1]."""
    rng = np.random.default_rng(seed)
    img = np.empty((channels, height, width), dtype=np.float64)
    img[:] = rng.uniform(0.1, 0.9, size=(channels, 1, 1))
    n_shapes = 16 + (width * height) // 16384
    smax = max(4, min(width, height) // 5)
    smin = max(2, min(width, height) // 48)
    yy = np.arange(height)[:, None]
    xx = np.arange(width)[None, :]
    for _ in range(n_shapes):
        kind = rng.integers(0, 2)
        level = rng.uniform(0.02, 0.98, size=(channels, 1, 1))
        cx = rng.uniform(0, width)
        cy = rng.uniform(0, height)
        if kind == 0:
            hw = rng.uniform(smin, smax) / 2.0
            hh = rng.uniform(smin, smax) / 2.0
            x0, x1 = int(max(0, cx - hw)), int(min(width, cx + hw))
            y0, y1 = int(max(0, cy - hh)), int(min(height, cy + hh))
            if x1 > x0 and y1 > y0:
                img[:, y0:y1, x0:x1] = level
        else:
            rad = rng.uniform(smin, smax) / 2.0
            x0, x1 = int(max(0, cx - rad)), int(min(width, cx + rad + 1))
            y0, y1 = int(max(0, cy - rad)), int(min(height, cy + rad + 1))
            if x1 > x0 and y1 > y0:
                m = ((xx[:, x0:x1] - cx) ** 2 + (yy[y0:y1, :] - cy) ** 2) <= rad * rad
                sub = img[:, y0:y1, x0:x1]
                sub[:, m] = np.broadcast_to(level, (channels, 1, 1))[:, :, 0]
    for _ in range(2):
        period = rng.uniform(4.0, 32.0)
        theta = rng.uniform(0.0, math.pi)
        amp = rng.uniform(0.1, 0.3)
        gw = int(rng.uniform(0.2, 0.5) * width)
        gh = int(rng.uniform(0.2, 0.5) * height)
        x0 = int(rng.uniform(0, max(1, width - gw)))
        y0 = int(rng.uniform(0, max(1, height - gh)))
        gx = np.arange(x0, x0 + gw)[None, :]
        gy = np.arange(y0, y0 + gh)[:, None]
        wave = np.sin(2.0 * math.pi * (gx * math.cos(theta) + gy * math.sin(theta)) / period)
        img[:, y0:y0 + gh, x0:x0 + gw] += amp * wave[None, :, :]
    np.clip(img, 0.0, 1.0, out=img)
    return img.astype(np.float32)


def add_gaussian(clean: np.ndarray, sigma: float, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    noisy = clean.astype(np.float64) + sigma * rng.standard_normal(clean.shape)
    return noisy.astype(np.float32)


def add_poisson_gaussian(clean: np.ndarray, gain: float, read_sigma: float, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    lam = np.maximum(clean.astype(np.float64), 0.0) / gain
    noisy = gain * rng.poisson(lam) + read_sigma * rng.standard_normal(clean.shape)
    return noisy.astype(np.float32)


def sigma_from_noise_profile(pairs) -> float:
    """Scalar sigma from (beta, lambda) pairs -- bm3d.cpp:49-54 (float terms, double mean)."""
    if not pairs:
        return 0.03
    var = 0.0
    for beta, lam in pairs:
        var += max(0.0, float(np.float32(beta) + np.float32(lam) * np.float32(0.25)))
    var /= len(pairs)
    return float(np.float32(math.sqrt(var)))


def workload(name: str, seed: int = 0):
    """Named workloads: returns (clean, noisy, sigma) as (C, H, W) float32 arrays.

    * ``gray512``  -- 512x512x1, Gaussian sigma 0.05
    * ``rgb12mp``  -- 4080x3072x3, Poisson-Gaussian gain 0.01 / read 0.002
    * ``rgbWxH``   -- same noise model at another size, e.g. ``rgb1024x768``
    * ``grayWxH``  -- Gaussian sigma 0.05 at another size
    """
    if name == "gray512":
        name = "gray512x512"
    if name == "rgb12mp":
        name = "rgb4080x3072"
    if name.startswith("gray"):
        w, h = (int(v) for v in name[4:].split("x"))
        clean = clean_scene(w, h, 1, seed)
        return clean, add_gaussian(clean, 0.05, seed + 1), 0.05
    if name.startswith("rgb"):
        w, h = (int(v) for v in name[3:].split("x"))
        clean = clean_scene(w, h, 3, seed)
        noisy = add_poisson_gaussian(clean, 0.01, 0.002, seed + 1)
        sigma = sigma_from_noise_profile([(0.002 ** 2, 0.01)] * 3)
        return clean, noisy, sigma
    raise ValueError(f"unknown workload {name!r}")


def psnr(a: np.ndarray, b: np.ndarray, peak: float = 1.0) -> float:
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0.0 else 10.0 * math.log10(peak * peak / mse)


def max_abs(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))))
