"""Regenerates the golden pairs tests/golden/denoise.bm3d/{in,out}.exr and
tests/golden/denoise.wavelet_bayes_shrink/{in,out}.exr.

The reference's pairs of the same names are stored as Git LFS pointer stubs, so
their pixels are rebuilt here from the formulas of the reference's generator
tests/fixtures/gen_golden_isp_nodes.cpp:
  denoise_sample()               :556-564  (edge-clamped, half-rounded sample)
  denoise_input()                :619-634  (8x8 ramp + chroma checkerboard + luma pattern)
  ycocg_at() / soft_threshold()  :656-673  (luma, co, cg; threshold 0.035)
  wavelet_chroma()               :675-697  (2x2 Haar bands, soft-thresholded, rebuilt)
  wavelet_bayes_shrink_output()  :699-715  (RGB = half(max(0, .)), alpha = half(alpha))
  bm3d_mean2x2() / bm3d_output() :717-740  (RGB = half(max(0, 2x2 mean)), alpha = half(alpha))
and written the way write_pair (:1100-1120) writes them: FLOAT pixels, 4
channels, in.exr holding the un-rounded float32 input. All arithmetic is float32
like the C++ generator; halves are IEEE round to nearest even.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from cpipe_amd import exr  # noqa: E402

f = np.float32


def half(x):
    return np.float32(np.float16(np.float32(x)))


def denoise_input():
    img = np.zeros((8, 8, 4), np.float32)
    for y in range(8):
        for x in range(8):
            base = f(0.08) + f(0.018) * f(x) + f(0.014) * f(y)
            chroma = f(0.035) if ((x + y) & 1) == 0 else f(-0.035)
            luma = f(((x * 3 + y * 5) % 5) - 2) * f(0.006)
            img[y, x] = [base + luma + chroma, base + f(0.5) * luma, base + luma - chroma, f(1.0)]
    return img


def _sampler(img):
    h, w = img.shape[:2]
    return lambda x, y, c: half(img[min(max(y, 0), h - 1), min(max(x, 0), w - 1), c])


def _soft(v):
    t = f(0.035)
    return f(v - t) if v > t else (f(v + t) if v < -t else f(0.0))


def wavelet_output(img):
    h, w = img.shape[:2]
    sample = _sampler(img)

    def ycocg(x, y):
        r, g, b = sample(x, y, 0), sample(x, y, 1), sample(x, y, 2)
        return [f(f(f(r + f(f(2.0) * g)) + b) * f(0.25)), f(r - b), f(g - f(f(r + b) * f(0.5)))]

    def chroma(x, y, k):
        bx, by = (x // 2) * 2, (y // 2) * 2
        p00, p10 = ycocg(bx, by)[k], ycocg(bx + 1, by)[k]
        p01, p11 = ycocg(bx, by + 1)[k], ycocg(bx + 1, by + 1)[k]
        ll = f(f(0.25) * f(f(f(p00 + p10) + p01) + p11))
        hh = _soft(f(f(0.25) * f(f(f(p00 - p10) + p01) - p11)))
        vv = _soft(f(f(0.25) * f(f(f(p00 + p10) - p01) - p11)))
        dd = _soft(f(f(0.25) * f(f(f(p00 - p10) - p01) + p11)))
        if x % 2 == 0 and y % 2 == 0:
            return f(f(f(ll + hh) + vv) + dd)
        if x % 2 == 1 and y % 2 == 0:
            return f(f(f(ll - hh) + vv) - dd)
        if x % 2 == 0:
            return f(f(f(ll + hh) - vv) - dd)
        return f(f(f(ll - hh) - vv) + dd)

    out = np.zeros_like(img)
    for y in range(h):
        for x in range(w):
            luma = ycocg(x, y)[0]
            co, cg = chroma(x, y, 1), chroma(x, y, 2)
            t = f(luma - f(f(0.5) * cg))
            out[y, x] = [half(max(f(0.0), f(t + f(f(0.5) * co)))), half(max(f(0.0), f(t + cg))),
                         half(max(f(0.0), f(t - f(f(0.5) * co)))), half(sample(x, y, 3))]
    return out


def bm3d_output(img):
    h, w = img.shape[:2]
    sample = lambda x, y, c: half(img[min(max(y, 0), h - 1), min(max(x, 0), w - 1), c])
    out = np.zeros_like(img)
    for y in range(h):
        for x in range(w):
            for c in range(4):
                if c == 3:
                    out[y, x, c] = half(sample(x, y, c))
                    continue
                bx, by = (x // 2) * 2, (y // 2) * 2
                m = f(0.25) * (((sample(bx, by, c) + sample(bx + 1, by, c)) + sample(bx, by + 1, c))
                               + sample(bx + 1, by + 1, c))
                out[y, x, c] = half(max(f(0.0), m))
    return out


def main():
    inp = denoise_input()
    root = os.path.dirname(os.path.abspath(__file__))
    for name, out in (("denoise.bm3d", bm3d_output(inp)), ("denoise.wavelet_bayes_shrink", wavelet_output(inp))):
        d = os.path.join(root, name)
        os.makedirs(d, exist_ok=True)
        exr.write(os.path.join(d, "in.exr"), inp, "float")
        exr.write(os.path.join(d, "out.exr"), out, "float")
        print("wrote", d, "out pixel(0,0) =", out[0, 0])


if __name__ == "__main__":
    main()
