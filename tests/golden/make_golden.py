"""Regenerates tests/golden/denoise_bm3d_golden.json.

The reference's golden pair tests/golden/denoise.bm3d/{in,out}.exr is stored as
Git LFS pointer stubs, so its pixels are rebuilt here from the formulas of the
reference's generator tests/fixtures/gen_golden_isp_nodes.cpp:
  denoise_input()  :619-634  (8x8 ramp + chroma checkerboard + luma pattern)
  bm3d_mean2x2()   :717-724  (2x2-aligned mean of half-rounded samples)
  bm3d_output()    :726-740  (RGB = half(max(0, mean)), alpha = half(alpha))
All arithmetic is float32 like the C++ generator; halves are IEEE RNE.
"""
import json
import os

import numpy as np

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
    in_bits = inp.astype(np.float16).view(np.uint16)
    out = bm3d_output(inp.astype(np.float16).astype(np.float32))
    out_bits = out.astype(np.float16).view(np.uint16)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "denoise_bm3d_golden.json")
    with open(path, "w") as fh:
        json.dump({"source": "tests/fixtures/gen_golden_isp_nodes.cpp:619-634,717-740",
                   "width": 8, "height": 8, "format": "R16G16B16A16_SFLOAT",
                   "in_f16_bits": in_bits.ravel().tolist(),
                   "out_f16_bits": out_bits.ravel().tolist()}, fh)
    print("wrote", path, "pixel(0,0) =", out[0, 0])


if __name__ == "__main__":
    main()
