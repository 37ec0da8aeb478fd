"""One BM3D pass for profilers (rocprofv3 kernel trace / PMC counters).

Runs the plan on the named workload `--runs` times, then a 256 MiB
device-to-device copy whose counter readings calibrate FETCH_SIZE and
WRITE_SIZE (the copy reads and writes exactly 268435456 bytes).

Usage: python -m cpipe_amd.profile_run --workload rgb12mp --precision fp32 [--pilot] [--bm mfma]
       python -m cpipe_amd.profile_run --wavelet   (the wavelet BayesShrink node's kernel instead)
"""
from __future__ import annotations

import argparse
import sys


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="rgb12mp")
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--runs", type=int, default=1)
    ap.add_argument("--pilot", action="store_true", help="learned pilot with the packaged weights")
    ap.add_argument("--noise-profile", default=None, metavar="BETA,LAMBDA",
                    help="per-reference sigma map of this noise profile (every channel)")
    ap.add_argument("--bm", choices=["exact", "mfma"], default=None,
                    help="block-matching kernel of a mixed-precision plan (default: the plan's)")
    ap.add_argument("--wavelet", action="store_true",
                    help="run the wavelet BayesShrink kernel on the frame as RGBA16F instead of BM3D")
    args = ap.parse_args(argv)
    import numpy as np
    import torch
    from . import engine, synth
    clean, noisy, sigma = synth.workload(args.workload, seed=0)
    c, h, w = noisy.shape
    if args.wavelet:
        rgba = np.ones((h, w, 4), np.float32)
        rgba[..., :c] = np.moveaxis(noisy, 0, -1)
        src = torch.from_numpy(rgba.astype(np.float16).view(np.int16)).cuda()
        dst = torch.empty_like(src)
        for i in range(args.runs):
            engine.wavelet_rgba16f(src, dst, 1.0)
            torch.cuda.synchronize()
            print(f"wavelet run {i} done", file=sys.stderr, flush=True)
        _calibration_copy(torch)
        return
    src = torch.from_numpy(noisy).cuda()
    dst = torch.empty_like(src)
    plan = engine.Plan(w, h, c, args.precision, 0)
    if args.noise_profile:
        beta, lam = (float(v) for v in args.noise_profile.split(","))
        plan.set_noise_profile(beta, lam)
    if args.bm is not None:
        plan.set_block_matching(args.bm)
    if args.pilot:
        from . import pilot
        plan.set_pilot(pilot.Pilot(seed=None, weights=pilot.DEFAULT_WEIGHTS))
    for i in range(args.runs):
        plan.run(src, dst, sigma)
        torch.cuda.synchronize()
        print(f"run {i} done", file=sys.stderr, flush=True)
    plan.check()
    _calibration_copy(torch)


def _calibration_copy(torch):
    a = torch.full((64 * 1024 * 1024,), 1.0, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print("calibration copy done (268435456 bytes read and written)", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
