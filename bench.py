"""BM3D node throughput on one MI355X (Mpixel/s), with roofline and CPU baseline.

Workload (default): 4080x3072 RGB float32 linear planes in [0, 1] from the
seeded shape+texture generator, Poisson-Gaussian noise (gain 0.01, read 0.002),
node sigma from the noise-profile rule (0.05004), both BM3D steps, plain FP32.
A step is one full BM3D pass over the frame with the input already resident in
HBM. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cpipe_amd import engine, synth  # noqa: E402

K, R, STRIDE = 8, 19, 3
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def ref_positions(n: int) -> np.ndarray:
    last = n - K
    return np.minimum(np.arange((last + STRIDE - 1) // STRIDE + 1) * STRIDE, last)


def bm_candidates(w: int, h: int) -> int:
    """Candidate positions over every reference's clipped 39x39 window (reference included)."""
    def axis(n):
        r = ref_positions(n)
        return np.minimum(n - K, r + R) - np.maximum(0, r - R) + 1
    return int(axis(w).sum()) * int(axis(h).sum())


def group_flops(counts: np.ndarray, channels: int, step: int) -> float:
    """Nominal FLOPs of the group kernels: per group and channel, a 3D transform costs
    n*4k^3 + 4k^2(n-1); step 1 = 2 transforms + aggregation 4nk^2;
    step 2 = 3 transforms + Wiener 4nk^2 + aggregation 4nk^2 (n = group size)."""
    c = counts.astype(np.int64)
    n = np.where(c >= 16, 16, np.where(c >= 8, 8, np.where(c >= 4, 4, np.where(c >= 2, 2, 1))))
    t = n * 4 * K ** 3 + 4 * K * K * (n - 1)
    agg = 4 * n * K * K
    per = 2 * t + agg if step == 1 else 3 * t + 2 * agg
    return float(per.sum()) * channels


def cpu_baseline(noisy: np.ndarray, sigma: float, sample_w: int, sample_h: int):
    from oracle import oracle
    threads = min(16, os.cpu_count() or 1)
    crop = np.ascontiguousarray(noisy[:, :sample_h, :sample_w])
    t0 = time.perf_counter()
    oracle.bm3d_f32(crop, sigma, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(sample_w * sample_h / dt / 1e6, 4), "unit": "Mpixel/s",
            "cores": threads, "kind": "port",
            "sample": f"{sample_w}x{sample_h}x3 top-left crop of the same noisy frame, "
                      f"CPU FP32 BM3D (same algorithm and operation order), {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rgb12mp")
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample", default="2048x1536")
    ap.add_argument("--pilot", default=None,
                    help="learned-pilot weight blob (default: the packaged weights if present)")
    ap.add_argument("--no-pilot", action="store_true", help="skip the learned-pilot form")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import torch
    dev = torch.device("cuda:0")
    clean, noisy, sigma = synth.workload(args.workload, seed=0)
    c, h, w = noisy.shape
    print(f"# workload {args.workload}: {w}x{h}x{c} sigma={sigma:.5f}", file=sys.stderr, flush=True)
    src = torch.from_numpy(noisy).to(dev)
    dst = torch.empty_like(src)
    plan = engine.Plan(w, h, c, args.precision, 0)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        plan.run(src, dst, sigma, stream)
    torch.cuda.synchronize()
    print("# warmup done", file=sys.stderr, flush=True)

    plan.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run(src, dst, sigma, stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    plan.check()
    stage_ms, runs = plan.timing()
    ms_per_step = wall * 1e3 / args.steps
    mpx = w * h / (ms_per_step / 1e3) / 1e6
    per_stage = {k: v / runs for k, v in stage_ms.items()}
    print(f"# per-stage ms: {json.dumps({k: round(v, 3) for k, v in per_stage.items()})}",
          file=sys.stderr, flush=True)

    out = dst.cpu().numpy()
    psnr_out = synth.psnr(out, clean)
    psnr_in = synth.psnr(noisy, clean)

    _, cnt1 = plan.matches(1)
    _, cnt2 = plan.matches(2)
    cnt1 = cnt1.cpu().numpy()
    cnt2 = cnt2.cpu().numpy()
    cand = bm_candidates(w, h)
    bm_flop = 192.0 * cand
    g1_flop = group_flops(cnt1, c, 1)
    g2_flop = group_flops(cnt2, c, 2)
    kernels = {
        "block_match": (bm_flop, (per_stage["bm1"] + per_stage["bm2"]) / 2),
        "group_ht": (g1_flop, per_stage["group1"]),
        "group_wiener": (g2_flop, per_stage["group2"]),
    }
    total_ms = {"block_match": per_stage["bm1"] + per_stage["bm2"],
                "group_ht": per_stage["group1"], "group_wiener": per_stage["group2"]}
    dominant = max(total_ms, key=total_ms.get)
    flop, ms = kernels[dominant]
    achieved = flop / (ms / 1e3) / 1e12
    total_flop = 2 * bm_flop + g1_flop + g2_flop
    roofline = {"bound": "mfma", "kernel": dominant, "achieved": round(achieved, 3),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                "flop_per_launch": flop, "launch_ms": round(ms, 4),
                "end_to_end_tflops": round(total_flop / (ms_per_step / 1e3) / 1e12, 3)}

    learned = None
    from cpipe_amd import pilot as pilot_mod
    blob = args.pilot or (pilot_mod.DEFAULT_WEIGHTS if os.path.exists(pilot_mod.DEFAULT_WEIGHTS) else None)
    if blob and not args.no_pilot and c == 3:
        net = pilot_mod.Pilot(seed=None, weights=blob)
        plan.set_pilot(net)
        dst2 = torch.empty_like(src)
        plan.set_timing(False)
        for _ in range(args.warmup):
            plan.run(src, dst2, sigma, stream)
        torch.cuda.synchronize()
        plan.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.run(src, dst2, sigma, stream)
        torch.cuda.synchronize()
        wall2 = time.perf_counter() - t0
        plan.check()
        st2, runs2 = plan.timing()
        rep = plan.pilot_report()
        out2 = dst2.cpu().numpy()
        ms2 = wall2 * 1e3 / args.steps
        psnr2 = synth.psnr(out2, clean)
        learned = {"value": round(w * h / (ms2 / 1e3) / 1e6, 3), "unit": "Mpixel/s",
                   "ms_per_step": round(ms2, 3), "psnr_out_db": round(psnr2, 3),
                   "delta_psnr_db_vs_plain": round(psnr2 - psnr_out, 4),
                   "speedup_vs_fp32": round(ms_per_step / ms2, 3),
                   "stage_ms": {k: round(v / runs2, 4) for k, v in st2.items()},
                   "spot_checks": rep, "weights": os.path.basename(blob)}
        plan.set_pilot(None)
        print(f"# learned pilot: {json.dumps(learned)}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        sw, sh = (int(v) for v in args.cpu_sample.split("x"))
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(noisy, sigma, min(sw, w), min(sh, h))

    line = {
        "metric": "BM3D Mpixel/s at 12.5MP RGB (plain FP32, both steps)",
        "value": round(mpx, 3), "unit": "Mpixel/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload} {w}x{h}x{c} float32 planes, Poisson-Gaussian "
                               f"gain 0.01 read 0.002, sigma {sigma:.5f}, both steps",
                   "precision": args.precision, "frames": 1},
        "quality": {"psnr_noisy_db": round(psnr_in, 3), "psnr_out_db": round(psnr_out, 3)},
        "stage_ms": {k: round(v, 4) for k, v in per_stage.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "learned_pilot": learned,
    }
    if cpu:
        line["gpu_over_cpu"] = round(mpx / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    plan.close()


if __name__ == "__main__":
    main()
