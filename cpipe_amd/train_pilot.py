"""Train the learned pilot on pairs the node generates itself.

Data: seeded 128x128 crops of scenes from the synthetic shape+texture
generator (cpipe_amd.synth), each with its own noise: half Gaussian with
sigma ~ U(0.01, 0.1), half Poisson-Gaussian (gain log-uniform in
[0.002, 0.03], read noise U(0.0005, 0.01)) with the node's scalar sigma from
the noise-profile rule sqrt(read^2 + gain / 4), as a camera frame's metadata
gives it. Targets: the node's own plain step-1 (hard-threshold)
basic estimate of each noisy crop, in opponent space
(cpipe_bm3d_plan_step1_opponent). Training: the library's FP16/MFMA
forward/backward and FP32 Adam (cpipe_bm3d_pilot_train_step), cosine learning
rate decay, seeded shuffling. Output: a versioned weight blob.

Usage: python -m cpipe_amd.train_pilot --crops 20000 --epochs 6 --out pilot.bin
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import engine, pilot, synth


def make_crops(n: int, size: int = 128, seed: int = 0, frame: int = 1024,
               sigma_range=(0.01, 0.1), poisson_fraction: float = 0.5):
    """Random crops of synthetic RGB scenes with per-crop Gaussian or Poisson-Gaussian noise."""
    rng = np.random.default_rng(seed)
    per_frame = max(1, min(n, 4 * (frame // size) ** 2))
    clean = np.empty((n, 3, size, size), np.float32)
    i, f = 0, 0
    while i < n:
        scene = synth.clean_scene(frame, frame, 3, seed=seed * 100003 + f)
        f += 1
        for _ in range(min(per_frame, n - i)):
            y = int(rng.integers(0, frame - size + 1))
            x = int(rng.integers(0, frame - size + 1))
            clean[i] = scene[:, y:y + size, x:x + size]
            i += 1
    sigma = rng.uniform(*sigma_range, size=n).astype(np.float32)
    noisy = clean + sigma[:, None, None, None] * rng.standard_normal(clean.shape).astype(np.float32)
    pg = np.flatnonzero(rng.random(n) < poisson_fraction)
    for i in pg:
        gain = float(np.exp(rng.uniform(np.log(0.002), np.log(0.03))))
        read = float(rng.uniform(0.0005, 0.01))
        lam = np.maximum(clean[i].astype(np.float64), 0.0) / gain
        noisy[i] = gain * rng.poisson(lam) + read * rng.standard_normal(clean[i].shape)
        sigma[i] = synth.sigma_from_noise_profile([(read * read, gain)] * 3)
    return clean, noisy.astype(np.float32), sigma


def build_pairs(noisy: np.ndarray, sigma: np.ndarray, device: int = 0):
    """Opponent noisy planes and plain step-1 targets on the GPU, (N, 3, S, S) each."""
    import torch
    n, _, s, _ = noisy.shape
    dev = torch.device(f"cuda:{device}")
    X = torch.empty((n, 3, s, s), dtype=torch.float32, device=dev)
    T = torch.empty_like(X)
    plan = engine.Plan(s, s, 3, "fp32", device)
    src = torch.from_numpy(noisy).to(dev)
    for i in range(n):
        plan.step1_opponent(src[i], X[i], T[i], float(sigma[i]))
    torch.cuda.synchronize()
    plan.check()
    plan.close()
    return X, T


def train(net: pilot.Pilot, X, T, sigma: np.ndarray, epochs: int, batch: int, lr: float,
          seed: int = 0, log=None):
    import torch
    n = X.shape[0]
    rng = np.random.default_rng(seed)
    steps_per_epoch = n // batch
    total = epochs * steps_per_epoch
    step = 0
    history = []
    for ep in range(epochs):
        perm = rng.permutation(n)
        acc = 0.0
        t0 = time.perf_counter()
        for b in range(steps_per_epoch):
            idx = perm[b * batch:(b + 1) * batch]
            ti = torch.from_numpy(idx).to(X.device)
            xb = X.index_select(0, ti).contiguous()
            tb = T.index_select(0, ti).contiguous()
            cur = lr * 0.5 * (1.0 + math.cos(math.pi * step / max(1, total)))
            acc += net.train_step(xb, tb, sigma[idx], cur)
            step += 1
        history.append(acc / steps_per_epoch)
        if log:
            log(f"epoch {ep}: loss {history[-1]:.3e} ({time.perf_counter() - t0:.1f} s)")
    return history


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--crops", type=int, default=20000)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=pilot.DEFAULT_WEIGHTS)
    ap.add_argument("--checkpoint", default=None, help="also write weights + Adam state here")
    ap.add_argument("--resume", default=None, help="start from this blob (weights or checkpoint)")
    ap.add_argument("--poisson-fraction", type=float, default=0.5,
                    help="share of crops with Poisson-Gaussian noise (the rest Gaussian)")
    ap.add_argument("--loss", choices=["mse", "sigma"], default="mse",
                    help="sigma: each crop's squared error scaled by (0.05 / sigma)^2")
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    _, noisy, sigma = make_crops(args.crops, args.size, args.seed, poisson_fraction=args.poisson_fraction)
    log(f"crops: {args.crops} x {args.size}^2 in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    X, T = build_pairs(noisy, sigma)
    log(f"step-1 targets in {time.perf_counter() - t0:.1f} s")
    net = pilot.Pilot(seed=args.seed, weights=args.resume)
    net.set_loss(args.loss, 0.05)
    t0 = time.perf_counter()
    hist = train(net, X, T, sigma, args.epochs, args.batch, args.lr, args.seed, log)
    train_s = time.perf_counter() - t0
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "wb") as f:
        f.write(net.save(False))
    if args.checkpoint:
        with open(args.checkpoint, "wb") as f:
            f.write(net.save(True))
    print(json.dumps({"crops": args.crops, "size": args.size, "epochs": args.epochs,
                      "batch": args.batch, "lr": args.lr, "seed": args.seed, "loss": args.loss,
                      "poisson_fraction": args.poisson_fraction,
                      "loss_per_epoch": hist, "train_seconds": round(train_s, 2),
                      "weights": os.path.abspath(args.out)}), flush=True)


if __name__ == "__main__":
    main()
