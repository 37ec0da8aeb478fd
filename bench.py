"""BM3D node throughput on one MI355X (Mpixel/s), with roofline and CPU baseline.

Workload (default): 4080x3072 RGB float32 linear planes in [0, 1] from the
seeded shape+texture generator, Poisson-Gaussian noise (gain 0.01, read 0.002),
node sigma from the noise-profile rule (0.05004), both BM3D steps. A step is
one full BM3D pass over the frame with the input already resident in HBM.

The headline `value` is the plain exact-FP32 form. The same frame is also run
in the mixed-precision forms (FP16 / BF16 MFMA group transforms) and with the
learned pilot replacing step 1; each form reports its throughput, its PSNR
against the clean frame and against the FP32 output, and, on a 512x384 crop
of the same noisy frame, its PSNR difference to the FP64 CPU oracle.
Prints one JSON line.
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
PEAK_F16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of each exact FP32 kernel (rocprofv3 FETCH_SIZE / WRITE_SIZE, rgb12mp)
TRAFFIC_FILES = {k: os.path.join(ROOT, "profiles", f"{k}_fp32_rgb12mp_traffic.json")
                 for k in ("block_match", "group_ht", "group_wiener")}


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


def host_cores():
    """CPU cores this job may use: the affinity mask, capped by a cgroup CPU quota
    when one is set (os.cpu_count() can report the whole machine)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(float(quota) / float(period))))
    except (OSError, ValueError):
        pass
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(noisy: np.ndarray, sigma: float, sample_w: int, sample_h: int):
    """The CPU FP32 BM3D (same algorithm and operation order as the GPU FP32 path,
    bit-identical output; pthreads) on every core available to the job, and on one
    core fewer (the reference Scheduler runs N-1 workers, Scheduler.cpp:116-119)."""
    from oracle import oracle
    cores = host_cores()
    crop = np.ascontiguousarray(noisy[:, :sample_h, :sample_w])
    res = {}
    for key, threads in (("all", cores), ("n_minus_1", max(1, cores - 1))):
        t0 = time.perf_counter()
        oracle.bm3d_f32(crop, sigma, nthreads=threads)
        dt = time.perf_counter() - t0
        res[key] = (threads, dt)
    threads, dt = res["all"]
    c = noisy.shape[0]
    full = sample_w == noisy.shape[2] and sample_h == noisy.shape[1]
    return {"value": round(sample_w * sample_h / dt / 1e6, 4), "unit": "Mpixel/s",
            "cores": threads, "kind": "port",
            "sample": (f"{'the full' if full else 'a ' + str(sample_w) + 'x' + str(sample_h) + ' top-left crop of the same'}"
                       f" {sample_w}x{sample_h}x{c} noisy frame, CPU FP32 BM3D (same algorithm and "
                       f"operation order, bit-identical output), {dt:.2f} s on {threads} threads"),
            "n_minus_1": {"value": round(sample_w * sample_h / res["n_minus_1"][1] / 1e6, 4),
                          "cores": res["n_minus_1"][0], "seconds": round(res["n_minus_1"][1], 2)},
            "cpu": cpu_model(), "os_cpu_count": os.cpu_count(), "flags": "gcc -O3 -march=x86-64-v3 -ffp-contract=off"}


def cpu_config0(torch):
    """Config 0 (512x512 grey, Gaussian sigma 0.05): the CPU FP32 BM3D on all
    available cores and the GPU FP32 plan, whose output must be bit-identical."""
    from oracle import oracle
    clean, noisy, sigma = synth.workload("gray512", seed=0)
    cores = host_cores()
    t0 = time.perf_counter()
    want = oracle.bm3d_f32(noisy, sigma, nthreads=cores)
    dt = time.perf_counter() - t0
    src = torch.from_numpy(noisy).cuda()
    dst = torch.empty_like(src)
    plan = engine.Plan(512, 512, 1, "fp32", 0)
    ms, _ = time_form(torch, plan, src, dst, sigma, 20, 3, torch.cuda.current_stream())
    got = dst.cpu().numpy()
    plan.close()
    return {"workload": "gray512 512x512x1 float32, Gaussian sigma 0.05, both steps",
            "cpu": {"value": round(0.262144 / dt, 3), "unit": "Mpixel/s", "seconds": round(dt, 3), "cores": cores},
            "gpu": {"value": round(0.262144 / (ms / 1e3), 1), "unit": "Mpixel/s", "ms": round(ms, 3)},
            "gpu_over_cpu": round(dt * 1e3 / ms, 1),
            "bit_exact": bool(np.array_equal(got.view(np.uint32), want.view(np.uint32))),
            "psnr_noisy_db": round(synth.psnr(noisy, clean), 3), "psnr_out_db": round(synth.psnr(got, clean), 3)}


def compat_context(torch, noisy, sigma):
    """The reference node's actual work at the frame size: the 2x2 box-mean blend
    (bm3d_step1_generator.cpp:42-56) on RGBA16F, restated on the CPU (row bands on
    every available core; the reference runs parallel(y) on N-1 workers) and the
    GPU's k_compat_rgba16f. Not comparable work (about 16 bytes per pixel)."""
    import concurrent.futures as cf
    import ctypes
    from oracle import oracle
    c, h, w = noisy.shape
    rgba = np.ones((h, w, 4), np.float32)
    rgba[..., :min(c, 3)] = np.moveaxis(noisy[:3], 0, -1)
    bits = np.ascontiguousarray(rgba.astype(np.float16).view(np.uint16))
    out = np.empty_like(bits)
    L = oracle.lib()
    cores = host_cores()
    band = ((h + cores - 1) // cores + 1) & ~1  # even heights: 2x2 blocks never straddle bands

    def run(y0):
        y1 = min(h, y0 + band)
        L.oracle_compat_rgba16f(bits.ctypes.data + y0 * w * 8, w, out.ctypes.data + y0 * w * 8, w, w, y1 - y0,
                                float(sigma), 1)

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(cores) as ex:
        list(ex.map(run, range(0, h, band)))
    dt = time.perf_counter() - t0
    src = torch.from_numpy(bits.view(np.int16)).cuda()
    dst = torch.empty_like(src)
    for _ in range(3):
        engine.compat_rgba16f(src, dst, sigma)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(20):
        engine.compat_rgba16f(src, dst, sigma)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 20
    same = bool(np.array_equal(dst.cpu().numpy().view(np.uint16), out))
    return {"work": "reference node arithmetic: 2x2 box-mean blend, RGBA16F in/out",
            "cpu": {"value": round(w * h / dt / 1e6, 1), "unit": "Mpixel/s", "seconds": round(dt, 3), "cores": cores},
            "gpu": {"value": round(w * h / (ms / 1e3) / 1e6, 1), "unit": "Mpixel/s", "ms": round(ms, 4),
                    "hbm_gbs": round(16.0 * w * h / (ms / 1e3) / 1e9, 1)},
            "bit_exact": same}


BATCH_SIZES = [(4080, 3072)] * 4 + [(2040, 1536)] * 4 + [(1024, 768)] * 4


def batch_bench(torch, steps, warmup, stream, weights=None):
    """12 frames of three sizes, per-frame Gaussian sigma ~ U(0.01, 0.1), mixed
    precision (FP16) + learned pilot: one batch (autotuned, one hipGraph launch
    per run) against the same frames through per-frame plans."""
    from cpipe_amd import pilot as pilot_mod
    rng = np.random.default_rng(4)
    sigmas = [float(v) for v in rng.uniform(0.01, 0.1, size=len(BATCH_SIZES))]
    t0 = time.perf_counter()
    cleans, srcs = [], []
    for i, (w, h) in enumerate(BATCH_SIZES):
        clean = synth.clean_scene(w, h, 3, seed=100 + i)
        noisy = synth.add_gaussian(clean, sigmas[i], seed=200 + i)
        cleans.append(clean)
        srcs.append(torch.from_numpy(noisy).cuda())
    gen_s = time.perf_counter() - t0
    net = pilot_mod.Pilot(seed=None, weights=weights or pilot_mod.DEFAULT_WEIGHTS)
    mpx = sum(w * h for w, h in BATCH_SIZES) / 1e6
    outs = [torch.empty_like(s) for s in srcs]

    def timed(fn, n):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / n * 1e3

    # per-frame launches: one plan per frame, run one after another
    plans = []
    for (w, h) in BATCH_SIZES:
        p = engine.Plan(w, h, 3, "fp16", 0)
        p.set_pilot(net)
        plans.append(p)

    def per_frame():
        for p, s, o, g in zip(plans, srcs, outs, sigmas):
            p.run(s, o, g, stream)
    for _ in range(warmup):
        per_frame()
    ms_frames = timed(per_frame, steps)
    ref = [o.cpu().numpy() for o in outs]
    for p in plans:
        p.close()

    res = {"frames": "4x4080x3072 + 4x2040x1536 + 4x1024x768 RGB, Gaussian sigma_f ~ U(0.01, 0.1), "
                     "fp16 + learned pilot", "mpixel": round(mpx, 2), "generation_s": round(gen_s, 1),
           "per_frame_launches": {"ms": round(ms_frames, 2), "value": round(mpx / (ms_frames / 1e3), 1)}}
    for graph in (False, True):
        b = engine.Batch(BATCH_SIZES, 3, "fp16", 0, use_graph=graph)
        b.set_pilot(net)
        tune = b.tune(srcs, outs, sigmas, iters=1, stream=stream) if graph else None
        for _ in range(max(1, warmup)):
            b.run(srcs, outs, sigmas, stream)
        ms = timed(lambda: b.run(srcs, outs, sigmas, stream), steps)
        b.check()
        got = [o.cpu().numpy() for o in outs]
        key = "batch_graph_tuned" if graph else "batch_eager"
        res[key] = {"ms": round(ms, 2), "value": round(mpx / (ms / 1e3), 1),
                    "speedup_vs_per_frame": round(ms_frames / ms, 3),
                    "min_psnr_vs_per_frame_db": round(min(synth.psnr(a, r) for a, r in zip(got, ref)), 2),
                    "mean_psnr_out_db": round(float(np.mean([synth.psnr(a, c) for a, c in zip(got, cleans)])), 3),
                    "spot_checks": b.pilot_report()}
        if tune:
            res[key]["config"] = b.config()
            res[key]["tune"] = tune["candidates"]
        b.close()
    return res


PILOT_FLOP_PER_PX = 2 * 9 * (4 * 32 + 4 * 32 * 32 + 32 * 3)  # 6 conv layers: 4(+level) -> 32 x5 -> 3


def form_roofline(prec, stage, bm_flop, g1_flop, g2_flop, npx):
    """Roofline of a form's kernels from its HIP-event stage times and the same
    counting rule as the headline: algorithmic FLOP per launch / launch time,
    against the dense FP16/BF16 MFMA peak (mixed-precision group transforms and
    the MFMA block matching, the learned pilot's convolutions) or the FP32 VALU
    peak (exact FP32 kernels). The dominant kernel is the longest launch."""
    mixed = prec in ("fp16", "bf16")
    peak = PEAK_F16_TFLOPS if mixed else PEAK_FP32_TFLOPS
    ks = {}
    for key, flop, st in (("block_match", bm_flop, "bm2"), ("group_ht", g1_flop, "group1"),
                          ("group_wiener", g2_flop, "group2")):
        ms = stage.get(st, 0.0)
        if ms > 0.1:  # stages skipped behind the pilot's gate take microseconds
            ach = flop / (ms / 1e3) / 1e12
            ks[key] = {"launch_ms": round(ms, 4), "flop_per_launch": flop, "achieved": round(ach, 2),
                       "peak": peak, "frac": round(ach / peak, 4), "bound": "mfma" if mixed else "valu"}
    if stage.get("pilot", 0.0) > 0.1:
        flop = float(PILOT_FLOP_PER_PX) * npx
        ach = flop / (stage["pilot"] / 1e3) / 1e12
        ks["pilot"] = {"launch_ms": round(stage["pilot"], 4), "flop_per_launch": flop, "achieved": round(ach, 2),
                       "peak": PEAK_F16_TFLOPS, "frac": round(ach / PEAK_F16_TFLOPS, 4), "bound": "mfma",
                       "note": "network + overlapped plain step-1 spot checks"}
    dom = max(ks, key=lambda k: ks[k]["launch_ms"])
    return dict(ks[dom], kernel=dom, unit="TFLOP/s", kernels=ks)


def time_form(torch, plan, src, dst, sigma, steps, warmup, stream):
    plan.set_timing(False)
    for _ in range(warmup):
        plan.run(src, dst, sigma, stream)
    torch.cuda.synchronize()
    plan.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run(src, dst, sigma, stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    plan.check()
    stage_ms, runs = plan.timing()
    plan.set_timing(False)
    return wall * 1e3 / steps, {k: v / runs for k, v in stage_ms.items() if v > 0}


def host_path_bench(noisy, sigma, runs):
    """The node's path (cpipe_bm3d_host_planes): pageable host planes in, H2D,
    BM3D, D2H, synchronise; plans are cached across calls, so the first call
    (plan creation) is reported separately. Not the headline: the headline
    keeps the frame resident in HBM."""
    import ctypes
    c, h, w = noisy.shape
    L = engine.lib()
    src = [np.ascontiguousarray(noisy[i]) for i in range(c)]
    dst = [np.empty((h, w), np.float32) for _ in range(c)]
    fp = ctypes.POINTER(ctypes.c_float)
    ins = (fp * c)(*[a.ctypes.data_as(fp) for a in src])
    outs = (fp * c)(*[a.ctypes.data_as(fp) for a in dst])
    res = {"frame": f"{w}x{h}x{c} float32 host planes (pageable)"}
    for name, prec, flags in (("fp32", 0, 0), ("learned_fp16", 2, 1)):
        if flags and c != 3:
            continue
        times = []
        for _ in range(runs + 1):
            t0 = time.perf_counter()
            rc = L.cpipe_bm3d_host_planes(ctypes.cast(ins, ctypes.c_void_p), w,
                                          ctypes.cast(outs, ctypes.c_void_p), w, w, h, c,
                                          float(sigma), prec, 0, flags)
            times.append((time.perf_counter() - t0) * 1e3)
            if rc != 0:
                raise engine.EngineError(rc, L.cpipe_bm3d_last_error().decode())
        ms = float(np.median(times[1:]))
        res[name] = {"first_call_ms": round(times[0], 2), "ms": round(ms, 2),
                     "value": round(w * h / (ms / 1e3) / 1e6, 2), "unit": "Mpixel/s"}
    L.cpipe_bm3d_host_release()
    return res


def node_chain_bench(noisy, sigma, runs):
    """The classic pipeline's chain bm3d -> wavelet_bayes_shrink -> precision_convert
    (to RGBA8) through the plugin ABI with the minimal host, on one RGBA16F frame:
    host buffers (each node copies in and out over PCIe) against device-resident
    buffers (the "hip_buffer" suite: device addresses, no host copy). The first
    chain of each kind (plan creation) is left out of the median."""
    from cpipe_amd import host
    c, h, w = noisy.shape
    if c != 3:
        return None
    rgba = np.ones((h, w, 4), np.float32)
    rgba[..., :3] = np.moveaxis(noisy, 0, -1)
    bits = rgba.astype(np.float16).view(np.uint16)
    nodes = (host.find_node(), host.find_node("com.cpipe.denoise.wavelet_bayes_shrink"),
             host.find_node("com.cpipe.precision_convert"))
    res = {"frame": f"{w}x{h} RGBA16F -> bm3d -> wavelet_bayes_shrink -> precision_convert RGBA8"}
    for form, params in (("fp32", {"mode": "bm3d", "sigma": float(sigma)}),
                         ("learned_fp16", {"mode": "bm3d", "sigma": float(sigma), "precision": "fp16",
                                           "pilot": "learned"})):
        for kind, dev in (("host_buffers", None), ("device_buffers", 0)):
            a = host.Buffer(host.FORMAT_RGBA16F, w, h, device=dev)
            a.write(bits)
            mid = host.Buffer(host.FORMAT_RGBA16F, w, h, device=dev)
            wav = host.Buffer(host.FORMAT_RGBA16F, w, h, device=dev)
            out = host.Buffer(host.FORMAT_RGBA8, w, h, device=dev)
            chain, per = [], []
            for _ in range(runs + 1):
                t = [time.perf_counter()]
                for d, i, o, prm in ((nodes[0], a, mid, params), (nodes[1], mid, wav, {}),
                                     (nodes[2], wav, out, {"target_format": "R8G8B8A8_UNORM"})):
                    rc = host.run_node(d, [i], [o], prm)
                    if rc != host.CPIPE_OK:
                        raise RuntimeError(f"node chain: status {rc}: {host.host_log()}")
                    t.append(time.perf_counter())
                chain.append((t[-1] - t[0]) * 1e3)
                per.append([(t[k + 1] - t[k]) * 1e3 for k in range(3)])
            ms = float(np.median(chain[1:]))
            nodes_ms = np.median(np.array(per[1:]), axis=0)
            res.setdefault(form, {})[kind] = {
                "first_chain_ms": round(chain[0], 2), "ms": round(ms, 2),
                "node_ms": {"bm3d": round(float(nodes_ms[0]), 2), "wavelet": round(float(nodes_ms[1]), 3),
                            "convert": round(float(nodes_ms[2]), 3)},
                "value": round(w * h / (ms / 1e3) / 1e6, 2), "unit": "Mpixel/s"}
            del a, mid, wav, out
        f = res[form]
        f["device_over_host"] = round(f["host_buffers"]["ms"] / f["device_buffers"]["ms"], 3)
    engine.lib().cpipe_bm3d_host_release()
    return res


def wavelet_bench(torch, noisy, stream, runs, want_cpu):
    """The downstream node com.cpipe.denoise.wavelet_bayes_shrink on the same frame as
    RGBA16F (alpha 1), resident in HBM: one kernel launch per run, timed with HIP
    events on the launch stream. HBM-bound: 16 algorithmic bytes per pixel (8 read,
    8 written). CPU baseline: the oracle (single thread) on a 1024x768 crop, whose
    output the GPU's must equal bit for bit."""
    c, h, w = noisy.shape
    if c != 3:
        return None
    rgba = np.empty((h, w, 4), np.float32)
    rgba[..., :3] = np.moveaxis(noisy, 0, -1)
    rgba[..., 3] = 1.0
    bits = rgba.astype(np.float16).view(np.uint16)
    src = torch.from_numpy(bits.view(np.int16)).cuda()
    dst = torch.empty_like(src)
    for _ in range(3):
        engine.wavelet_rgba16f(src, dst, 1.0, stream)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for _ in range(runs):
        engine.wavelet_rgba16f(src, dst, 1.0, stream)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / runs
    gbs = 16.0 * w * h / (ms / 1e3) / 1e9
    res = {"node": "com.cpipe.denoise.wavelet_bayes_shrink", "frame": f"{w}x{h} RGBA16F (alpha 1), chroma_strength 1",
           "value": round(w * h / (ms / 1e3) / 1e6, 1), "unit": "Mpixel/s", "ms": round(ms, 4),
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": None,
                        "bytes_per_launch": 16 * w * h}}
    tfile = os.path.join(ROOT, "profiles", "wavelet_rgba16f_rgb12mp_traffic.json")
    if (w, h) == (4080, 3072) and os.path.exists(tfile):
        with open(tfile) as f:
            res["roofline"]["traffic"] = json.load(f).get("hbm_bytes_per_launch")
    if want_cpu:
        from oracle import oracle
        cw, ch = min(1024, w), min(768, h)
        crop = np.ascontiguousarray(bits[:ch, :cw])
        t = time.perf_counter()
        want = oracle.wavelet_rgba16f(crop, 1.0)
        dt = time.perf_counter() - t
        csrc = torch.from_numpy(crop.view(np.int16)).cuda()
        cdst = torch.empty_like(csrc)
        engine.wavelet_rgba16f(csrc, cdst, 1.0, stream)
        torch.cuda.synchronize()
        res["bit_exact_vs_oracle_crop"] = bool(np.array_equal(cdst.cpu().numpy().view(np.uint16), want))
        res["cpu_baseline"] = {"value": round(cw * ch / dt / 1e6, 2), "unit": "Mpixel/s", "cores": 1,
                               "kind": "port", "sample": f"{cw}x{ch} top-left crop, oracle_wavelet_rgba16f, "
                                                         f"{dt:.3f} s"}
        res["gpu_over_cpu"] = round(res["value"] / res["cpu_baseline"]["value"], 1)
    return res


def convert_bench(torch, noisy, stream, runs):
    """The precision-convert node's kernels on the same frame (resident in HBM), one
    launch per run timed with HIP events: R16G16B16A16_SFLOAT -> R8G8B8A8_UNORM
    (12 bytes per pixel) and R32_SFLOAT -> R16G16B16A16_SFLOAT (12 bytes per pixel),
    both HBM-bound."""
    c, h, w = noisy.shape
    rgba = np.ones((h, w, 4), np.float32)
    rgba[..., :c] = np.moveaxis(noisy, 0, -1)
    f16 = torch.from_numpy(rgba.astype(np.float16).view(np.int16)).cuda()
    grey = torch.from_numpy(np.ascontiguousarray(noisy[0])).cuda()
    cases = {"rgba16f_to_rgba8": (f16, "R16G16B16A16_SFLOAT", torch.empty((h, w, 4), dtype=torch.uint8, device="cuda"),
                                  "R8G8B8A8_UNORM"),
             "f32_to_rgba16f": (grey, "R32_SFLOAT", torch.empty((h, w, 4), dtype=torch.int16, device="cuda"),
                                "R16G16B16A16_SFLOAT")}
    res = {"node": "com.cpipe.precision_convert", "frame": f"{w}x{h}"}
    for name, (src, sf, dst, df) in cases.items():
        for _ in range(3):
            engine.convert(src, sf, dst, df, stream)
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0.record(stream)
        for _ in range(runs):
            engine.convert(src, sf, dst, df, stream)
        t1.record(stream)
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / runs
        gbs = 12.0 * w * h / (ms / 1e3) / 1e9
        res[name] = {"value": round(w * h / (ms / 1e3) / 1e6, 1), "unit": "Mpixel/s", "ms": round(ms, 4),
                     "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS,
                                  "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": None,
                                  "bytes_per_launch": 12 * w * h}}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rgb12mp")
    ap.add_argument("--forms", default="fp16,bf16,fp16_exactbm,learned_fp16,learned_fp32",
                    help="extra forms besides the plain FP32 headline ('' for none): precision "
                         "fp16|bf16|fp32, optional learned_ (pilot) and _exactbm (exact FP32 block "
                         "matching in a mixed-precision plan, whose default is the MFMA cross term)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-oracle crop check")
    ap.add_argument("--no-batch", action="store_true", help="skip the 12-frame batch leg")
    ap.add_argument("--no-noise-map", action="store_true",
                    help="skip the noise-profile sigma map leg")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer (node path) and node-chain legs")
    ap.add_argument("--no-wavelet", action="store_true",
                    help="skip the downstream wavelet BayesShrink and precision-convert node legs")
    ap.add_argument("--cpu-sample", default="full",
                    help="CPU baseline sample: 'full' (the whole frame) or WxH (top-left crop)")
    ap.add_argument("--fp64-crop", default="512x384")
    ap.add_argument("--pilot", default=None,
                    help="learned-pilot weight blob (default: the packaged weights)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the multi-GPU path is not implemented (asked for "
              f"--gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import torch
    dev = torch.device("cuda:0")
    clean, noisy, sigma = synth.workload(args.workload, seed=0)
    c, h, w = noisy.shape
    print(f"# workload {args.workload}: {w}x{h}x{c} sigma={sigma:.5f}", file=sys.stderr, flush=True)
    src = torch.from_numpy(noisy).to(dev)
    stream = torch.cuda.current_stream()
    psnr_in = synth.psnr(noisy, clean)

    # ---- headline: plain exact FP32
    plan = engine.Plan(w, h, c, "fp32", 0)
    dst = torch.empty_like(src)
    ms_per_step, per_stage = time_form(torch, plan, src, dst, sigma, args.steps, args.warmup, stream)
    mpx = w * h / (ms_per_step / 1e3) / 1e6
    print(f"# fp32 {ms_per_step:.3f} ms; per-stage ms: "
          f"{json.dumps({k: round(v, 3) for k, v in per_stage.items()})}", file=sys.stderr, flush=True)
    out32 = dst.cpu().numpy()
    psnr_out = synth.psnr(out32, clean)
    _, cnt1 = plan.matches(1)
    _, cnt2 = plan.matches(2)
    cnt1 = cnt1.cpu().numpy()
    cnt2 = cnt2.cpu().numpy()
    plan.close()

    cand = bm_candidates(w, h)
    bm_flop = 192.0 * cand
    g1_flop = group_flops(cnt1, c, 1)
    g2_flop = group_flops(cnt2, c, 2)
    kernels = {
        "block_match": (bm_flop, (per_stage["bm1"] + per_stage["bm2"]) / 2),
        "group_ht": (g1_flop, per_stage["group1"]),
        "group_wiener": (g2_flop, per_stage["group2"]),
    }
    # the dominant kernel is the longest single launch (the Wiener groups); block
    # matching's two launches per frame tie with it in total, so every exact
    # kernel's figures are listed as well
    dominant = max(kernels, key=lambda k: kernels[k][1])
    total_flop = 2 * bm_flop + g1_flop + g2_flop

    def traffic_of(name):
        tfile = TRAFFIC_FILES.get(name)
        if args.workload == "rgb12mp" and tfile and os.path.exists(tfile):
            with open(tfile) as f:
                return json.load(f).get("hbm_bytes_per_launch")
        return None

    every = {}
    for name, (kflop, kms) in kernels.items():
        ach = kflop / (kms / 1e3) / 1e12
        every[name] = {"achieved": round(ach, 3), "frac": round(ach / PEAK_FP32_TFLOPS, 4),
                       "launch_ms": round(kms, 4), "flop_per_launch": kflop, "traffic": traffic_of(name)}
    flop, ms = kernels[dominant]
    achieved = flop / (ms / 1e3) / 1e12
    roofline = {"bound": "valu", "kernel": dominant, "achieved": round(achieved, 3),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic_of(dominant),
                "flop_per_launch": flop, "launch_ms": round(ms, 4),
                "end_to_end_tflops": round(total_flop / (ms_per_step / 1e3) / 1e12, 3),
                "kernels": every}

    # ---- other forms on the same frame
    forms = {}
    names = [f for f in args.forms.split(",") if f]
    net = None
    if any(f.startswith("learned") for f in names) and c == 3:
        from cpipe_amd import pilot as pilot_mod
        blob = args.pilot or pilot_mod.DEFAULT_WEIGHTS
        net = pilot_mod.Pilot(seed=None, weights=blob)
    def make_plan(name, pw, ph):
        toks = name.split("_")
        prec = next(t for t in toks if t in ("fp32", "fp16", "bf16"))
        pl = engine.Plan(pw, ph, c, prec, 0)
        if "exactbm" in toks:
            pl.set_block_matching("exact")
        if "learned" in toks:
            pl.set_pilot(net)
        return pl, prec

    for name in names:
        learned = name.startswith("learned")
        if learned and net is None:
            continue
        fplan, prec = make_plan(name, w, h)
        fdst = torch.empty_like(src)
        fms, fstage = time_form(torch, fplan, src, fdst, sigma, args.steps, args.warmup, stream)
        fout = fdst.cpu().numpy()
        entry = {"value": round(w * h / (fms / 1e3) / 1e6, 3), "unit": "Mpixel/s",
                 "ms_per_step": round(fms, 3), "precision": prec,
                 "psnr_out_db": round(synth.psnr(fout, clean), 3),
                 "delta_psnr_db_vs_fp32": round(synth.psnr(fout, clean) - psnr_out, 4),
                 "psnr_vs_fp32_db": round(synth.psnr(fout, out32), 2),
                 "max_abs_vs_fp32": round(synth.max_abs(fout, out32), 5),
                 "speedup_vs_fp32": round(ms_per_step / fms, 3),
                 "stage_ms": {k: round(v, 4) for k, v in fstage.items()}}
        if learned:
            entry["spot_checks"] = fplan.pilot_report()
        entry["roofline"] = form_roofline(prec, fstage, bm_flop, g1_flop, g2_flop, w * h)
        fplan.close()
        forms[name] = entry
        print(f"# {name}: {json.dumps(entry)}", file=sys.stderr, flush=True)

    # ---- noise-profile sigma map: the workload's own profile (read 0.002, gain 0.01)
    noise_map = None
    if not args.no_noise_map and args.workload.startswith(("rgb", "gray")):
        beta, lam = 0.002 ** 2, 0.01
        if args.workload.startswith("gray") or args.workload == "gray512":
            beta, lam = 0.05 ** 2, 0.0  # Gaussian workloads: a signal-independent profile
        noise_map = {"profile": {"beta": beta, "lambda": lam}}
        for prec in ("fp32", "fp16"):
            mplan = engine.Plan(w, h, c, prec, 0)
            mplan.set_noise_profile(beta, lam)
            mdst = torch.empty_like(src)
            mms, mstage = time_form(torch, mplan, src, mdst, sigma, args.steps, args.warmup, stream)
            mout = mdst.cpu().numpy()
            scalar_psnr = psnr_out if prec == "fp32" else forms.get("fp16", {}).get("psnr_out_db")
            entry = {"value": round(w * h / (mms / 1e3) / 1e6, 3), "unit": "Mpixel/s",
                     "ms_per_step": round(mms, 3), "psnr_out_db": round(synth.psnr(mout, clean), 3),
                     "psnr_scalar_sigma_db": None if scalar_psnr is None else round(scalar_psnr, 3),
                     "stage_ms": {k: round(v, 4) for k, v in mstage.items()}}
            if prec == "fp32" and not args.no_cpu:
                from oracle import oracle
                cw, ch = min(512, w), min(384, h)
                cn = np.ascontiguousarray(noisy[:, :ch, :cw])
                csrc = torch.from_numpy(cn).to(dev)
                cplan = engine.Plan(cw, ch, c, "fp32", 0)
                cplan.set_noise_profile(beta, lam)
                cdst = torch.empty_like(csrc)
                cplan.run(csrc, cdst, sigma)
                torch.cuda.synchronize()
                cplan.check()
                want = oracle.bm3d_f32_profile(cn, beta, lam, nthreads=min(16, os.cpu_count() or 1))
                entry["crop_bit_exact_vs_oracle"] = bool(np.array_equal(cdst.cpu().numpy().view(np.uint32),
                                                                        want.view(np.uint32)))
                cplan.close()
            mplan.close()
            noise_map[prec] = entry
        print(f"# noise map: {json.dumps(noise_map)}", file=sys.stderr, flush=True)

    # ---- FP64 oracle on a crop: delta PSNR of every form
    fp64 = None
    if not args.no_fp64:
        from oracle import oracle
        cw, ch = (int(v) for v in args.fp64_crop.split("x"))
        cw, ch = min(cw, w), min(ch, h)
        cn = np.ascontiguousarray(noisy[:, :ch, :cw])
        cc = clean[:, :ch, :cw]
        t0 = time.perf_counter()
        o64 = oracle.bm3d_f64(cn, sigma)
        t64 = time.perf_counter() - t0
        p64 = synth.psnr(o64, cc)
        csrc = torch.from_numpy(cn).to(dev)
        fp64 = {"crop": f"{cw}x{ch}x{c} top-left of the same noisy frame", "psnr_fp64_db": round(p64, 4),
                "oracle_seconds": round(t64, 2), "forms": {}}
        for name in ["fp32"] + names:
            learned = name.startswith("learned")
            if learned and net is None:
                continue
            cplan, _ = make_plan(name, cw, ch)
            cdst = torch.empty_like(csrc)
            cplan.run(csrc, cdst, sigma)
            torch.cuda.synchronize()
            cplan.check()
            co = cdst.cpu().numpy()
            cplan.close()
            fp64["forms"][name] = {"delta_psnr_db": round(synth.psnr(co, cc) - p64, 4),
                                   "psnr_vs_fp64_db": round(synth.psnr(co, o64), 2),
                                   "max_abs_vs_fp64": round(synth.max_abs(co, o64), 5)}
        print(f"# fp64 crop: {json.dumps(fp64)}", file=sys.stderr, flush=True)

    # ---- the node's path: host (pageable) planes in and out, PCIe copies included
    host_path = node_chain = None
    if not args.no_host:
        host_path = host_path_bench(noisy, sigma, 3)
        print(f"# host path: {json.dumps(host_path)}", file=sys.stderr, flush=True)
        node_chain = node_chain_bench(noisy, sigma, 3)
        print(f"# node chain: {json.dumps(node_chain)}", file=sys.stderr, flush=True)

    wavelet = convert = None
    if not args.no_wavelet:
        wavelet = wavelet_bench(torch, noisy, stream, 20, not args.no_cpu)
        print(f"# wavelet node: {json.dumps(wavelet)}", file=sys.stderr, flush=True)
        convert = convert_bench(torch, noisy, stream, 20)
        print(f"# precision_convert node: {json.dumps(convert)}", file=sys.stderr, flush=True)

    batch = None
    if not args.no_batch and c == 3:
        print("# batch of 12 frames ...", file=sys.stderr, flush=True)
        batch = batch_bench(torch, max(2, args.steps // 3), 1, stream, args.pilot)
        print(f"# batch: {json.dumps(batch)}", file=sys.stderr, flush=True)

    cpu = config0 = compat = None
    if not args.no_cpu:
        sw, sh = (w, h) if args.cpu_sample == "full" else (int(v) for v in args.cpu_sample.split("x"))
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(noisy, sigma, min(sw, w), min(sh, h))
        print(f"# cpu baseline: {json.dumps(cpu)}", file=sys.stderr, flush=True)
        config0 = cpu_config0(torch)
        print(f"# config 0: {json.dumps(config0)}", file=sys.stderr, flush=True)
        compat = compat_context(torch, noisy, sigma)
        print(f"# compat: {json.dumps(compat)}", file=sys.stderr, flush=True)

    line = {
        "metric": "BM3D Mpixel/s at 12.5MP RGB (plain FP32, both steps)",
        "value": round(mpx, 3), "unit": "Mpixel/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload} {w}x{h}x{c} float32 planes, Poisson-Gaussian "
                               f"gain 0.01 read 0.002, sigma {sigma:.5f}, both steps",
                   "precision": "fp32", "frames": 1},
        "quality": {"psnr_noisy_db": round(psnr_in, 3), "psnr_out_db": round(psnr_out, 3)},
        "stage_ms": {k: round(v, 4) for k, v in per_stage.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "config0_gray512": config0,
        "compat_step1": compat,
        "forms": forms,
        "fp64_crop": fp64,
        "batch12": batch,
        "host_path": host_path,
        "node_chain": node_chain,
        "noise_map": noise_map,
        "wavelet": wavelet,
        "precision_convert": convert,
    }
    if cpu:
        line["gpu_over_cpu"] = round(mpx / cpu["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
