"""Batches of uneven frames (one launch per stage, hipGraph replay, autotuner)
through the C ABI.

Bar: every frame of a batch is bit-identical to the same frame run alone
through a single-frame plan of the same precision and tile shape (exact FP32:
and therefore to the CPU FP32 restatement), for eager runs and for every
replay of the captured graph (with new sigmas and new output buffers, with
eager work in between); the autotuner's work-list orders leave exact results
unchanged.
"""
import numpy as np
import pytest

from cpipe_amd import engine, synth

pytestmark = pytest.mark.gpu

SIZES = [(200, 136), (96, 80), (61, 47), (7, 5), (130, 70)]


def _frames(torch, sizes, seed=0, channels=3):
    rng = np.random.default_rng(seed)
    srcs, sigmas = [], []
    for i, (w, h) in enumerate(sizes):
        name = f"{'rgb' if channels == 3 else 'gray'}{w}x{h}"
        _, noisy, _ = synth.workload(name, seed=seed + i)
        srcs.append(torch.from_numpy(noisy).cuda())
        sigmas.append(float(rng.uniform(0.01, 0.1)))
    return srcs, sigmas


def _single(torch, src, sigma, precision, net=None):
    c, h, w = src.shape
    plan = engine.Plan(w, h, c, precision, 0)
    if net is not None and w >= 8 and h >= 8:
        plan.set_pilot(net)
    out = torch.empty_like(src)
    plan.run(src, out, sigma)
    torch.cuda.synchronize()
    plan.check()
    plan.close()
    return out.cpu().numpy()


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_frames_equal_single_plans(cuda, precision, use_graph):
    torch = cuda
    srcs, sigmas = _frames(torch, SIZES, seed=1)
    print(f"batch {SIZES} precision {precision} graph {use_graph}", flush=True)
    want = [_single(torch, s, g, precision) for s, g in zip(srcs, sigmas)]
    b = engine.Batch(SIZES, 3, precision, 0, use_graph=use_graph)
    for rep in range(3):  # eager (+ capture), then graph replays
        sig = [g * (1.0 + 0.25 * rep) for g in sigmas]
        if rep > 0:
            want = [_single(torch, s, g, precision) for s, g in zip(srcs, sig)]
        outs = [torch.empty_like(s) for s in srcs]
        b.run(srcs, outs, sig)
        torch.cuda.synchronize()
        b.check()
        for f, (o, w) in enumerate(zip(outs, want)):
            assert np.array_equal(_bits(o.cpu().numpy()), _bits(w)), (rep, f)
        _ = (srcs[0] * 2.0).sum().item()  # eager work between replays
    b.close()


# Block matching launches with >= 1024 items match two refs per lane, smaller
# ones one ref per lane. These eight frames launch together with 1197 items;
# seven have their forced last ref column as the second of a pair (tiles that
# sum every ref themselves), three a forced last ref row, and one is paired
# throughout. Each frame alone takes the one-ref path.
PAIRED_SIZES = [(700, 500), (640, 481), (604, 422), (700, 347), (520, 499), (688, 500), (580, 460),
                (661, 389)]


def test_paired_block_matching_bit_exact(cuda, oracle_lib):
    torch = cuda
    srcs, sigmas = _frames(torch, PAIRED_SIZES, seed=5, channels=1)
    print(f"paired block matching batch {PAIRED_SIZES}", flush=True)
    b = engine.Batch(PAIRED_SIZES, 1, "fp32", 0, use_graph=False)
    outs = [torch.empty_like(s) for s in srcs]
    b.run(srcs, outs, sigmas)
    torch.cuda.synchronize()
    b.check()
    b.close()
    for f, (s, o, g) in enumerate(zip(srcs, outs, sigmas)):
        ref = oracle_lib.bm3d_f32(s.cpu().numpy(), g, nthreads=16)
        assert np.array_equal(_bits(o.cpu().numpy()), _bits(ref)), PAIRED_SIZES[f]
        assert np.array_equal(_bits(o.cpu().numpy()), _bits(_single(torch, s, g, "fp32"))), PAIRED_SIZES[f]


def test_batch_grey_frames(cuda):
    torch = cuda
    sizes = [(64, 64), (33, 90)]
    srcs, sigmas = _frames(torch, sizes, seed=3, channels=1)
    want = [_single(torch, s, g, "fp32") for s, g in zip(srcs, sigmas)]
    b = engine.Batch(sizes, 1, "fp32", 0, use_graph=True)
    for _ in range(2):
        outs = [torch.empty_like(s) for s in srcs]
        b.run(srcs, outs, sigmas)
        torch.cuda.synchronize()
        for o, w in zip(outs, want):
            assert np.array_equal(_bits(o.cpu().numpy()), _bits(w))
    b.close()


def test_batch_with_learned_pilot_equals_single_plans(cuda):
    torch = cuda
    from cpipe_amd import pilot
    net = pilot.Pilot(seed=None, weights=pilot.DEFAULT_WEIGHTS)
    sizes = [(300, 200), (200, 136), (96, 80)]
    srcs, sigmas = _frames(torch, sizes, seed=5)
    want = [_single(torch, s, g, "fp16", net) for s, g in zip(srcs, sigmas)]
    b = engine.Batch(sizes, 3, "fp16", 0, use_graph=True)
    b.set_pilot(net)
    for _ in range(3):
        outs = [torch.empty_like(s) for s in srcs]
        b.run(srcs, outs, sigmas)
        torch.cuda.synchronize()
        b.check()
        for f, (o, w) in enumerate(zip(outs, want)):
            assert np.array_equal(_bits(o.cpu().numpy()), _bits(w)), f
    print(b.pilot_report())
    b.close()


def test_tuning_keeps_exact_results(cuda):
    torch = cuda
    sizes = [(200, 136), (130, 70), (96, 80)]
    srcs, sigmas = _frames(torch, sizes, seed=7)
    want = [_single(torch, s, g, "fp32") for s, g in zip(srcs, sigmas)]
    b = engine.Batch(sizes, 3, "fp32", 0, use_graph=True)
    outs = [torch.empty_like(s) for s in srcs]
    rep = b.tune(srcs, outs, sigmas, iters=1)
    print(rep, b.config())
    assert len(rep["candidates"]) == 4 and 0 <= rep["best"] < 4
    assert all(c["bm"] == "exact" for c in rep["candidates"])
    assert all(c["tile_y"] == 4 for c in rep["candidates"])  # exact plans: 16 x 4-ref tiles
    for _ in range(2):
        outs = [torch.empty_like(s) for s in srcs]
        b.run(srcs, outs, sigmas)
        torch.cuda.synchronize()
        for o, w in zip(outs, want):
            assert np.array_equal(_bits(o.cpu().numpy()), _bits(w))
    b.close()
    # a new batch of the same shape starts from the cached choice
    best = rep["candidates"][rep["best"]]
    b2 = engine.Batch(sizes, 3, "fp32", 0)
    cfg = b2.config()
    assert cfg["order"] == best["order"] and cfg["band_rows"] == best["band_rows"]
    b2.close()


def test_tuning_mixed_tile_shapes(cuda):
    torch = cuda
    sizes = [(256, 192), (96, 80)]
    srcs, sigmas = _frames(torch, sizes, seed=9)
    b = engine.Batch(sizes, 3, "fp16", 0, use_graph=True)
    outs = [torch.empty_like(s) for s in srcs]
    b.run(srcs, outs, sigmas)
    torch.cuda.synchronize()
    before = [o.cpu().numpy() for o in outs]
    rep = b.tune(srcs, outs, sigmas, iters=1)
    print(rep, b.config())
    assert sorted({c["tile_y"] for c in rep["candidates"]}) == [8, 16]
    assert sorted({c["bm"] for c in rep["candidates"]}) == ["exact", "mfma"]
    outs = [torch.empty_like(s) for s in srcs]
    b.run(srcs, outs, sigmas)
    torch.cuda.synchronize()
    b.check()
    for o, w in zip(outs, before):
        assert synth.psnr(o.cpu().numpy(), w) >= 55.0
    b.close()
