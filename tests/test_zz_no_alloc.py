"""The entry point that allocates nothing (cpipe_bm3d_noalloc_planes).

GPU: its output is bit-identical to the plan path's (cpipe_bm3d_plan_run,
exact FP32) on two calls in a row, the second with a smaller frame, so the
second also shows that nothing the first left in the shared scratch leaks in;
a third call on a frame smaller than one patch is copied through.
CPU: frames above the stated maximum are refused before any device work, and
the largest admitted shapes pass the size check.
"""
import ctypes

import numpy as np
import pytest

from cpipe_amd import engine, synth

CPIPE_BM3D_BAD_INDEX = 5
CPIPE_BM3D_UNSUPPORTED = 8


def _call(src_ptrs, in_pitch, dst_ptrs, out_pitch, w, h, c, sigma, stream=None):
    L = engine.lib()
    return L.cpipe_bm3d_noalloc_planes(engine._ptr_array(src_ptrs), in_pitch, engine._ptr_array(dst_ptrs),
                                       out_pitch, w, h, c, float(sigma),
                                       0 if stream is None else engine._stream_handle(stream))


def _noalloc(src, dst, sigma):
    c, h, w = src.shape
    rc = _call([src[i].data_ptr() for i in range(c)], src.stride(1), [dst[i].data_ptr() for i in range(c)],
               dst.stride(1), w, h, c, sigma)
    if rc != 0:
        raise engine.EngineError(rc, engine.lib().cpipe_bm3d_last_error().decode())


def _plan(torch, src, sigma):
    c, h, w = src.shape
    plan = engine.Plan(w, h, c, "fp32", 0)
    out = torch.empty_like(src)
    plan.run(src, out, sigma)
    torch.cuda.synchronize()
    plan.check()
    plan.close()
    return out.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_matches_plan_on_calls_in_a_row(cuda):
    torch = cuda
    cases = [("rgb640x480", 1), ("gray300x200", 2)]  # the second call's frame is smaller
    for name, seed in cases:
        clean, noisy, sigma = synth.workload(name, seed=seed)
        c, h, w = noisy.shape
        print(f"no-alloc call: {name} seed {seed} sigma {sigma}", flush=True)
        src = torch.from_numpy(noisy).cuda()
        dst = torch.full_like(src, float("nan"))
        _noalloc(src, dst, sigma)
        torch.cuda.synchronize()
        got = dst.cpu().numpy()
        want = _plan(torch, src, sigma)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), name
        assert synth.psnr(got, clean) > synth.psnr(noisy, clean) + 3.0
    # smaller than one patch: the input comes back unchanged
    tiny = torch.from_numpy(np.random.default_rng(3).uniform(size=(3, 5, 7)).astype(np.float32)).cuda()
    out = torch.zeros_like(tiny)
    print("no-alloc call: 7x5x3 (copied through)", flush=True)
    _noalloc(tiny, out, 0.05)
    torch.cuda.synchronize()
    assert torch.equal(out, tiny)


def test_noalloc_refuses_frames_above_its_maximum():
    """Checked before any HIP call, so this runs without a GPU."""
    buf = np.zeros(16, np.float32)
    ptrs = [buf.ctypes.data] * 3
    for w, h in ((1025, 1024), (8193, 8), (8, 8193), (4096, 4096)):
        assert _call(ptrs, w, ptrs, w, w, h, 3, 0.05) == CPIPE_BM3D_BAD_INDEX, (w, h)
        msg = engine.lib().cpipe_bm3d_last_error().decode()
        assert "maximum" in msg and f"{w}x{h}" in msg
    # the largest admitted shapes pass the size check (without a GPU the call then
    # stops at the device lookup; on a GPU box it is not run here)
    import torch
    if not torch.cuda.is_available():
        for w, h in ((1024, 1024), (150, 6990), (8192, 128), (128, 8192)):
            rc = _call(ptrs, w, ptrs, w, w, h, 3, 0.05)
            assert rc == CPIPE_BM3D_UNSUPPORTED, (w, h, rc, engine.lib().cpipe_bm3d_last_error().decode())
