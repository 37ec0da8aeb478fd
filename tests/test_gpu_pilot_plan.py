"""The learned pilot inside a BM3D plan: spot checks and fallbacks.

- every block checked with a 0 dB threshold: all fall back, the device gate
  runs the plain step 1 on the whole frame, and the output is bit-identical
  to the plain plan;
- no fallback: the step-2 input is exactly the network's output;
- partial fallback: checked blocks that fall back carry the plain basic
  estimate exactly;
- an identity pilot (zero network: it leaves all the noise in) falls back on
  the under-denoising bound, and only on that bound;
- with the default options every frame gets at least one checked block (its
  central one when the one-in-16 pattern misses a small frame), so the packaged
  network never runs unchecked: on small Gaussian frames it trails step 1 by
  1-5 dB and the checks send them back to the plain step 1.
"""
import numpy as np
import pytest

from cpipe_amd import engine, pilot, synth

pytestmark = pytest.mark.gpu


def _frame(torch, name="rgb200x136", seed=4):
    clean, noisy, sigma = synth.workload(name, seed=seed)
    return clean, noisy, sigma, torch.from_numpy(noisy).cuda()


def _run(torch, src, sigma, net=None, **opts):
    c, h, w = src.shape
    plan = engine.Plan(w, h, 3)
    if net is not None:
        plan.set_pilot(net, **opts)
    out = torch.empty_like(src)
    basic = torch.empty_like(src)
    plan.run(src, out, sigma)
    plan.basic(basic)
    torch.cuda.synchronize()
    plan.check()
    rep = plan.pilot_report()
    res = out.cpu().numpy(), basic.cpu().numpy(), rep
    plan.close()
    return res


def test_full_fallback_is_bit_identical_to_plain(cuda):
    torch = cuda
    _, _, sigma, src = _frame(torch)
    plain, plain_basic, _ = _run(torch, src, sigma)
    net = pilot.Pilot(seed=1)  # untrained: a poor pilot
    out, basic, rep = _run(torch, src, sigma, net, fallback_db=0.0, under_db=0.0, check_block=64,
                           check_every=1)
    print(rep)
    assert rep["checked"] == 4 * 3 and rep["fallback"] == rep["checked"] and rep["full_fallback"]
    assert np.array_equal(out.view(np.uint32), plain.view(np.uint32))
    assert np.array_equal(basic.view(np.uint32), plain_basic.view(np.uint32))


def test_no_fallback_uses_the_network(cuda):
    torch = cuda
    _, noisy, sigma, src = _frame(torch)
    net = pilot.Pilot(seed=2)
    out, basic, rep = _run(torch, src, sigma, net, fallback_db=1e9, under_db=1e9, check_block=64,
                           check_every=2)
    assert rep["checked"] > 0 and rep["fallback"] == 0 and not rep["full_fallback"]
    # the basic estimate is the network applied to the opponent planes
    plan = engine.Plan(200, 136, 3)
    z = torch.empty_like(src)
    b = torch.empty_like(src)
    plan.step1_opponent(src, z, b, sigma)
    y = net.forward(z.unsqueeze(0).contiguous(), sigma)[0]
    torch.cuda.synchronize()
    y = y.cpu().numpy()
    c13, c23, c43 = np.float32(1 / 3), np.float32(2 / 3), np.float32(4 / 3)
    rgb = np.stack([(y[0] + y[1]) + y[2] * c23, y[0] - y[2] * c43, (y[0] - y[1]) + y[2] * c23])
    np.testing.assert_array_equal(basic, rgb)
    plan.close()


def test_partial_fallback_blocks_carry_the_plain_estimate(cuda):
    torch = cuda
    _, _, sigma, src = _frame(torch, "rgb256x192", seed=6)
    _, plain_basic, _ = _run(torch, src, sigma)
    net = pilot.Pilot(seed=3)
    out, basic, rep = _run(torch, src, sigma, net, fallback_db=0.0, under_db=0.0, check_block=64,
                           check_every=3, full_fraction=2.0)
    # every checked block falls back (threshold 0); whole-frame escalation disabled
    assert rep["fallback"] == rep["checked"] == 4 and not rep["full_fallback"]
    nbx = 4
    for b in range(12):
        if (b + b // nbx) % 3 != 1:
            continue
        x0, y0 = (b % nbx) * 64, (b // nbx) * 64
        np.testing.assert_array_equal(basic[:, y0:y0 + 64, x0:x0 + 64],
                                      plain_basic[:, y0:y0 + 64, x0:x0 + 64])


def test_identity_pilot_falls_back_on_the_under_denoising_bound(cuda):
    torch = cuda
    _, _, sigma, src = _frame(torch)
    plain, _, _ = _run(torch, src, sigma)
    net = pilot.Pilot(seed=5)
    blob = bytearray(net.save(False))
    blob[32:] = bytes(len(blob) - 32)  # zero weights and biases: pilot = noisy input
    net.load(bytes(blob))
    out, _, rep = _run(torch, src, sigma, net, check_block=64, check_every=1)
    print(rep)
    assert rep["fallback"] == rep["checked"] == 12 and rep["full_fallback"]
    assert np.array_equal(out.view(np.uint32), plain.view(np.uint32))
    # the over-smoothing bound alone (under-denoising bound disabled) keeps it
    _, _, rep = _run(torch, src, sigma, net, fallback_db=0.0, under_db=1e31, check_block=64,
                     check_every=1)
    assert rep["fallback"] == 0


def test_step2_on_the_step1_basic_reproduces_the_plain_run(cuda):
    """cpipe_bm3d_plan_step2_opponent fed the basic planes of
    cpipe_bm3d_plan_step1_opponent gives the plain run's output bit for bit, so
    a candidate pilot scored through it is scored by the Wiener step the node runs."""
    torch = cuda
    _, _, sigma, src = _frame(torch, "rgb200x136", seed=7)
    plain, _, _ = _run(torch, src, sigma)
    c, h, w = src.shape
    plan = engine.Plan(w, h, 3)
    z = torch.empty_like(src)
    basic = torch.empty_like(src)
    out = torch.empty_like(src)
    plan.step1_opponent(src, z, basic, sigma)
    plan.step2_opponent(src, basic, out, sigma)
    torch.cuda.synchronize()
    plan.check()
    plan.close()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), plain.view(np.uint32))


def test_plans_sharing_one_network_run_on_separate_streams(cuda):
    """Four plans that share one network, run at once on four streams, give the
    outputs they give one after another: each plan runs the network on its own
    activation workspace."""
    torch = cuda
    net = pilot.Pilot(seed=None, weights=pilot.DEFAULT_WEIGHTS)
    srcs, plans, outs = [], [], []
    for i in range(4):
        _, _, sigma, src = _frame(torch, "rgb1024x768", seed=20 + i)
        plan = engine.Plan(1024, 768, 3, "fp16", 0)
        plan.set_pilot(net)
        srcs.append((src, sigma))
        plans.append(plan)
        outs.append(torch.empty_like(src))
    for (src, sigma), plan, out in zip(srcs, plans, outs):
        plan.run(src, out, sigma)
    torch.cuda.synchronize()
    want = [o.cpu().numpy() for o in outs]
    for o in outs:
        o.zero_()
    streams = [torch.cuda.Stream() for _ in plans]
    s0 = torch.cuda.current_stream()
    ev = torch.cuda.Event()
    ev.record(s0)
    for (src, sigma), plan, out, st in zip(srcs, plans, outs, streams):
        st.wait_event(ev)
        plan.run(src, out, sigma, stream=st)
    torch.cuda.synchronize()
    for plan in plans:
        plan.check()
        plan.close()
    for w_, o in zip(want, outs):
        assert np.array_equal(o.cpu().numpy().view(np.uint32), w_.view(np.uint32))


@pytest.mark.parametrize("w,h,sigma", [(417, 402, 0.02), (332, 374, 0.007), (217, 372, 0.081), (256, 256, 0.05)])
def test_small_frames_are_always_checked(cuda, w, h, sigma):
    torch = cuda
    clean = synth.clean_scene(w, h, 3, seed=w + h)
    noisy = synth.add_gaussian(clean, sigma, seed=7).astype(np.float32)
    src = torch.from_numpy(noisy).cuda()
    net = pilot.Pilot(seed=None, weights=pilot.DEFAULT_WEIGHTS)
    out, _, rep = _run(torch, src, sigma, net)
    plain, _, _ = _run(torch, src, sigma)
    print(f"{w}x{h} sigma {sigma}: {rep}, psnr pilot {synth.psnr(out, clean):.2f} dB, "
          f"plain {synth.psnr(plain, clean):.2f} dB", flush=True)
    assert rep["checked"] >= 1
    assert synth.psnr(out, clean) >= synth.psnr(plain, clean) - 0.1
    net.close()
