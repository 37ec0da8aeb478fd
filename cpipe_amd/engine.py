"""Python binding of the MI355X BM3D engine (include/cpipe_bm3d.h) via ctypes.

torch is used only for device memory and streams. Every call goes to the
native library cpipe_amd/lib/libcpipe_bm3d.so; if it is missing or no HIP
device is visible, the calls raise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
ENGINE_SO = os.path.join(PKG, "lib", "libcpipe_bm3d.so")

PRECISION_FP32 = 0
PRECISION_BF16 = 1
PRECISION_FP16 = 2
_PRECISIONS = {"fp32": PRECISION_FP32, "bf16": PRECISION_BF16, "fp16": PRECISION_FP16}

STAGES = ("color", "bm1", "group1", "merge1", "bm2", "group2", "merge2", "pilot")

# Every symbol the C ABI header declares (checked by the CPU test suite).
EXPORTS = (
    "cpipe_bm3d_last_error", "cpipe_bm3d_device_count", "cpipe_bm3d_plan_create",
    "cpipe_bm3d_plan_destroy", "cpipe_bm3d_plan_run", "cpipe_bm3d_plan_run_rgba16f",
    "cpipe_bm3d_plan_basic", "cpipe_bm3d_plan_matches", "cpipe_bm3d_plan_ref_grid",
    "cpipe_bm3d_plan_check", "cpipe_bm3d_plan_debug_flag", "cpipe_bm3d_plan_step1_opponent",
    "cpipe_bm3d_plan_step2_opponent", "cpipe_bm3d_plan_set_pilot",
    "cpipe_bm3d_plan_pilot_report", "cpipe_bm3d_pilot_create", "cpipe_bm3d_pilot_destroy",
    "cpipe_bm3d_pilot_init", "cpipe_bm3d_pilot_save", "cpipe_bm3d_pilot_load",
    "cpipe_bm3d_pilot_forward", "cpipe_bm3d_pilot_gradients", "cpipe_bm3d_pilot_train_step",
    "cpipe_bm3d_pilot_set_loss",
    "cpipe_bm3d_pilot_blob_size", "cpipe_bm3d_pilot_param_count", "cpipe_bm3d_pilot_last_error",
    "cpipe_bm3d_plan_set_timing", "cpipe_bm3d_plan_timing", "cpipe_bm3d_compat_rgba16f",
    "cpipe_bm3d_host_planes", "cpipe_bm3d_host_rgba16f", "cpipe_bm3d_host_release",
    "cpipe_bm3d_device_planes", "cpipe_bm3d_device_rgba16f", "cpipe_bm3d_noalloc_planes",
    "cpipe_dng_noise_profile",
    "cpipe_bm3d_pilot_embedded", "cpipe_bm3d_node_descriptor",
    "cpipe_registry_entries", "cpipe_link_builtin_denoise_bm3d", "cpipe_bm3d_batch_create",
    "cpipe_bm3d_batch_destroy", "cpipe_bm3d_batch_run", "cpipe_bm3d_batch_tune",
    "cpipe_bm3d_batch_config", "cpipe_bm3d_plan_set_block_matching",
    "cpipe_bm3d_plan_set_noise_profile", "cpipe_bm3d_host_noise_profile",
    "cpipe_bm3d_batch_set_noise_profiles",
    # include/cpipe_wavelet.h and the second node of the registry section
    "cpipe_wavelet_rgba16f", "cpipe_wavelet_host_rgba16f", "cpipe_wavelet_last_error",
    "cpipe_wavelet_node_descriptor", "cpipe_link_builtin_denoise_wavelet_bayes_shrink",
    # include/cpipe_convert.h and the third node
    "cpipe_convert", "cpipe_convert_host", "cpipe_convert_supported", "cpipe_convert_last_error",
    "cpipe_convert_node_descriptor", "cpipe_link_builtin_precision_convert",
)


class EngineError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"cpipe_bm3d status {status}: {msg}")
        self.status = status


class PlanDesc(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("channels", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("device", ctypes.c_int32)]


class BatchDesc(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("n_frames", ctypes.c_int32),
                ("widths", ctypes.POINTER(ctypes.c_int32)), ("heights", ctypes.POINTER(ctypes.c_int32)),
                ("channels", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("device", ctypes.c_int32), ("use_graph", ctypes.c_int32)]


TUNE_MAX = 32


class TuneReport(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("best", ctypes.c_int32),
                ("tile_y", ctypes.c_int32 * TUNE_MAX), ("order", ctypes.c_int32 * TUNE_MAX),
                ("band_rows", ctypes.c_int32 * TUNE_MAX), ("bm_mfma", ctypes.c_int32 * TUNE_MAX),
                ("ms", ctypes.c_float * TUNE_MAX)]


_lib = None


def lib() -> ctypes.CDLL:
    """Load the native engine; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ENGINE_SO):
        raise RuntimeError(f"native engine missing: {ENGINE_SO} (run cpipe_amd.build.build())")
    L = ctypes.CDLL(ENGINE_SO)
    P, I, F, S = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_size_t
    L.cpipe_bm3d_last_error.restype = ctypes.c_char_p
    L.cpipe_bm3d_device_count.restype = I
    L.cpipe_bm3d_plan_create.argtypes = [ctypes.POINTER(PlanDesc), ctypes.POINTER(P)]
    L.cpipe_bm3d_plan_destroy.argtypes = [P]
    L.cpipe_bm3d_plan_destroy.restype = None
    L.cpipe_bm3d_plan_run.argtypes = [P, P, S, P, S, F, P]
    L.cpipe_bm3d_plan_run_rgba16f.argtypes = [P, P, S, P, S, F, P]
    L.cpipe_bm3d_plan_basic.argtypes = [P, P, S, P]
    L.cpipe_bm3d_plan_matches.argtypes = [P, I, P, P, P]
    L.cpipe_bm3d_plan_ref_grid.argtypes = [P, ctypes.POINTER(I), ctypes.POINTER(I)]
    L.cpipe_bm3d_plan_check.argtypes = [P]
    L.cpipe_bm3d_plan_debug_flag.argtypes = [P, P]
    L.cpipe_bm3d_plan_step1_opponent.argtypes = [P, P, S, F, P, P, S, P]
    L.cpipe_bm3d_plan_step2_opponent.argtypes = [P, P, S, F, P, S, P, S, P]
    L.cpipe_bm3d_plan_set_pilot.argtypes = [P, P, P]
    L.cpipe_bm3d_plan_pilot_report.argtypes = [P, ctypes.POINTER(I), ctypes.POINTER(I),
                                               ctypes.POINTER(F), ctypes.POINTER(I)]
    L.cpipe_bm3d_plan_set_timing.argtypes = [P, I]
    L.cpipe_bm3d_plan_timing.argtypes = [P, ctypes.POINTER(F), I, ctypes.POINTER(I)]
    L.cpipe_bm3d_compat_rgba16f.argtypes = [P, S, P, S, I, I, F, I, P]
    L.cpipe_dng_noise_profile.argtypes = [P, S, P, P, I, ctypes.POINTER(I)]
    L.cpipe_bm3d_noalloc_planes.argtypes = [P, S, P, S, I, I, I, F, P]
    L.cpipe_bm3d_host_planes.argtypes = [P, S, P, S, I, I, I, F, I, I, I]
    L.cpipe_bm3d_host_rgba16f.argtypes = [P, S, P, S, I, I, F, I, I, I, I]
    L.cpipe_bm3d_device_planes.argtypes = [P, S, P, S, I, I, I, F, I, I, I, P]
    L.cpipe_bm3d_device_rgba16f.argtypes = [P, S, P, S, I, I, F, I, I, I, I, P]
    L.cpipe_bm3d_host_release.argtypes = []
    L.cpipe_bm3d_host_release.restype = None
    L.cpipe_bm3d_pilot_embedded.argtypes = [ctypes.POINTER(ctypes.c_size_t)]
    L.cpipe_bm3d_pilot_embedded.restype = ctypes.c_void_p
    L.cpipe_bm3d_node_descriptor.restype = P
    L.cpipe_bm3d_batch_create.argtypes = [ctypes.POINTER(BatchDesc), ctypes.POINTER(P)]
    L.cpipe_bm3d_batch_destroy.argtypes = [P]
    L.cpipe_bm3d_batch_destroy.restype = None
    L.cpipe_bm3d_batch_run.argtypes = [P, P, P, P, P, P, P]
    L.cpipe_bm3d_batch_tune.argtypes = [P, P, P, P, P, P, I, P, ctypes.POINTER(TuneReport)]
    L.cpipe_bm3d_batch_config.argtypes = [P, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I),
                                          ctypes.POINTER(I)]
    L.cpipe_bm3d_plan_set_block_matching.argtypes = [P, I]
    L.cpipe_bm3d_plan_set_noise_profile.argtypes = [P, P, P, I]
    L.cpipe_bm3d_host_noise_profile.argtypes = [P, P, I]
    L.cpipe_bm3d_batch_set_noise_profiles.argtypes = [P, P, P, I]
    L.cpipe_wavelet_rgba16f.argtypes = [P, S, P, S, I, I, F, P]
    L.cpipe_wavelet_host_rgba16f.argtypes = [P, S, P, S, I, I, F]
    L.cpipe_wavelet_last_error.restype = ctypes.c_char_p
    L.cpipe_wavelet_node_descriptor.restype = P
    L.cpipe_convert.argtypes = [I, P, S, I, P, S, I, I, P]
    L.cpipe_convert_host.argtypes = [I, P, S, I, P, S, I, I]
    L.cpipe_convert_supported.argtypes = [I, I]
    L.cpipe_convert_last_error.restype = ctypes.c_char_p
    L.cpipe_convert_node_descriptor.restype = P
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise EngineError(rc, lib().cpipe_bm3d_last_error().decode(errors="replace"))


def device_count() -> int:
    return int(lib().cpipe_bm3d_device_count())


def _ptr_array(ptrs):
    return (ctypes.c_void_p * 3)(*(list(ptrs) + [0] * (3 - len(ptrs))))


def _stream_handle(stream) -> int:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class PilotOpts(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("fallback_db", ctypes.c_float),
                ("check_block", ctypes.c_int32), ("check_every", ctypes.c_int32),
                ("full_fraction", ctypes.c_float), ("under_db", ctypes.c_float)]


class Plan:
    """Device scratch for one image size; runs both BM3D steps on device planes."""

    def __init__(self, width: int, height: int, channels: int = 3, precision="fp32",
                 device: int = 0):
        if isinstance(precision, str):
            precision = _PRECISIONS[precision]
        self.width, self.height, self.channels = int(width), int(height), int(channels)
        self.precision, self.device = int(precision), int(device)
        d = PlanDesc(ctypes.sizeof(PlanDesc), self.width, self.height, self.channels,
                     self.precision, self.device)
        h = ctypes.c_void_p()
        _check(lib().cpipe_bm3d_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().cpipe_bm3d_plan_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    @property
    def ref_grid(self):
        nx, ny = ctypes.c_int(), ctypes.c_int()
        _check(lib().cpipe_bm3d_plan_ref_grid(self._h, ctypes.byref(nx), ctypes.byref(ny)))
        return nx.value, ny.value

    def run_ptrs(self, in_ptrs, in_pitch, out_ptrs, out_pitch, sigma, stream=None) -> None:
        _check(lib().cpipe_bm3d_plan_run(self._h, _ptr_array(in_ptrs), in_pitch,
                                         _ptr_array(out_ptrs), out_pitch, float(sigma),
                                         _stream_handle(stream)))

    def run(self, src, dst, sigma, stream=None) -> None:
        """src, dst: (C, H, W) float32 CUDA tensors with a unit stride along W."""
        self._check_planes(src)
        self._check_planes(dst)
        self.run_ptrs([src[c].data_ptr() for c in range(self.channels)], src.stride(1),
                      [dst[c].data_ptr() for c in range(self.channels)], dst.stride(1), sigma,
                      stream)

    def run_rgba16f(self, src, dst, sigma, stream=None) -> None:
        """src, dst: (H, W, 4) int16/float16 CUDA tensors (interleaved RGBA16F)."""
        for t in (src, dst):
            if tuple(t.shape) != (self.height, self.width, 4) or t.stride(2) != 1 or t.stride(1) != 4:
                raise ValueError("expected (H, W, 4) interleaved RGBA16F")
        _check(lib().cpipe_bm3d_plan_run_rgba16f(self._h, src.data_ptr(), src.stride(0) // 4,
                                                 dst.data_ptr(), dst.stride(0) // 4, float(sigma),
                                                 _stream_handle(stream)))

    def basic(self, dst, stream=None) -> None:
        self._check_planes(dst)
        _check(lib().cpipe_bm3d_plan_basic(self._h, _ptr_array([dst[c].data_ptr() for c in range(self.channels)]),
                                           dst.stride(1), _stream_handle(stream)))

    def matches(self, step: int, stream=None):
        import torch
        nx, ny = self.ref_grid
        pos = torch.zeros((nx * ny, 16), dtype=torch.int32, device=f"cuda:{self.device}")
        cnt = torch.zeros((nx * ny,), dtype=torch.uint8, device=f"cuda:{self.device}")
        _check(lib().cpipe_bm3d_plan_matches(self._h, int(step), pos.data_ptr(), cnt.data_ptr(),
                                             _stream_handle(stream)))
        return pos, cnt

    def step1_opponent(self, src, noisy_out, basic_out, sigma, stream=None) -> None:
        """Plain step 1 for training pairs: opponent noisy and basic planes (3, H, W)."""
        self._check_planes(src)
        self._check_planes(noisy_out)
        self._check_planes(basic_out)
        _check(lib().cpipe_bm3d_plan_step1_opponent(
            self._h, _ptr_array([src[c].data_ptr() for c in range(3)]), src.stride(1), float(sigma),
            _ptr_array([noisy_out[c].data_ptr() for c in range(3)]),
            _ptr_array([basic_out[c].data_ptr() for c in range(3)]), noisy_out.stride(1),
            _stream_handle(stream)))

    def step2_opponent(self, src, basic, out, sigma, stream=None) -> None:
        """Step 2 on a given opponent basic estimate (3, H, W): the output a pilot leads to."""
        self._check_planes(src)
        self._check_planes(basic)
        self._check_planes(out)
        _check(lib().cpipe_bm3d_plan_step2_opponent(
            self._h, _ptr_array([src[c].data_ptr() for c in range(3)]), src.stride(1), float(sigma),
            _ptr_array([basic[c].data_ptr() for c in range(3)]), basic.stride(1),
            _ptr_array([out[c].data_ptr() for c in range(3)]), out.stride(1), _stream_handle(stream)))

    def set_pilot(self, net, fallback_db: float = 0.5, check_block: int = 256,
                  check_every: int = 16, full_fraction: float = 0.5, under_db: float = 0.25) -> None:
        """Use a learned pilot (cpipe_amd.pilot.Pilot) in place of step 1; None restores it."""
        opts = PilotOpts(ctypes.sizeof(PilotOpts), fallback_db, check_block, check_every,
                         full_fraction, under_db)
        _check(lib().cpipe_bm3d_plan_set_pilot(self._h, net.handle if net is not None else None,
                                               ctypes.byref(opts)))
        self._pilot = net  # the plan references the network

    def set_block_matching(self, kind: str) -> None:
        """Mixed-precision plans: "mfma" (cross term on MFMA + FP32 re-rank) or "exact"."""
        _check(lib().cpipe_bm3d_plan_set_block_matching(self._h, 1 if kind == "mfma" else 0))

    def set_noise_profile(self, beta=None, lam=None) -> None:
        """Per-reference sigma from a noise profile (variance beta + lam * x per
        channel; scalars apply to every channel); None returns to the scalar sigma."""
        import numpy as np
        if beta is None:
            _check(lib().cpipe_bm3d_plan_set_noise_profile(self._h, None, None, 0))
            return
        b = np.atleast_1d(np.asarray(beta, np.float32)).ravel()
        l_ = np.atleast_1d(np.asarray(lam, np.float32)).ravel()
        if b.size == 1:
            b = np.repeat(b, self.channels)
        if l_.size == 1:
            l_ = np.repeat(l_, self.channels)
        if b.size != l_.size:
            raise ValueError("beta and lambda need the same number of entries")
        b, l_ = np.ascontiguousarray(b), np.ascontiguousarray(l_)
        _check(lib().cpipe_bm3d_plan_set_noise_profile(self._h, b.ctypes.data, l_.ctypes.data, int(b.size)))

    def pilot_report(self):
        nc, nf, full = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        md = ctypes.c_float()
        _check(lib().cpipe_bm3d_plan_pilot_report(self._h, ctypes.byref(nc), ctypes.byref(nf),
                                                  ctypes.byref(md), ctypes.byref(full)))
        return {"checked": nc.value, "fallback": nf.value, "max_delta_db": md.value,
                "full_fallback": bool(full.value)}

    def check(self) -> None:
        """Wait for the last run and raise if it hit an internal consistency check."""
        _check(lib().cpipe_bm3d_plan_check(self._h))

    def debug_flag(self, stream=None) -> None:
        """Tests only: mark the last run as failed (the next check() raises until
        the plan runs again)."""
        _check(lib().cpipe_bm3d_plan_debug_flag(self._h, _stream_handle(stream)))

    def set_timing(self, enable: bool) -> None:
        _check(lib().cpipe_bm3d_plan_set_timing(self._h, 1 if enable else 0))

    def timing(self):
        ms = (ctypes.c_float * len(STAGES))()
        runs = ctypes.c_int()
        _check(lib().cpipe_bm3d_plan_timing(self._h, ms, len(STAGES), ctypes.byref(runs)))
        return {k: float(v) for k, v in zip(STAGES, ms)}, runs.value

    def _check_planes(self, t) -> None:
        if tuple(t.shape) != (self.channels, self.height, self.width):
            raise ValueError(f"expected planes of shape {(self.channels, self.height, self.width)}, "
                             f"got {tuple(t.shape)}")
        if t.dtype.__repr__() != "torch.float32" or t.stride(2) != 1 or not t.is_cuda:
            raise ValueError("expected float32 CUDA planes with unit stride along W")
        if self.channels > 1 and t.stride(0) < t.stride(1) * self.height:
            raise ValueError("planes overlap")


class Batch(Plan):
    """Frames of uneven sizes in one plan: one launch per stage over all frames,
    optionally replayed as one captured hipGraph per run."""

    def __init__(self, sizes, channels: int = 3, precision="fp32", device: int = 0,
                 use_graph: bool = True):
        if isinstance(precision, str):
            precision = _PRECISIONS[precision]
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        self.channels, self.precision, self.device = int(channels), int(precision), int(device)
        n = len(self.sizes)
        ws = (ctypes.c_int32 * n)(*[w for w, _ in self.sizes])
        hs = (ctypes.c_int32 * n)(*[h for _, h in self.sizes])
        d = BatchDesc(ctypes.sizeof(BatchDesc), n, ws, hs, self.channels, self.precision,
                      self.device, 1 if use_graph else 0)
        h = ctypes.c_void_p()
        _check(lib().cpipe_bm3d_batch_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().cpipe_bm3d_batch_destroy(self._h)
            self._h = None

    def set_noise_profiles(self, betas=None, lams=None) -> None:
        """Per-frame noise profiles: (n_frames, channels) arrays (a row per frame,
        scalars broadcast); None returns every frame to its scalar sigma."""
        import numpy as np
        n, C = len(self.sizes), self.channels
        if betas is None:
            _check(lib().cpipe_bm3d_batch_set_noise_profiles(self._h, None, None, 0))
            return
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, np.float32), (n, C)))
        l_ = np.ascontiguousarray(np.broadcast_to(np.asarray(lams, np.float32), (n, C)))
        _check(lib().cpipe_bm3d_batch_set_noise_profiles(self._h, b.ctypes.data, l_.ctypes.data, C))

    def _args(self, srcs, dsts, sigmas):
        n, C = len(self.sizes), self.channels
        if len(srcs) != n or len(dsts) != n or len(sigmas) != n:
            raise ValueError(f"expected {n} frames")
        ins = (ctypes.c_void_p * (n * C))()
        outs = (ctypes.c_void_p * (n * C))()
        ip = (ctypes.c_size_t * n)()
        op = (ctypes.c_size_t * n)()
        for f, ((w, h), s, d) in enumerate(zip(self.sizes, srcs, dsts)):
            for t in (s, d):
                if tuple(t.shape) != (C, h, w) or t.stride(2) != 1 or not t.is_cuda:
                    raise ValueError(f"frame {f}: expected ({C}, {h}, {w}) float32 CUDA planes")
            for c in range(C):
                ins[f * C + c] = s[c].data_ptr()
                outs[f * C + c] = d[c].data_ptr()
            ip[f] = s.stride(1)
            op[f] = d.stride(1)
        sg = (ctypes.c_float * n)(*[float(v) for v in sigmas])
        return ins, ip, outs, op, sg

    def run(self, srcs, dsts, sigmas, stream=None) -> None:
        ins, ip, outs, op, sg = self._args(srcs, dsts, sigmas)
        _check(lib().cpipe_bm3d_batch_run(self._h, ins, ip, outs, op, sg, _stream_handle(stream)))

    def tune(self, srcs, dsts, sigmas, iters: int = 2, stream=None):
        ins, ip, outs, op, sg = self._args(srcs, dsts, sigmas)
        rep = TuneReport()
        _check(lib().cpipe_bm3d_batch_tune(self._h, ins, ip, outs, op, sg, int(iters),
                                           _stream_handle(stream), ctypes.byref(rep)))
        cands = [{"tile_y": rep.tile_y[i], "order": rep.order[i], "band_rows": rep.band_rows[i],
                  "bm": "mfma" if rep.bm_mfma[i] else "exact", "ms": round(float(rep.ms[i]), 3)}
                 for i in range(rep.n)]
        return {"candidates": cands, "best": rep.best}

    def config(self):
        t, o, b, m = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().cpipe_bm3d_batch_config(self._h, ctypes.byref(t), ctypes.byref(o), ctypes.byref(b),
                                             ctypes.byref(m)))
        return {"tile_y": t.value, "order": o.value, "band_rows": b.value,
                "bm": "mfma" if m.value else "exact"}


def compat_rgba16f(src, dst, sigma: float, stage: int = 1, stream=None) -> None:
    """Reference-compatible filter on (H, W, 4) RGBA16F CUDA tensors."""
    h, w, c = src.shape
    if c != 4 or tuple(dst.shape) != (h, w, 4):
        raise ValueError("expected (H, W, 4) RGBA16F tensors")
    _check(lib().cpipe_bm3d_compat_rgba16f(src.data_ptr(), src.stride(0) // 4, dst.data_ptr(),
                                           dst.stride(0) // 4, w, h, float(sigma), int(stage),
                                           _stream_handle(stream)))


def _wavelet_check(rc: int) -> None:
    if rc != 0:
        raise EngineError(rc, lib().cpipe_wavelet_last_error().decode(errors="replace"))


def wavelet_rgba16f(src, dst, chroma_strength: float = 1.0, stream=None) -> None:
    """Wavelet BayesShrink chroma denoise (com.cpipe.denoise.wavelet_bayes_shrink) on
    (H, W, 4) RGBA16F CUDA tensors (float16, or int16/uint16 holding binary16 bits)."""
    h, w, c = src.shape
    if c != 4 or tuple(dst.shape) != (h, w, 4):
        raise ValueError("expected (H, W, 4) RGBA16F tensors")
    for t in (src, dst):
        if t.element_size() != 2 or t.stride(2) != 1 or t.stride(1) != 4 or not t.is_cuda:
            raise ValueError("expected 2-byte CUDA tensors with interleaved pixels")
    _wavelet_check(lib().cpipe_wavelet_rgba16f(src.data_ptr(), src.stride(0) // 4, dst.data_ptr(),
                                               dst.stride(0) // 4, w, h, float(chroma_strength),
                                               _stream_handle(stream)))


def wavelet_host_rgba16f(src, chroma_strength: float = 1.0):
    """The same on a host (H, W, 4) uint16 numpy array (binary16 bits); returns the output."""
    import numpy as np
    src = np.ascontiguousarray(src, dtype=np.uint16)
    h, w, c = src.shape
    if c != 4:
        raise ValueError("expected (H, W, 4) RGBA16F bits")
    out = np.empty_like(src)
    _wavelet_check(lib().cpipe_wavelet_host_rgba16f(src.ctypes.data, w, out.ctypes.data, w, w, h,
                                                    float(chroma_strength)))
    return out


# pixel formats, numbered as the reference's PixelFormat (include/cpipe_convert.h)
CONVERT_FORMATS = {"R16_UINT": 1, "R8G8B8A8_UNORM": 3, "R16G16B16A16_SFLOAT": 6, "R32_SFLOAT": 10,
                   "R16G16B16A16_UNORM": 16}


def _convert_check(rc: int) -> None:
    if rc != 0:
        raise EngineError(rc, lib().cpipe_convert_last_error().decode(errors="replace"))


def convert(src, src_format: str, dst, dst_format: str, stream=None) -> None:
    """Precision conversion (com.cpipe.precision_convert) between CUDA tensors laid out
    as (H, W) or (H, W, 4) with unit stride along the pixel's channels."""
    h, w = src.shape[:2]
    if tuple(dst.shape[:2]) != (h, w):
        raise ValueError("source and destination sizes differ")
    for t in (src, dst):
        if not t.is_cuda or t.stride(-1) != 1 or (t.dim() == 3 and t.stride(1) != t.shape[2]):
            raise ValueError("expected CUDA tensors with interleaved pixels")
    _convert_check(lib().cpipe_convert(CONVERT_FORMATS[src_format], src.data_ptr(),
                                       src.stride(0) * src.element_size(), CONVERT_FORMATS[dst_format],
                                       dst.data_ptr(), dst.stride(0) * dst.element_size(), w, h,
                                       _stream_handle(stream)))


def denoise(planes, sigma: float, precision="fp32", stream=None):
    """BM3D both steps on (C, H, W) float32 CUDA planes; returns new planes."""
    import torch
    c, h, w = planes.shape
    plan = Plan(w, h, c, precision, planes.device.index or 0)
    out = torch.empty_like(planes)
    plan.run(planes.contiguous(), out, sigma, stream)
    return out
