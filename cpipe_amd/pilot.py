"""Python binding of the learned pilot (include/cpipe_bm3d.h, cpipe_bm3d_pilot_*).

The network, its training step and its inference all run in the native
library; this module only moves tensors (torch, for device memory) and
host-side arrays in and out.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import engine

WEIGHTS_DIR = os.path.join(engine.PKG, "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "pilot_v1.bin")

_bound = False


def _lib():
    global _bound
    L = engine.lib()
    if not _bound:
        P, I, F, S = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_size_t
        L.cpipe_bm3d_pilot_last_error.restype = ctypes.c_char_p
        L.cpipe_bm3d_pilot_param_count.restype = S
        L.cpipe_bm3d_pilot_create.argtypes = [I, ctypes.POINTER(P)]
        L.cpipe_bm3d_pilot_destroy.argtypes = [P]
        L.cpipe_bm3d_pilot_destroy.restype = None
        L.cpipe_bm3d_pilot_init.argtypes = [P, ctypes.c_uint64]
        L.cpipe_bm3d_pilot_blob_size.argtypes = [I]
        L.cpipe_bm3d_pilot_blob_size.restype = S
        L.cpipe_bm3d_pilot_save.argtypes = [P, P, S, I]
        L.cpipe_bm3d_pilot_load.argtypes = [P, P, S]
        L.cpipe_bm3d_pilot_forward.argtypes = [P, P, S, S, I, I, I, F, P, P]
        L.cpipe_bm3d_pilot_gradients.argtypes = [P, P, P, I, I, I, P, P, ctypes.POINTER(F)]
        L.cpipe_bm3d_pilot_train_step.argtypes = [P, P, P, I, I, I, P, F, ctypes.POINTER(F)]
        L.cpipe_bm3d_pilot_set_loss.argtypes = [P, I, F]
        _bound = True
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise engine.EngineError(rc, _lib().cpipe_bm3d_pilot_last_error().decode(errors="replace"))


def param_count() -> int:
    return int(_lib().cpipe_bm3d_pilot_param_count())


def _planes3(t):
    """(N, 3, H, W) contiguous CUDA float32 -> plane pointers of image 0."""
    if t.dim() != 4 or t.shape[1] != 3 or not t.is_contiguous():
        raise ValueError("expected a contiguous (N, 3, H, W) tensor")
    return engine._ptr_array([t[0, c].data_ptr() for c in range(3)])


class Pilot:
    def __init__(self, device: int = 0, seed: int | None = 0, weights: str | bytes | None = None):
        h = ctypes.c_void_p()
        _check(_lib().cpipe_bm3d_pilot_create(int(device), ctypes.byref(h)))
        self._h = h
        self.device = device
        if weights is not None:
            self.load(weights)
        elif seed is not None:
            _check(_lib().cpipe_bm3d_pilot_init(self._h, int(seed)))

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib().cpipe_bm3d_pilot_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def save(self, with_optimizer: bool = False) -> bytes:
        n = _lib().cpipe_bm3d_pilot_blob_size(1 if with_optimizer else 0)
        buf = ctypes.create_string_buffer(n)
        _check(_lib().cpipe_bm3d_pilot_save(self._h, buf, n, 1 if with_optimizer else 0))
        return buf.raw

    def load(self, blob) -> None:
        if isinstance(blob, str):
            with open(blob, "rb") as f:
                blob = f.read()
        buf = ctypes.create_string_buffer(blob, len(blob))
        _check(_lib().cpipe_bm3d_pilot_load(self._h, buf, len(blob)))

    def params(self) -> np.ndarray:
        """FP32 master weights in the flat layout (per layer W[tap][cin][cout], then bias)."""
        blob = self.save(False)
        return np.frombuffer(blob, dtype=np.float32, offset=32, count=param_count()).copy()

    def forward(self, planes, sigma: float, out=None, stream=None):
        """planes: (N, 3, H, W) float32 CUDA opponent planes; returns the pilot planes."""
        import torch
        n, _, h, w = planes.shape
        if out is None:
            out = torch.empty_like(planes)
        _check(_lib().cpipe_bm3d_pilot_forward(self._h, _planes3(planes), w, 3 * h * w, w, h, n,
                                               float(sigma), _planes3(out),
                                               engine._stream_handle(stream)))
        return out

    def gradients(self, planes, target, sigma):
        n, _, h, w = planes.shape
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma, np.float32), (n,)))
        g = np.zeros(param_count(), np.float32)
        loss = ctypes.c_float()
        _check(_lib().cpipe_bm3d_pilot_gradients(self._h, _planes3(planes), _planes3(target), w, h, n,
                                                 s.ctypes.data, g.ctypes.data, ctypes.byref(loss)))
        return loss.value, g

    def set_loss(self, mode: str = "mse", ref_sigma: float = 0.05) -> None:
        """'mse' (default) or 'sigma': each image's squared error scaled by
        (ref_sigma / sigma_i)^2 (cpipe_bm3d_pilot_set_loss)."""
        m = {"mse": 0, "sigma": 1}[mode]
        _check(_lib().cpipe_bm3d_pilot_set_loss(self._h, m, float(ref_sigma)))

    def train_step(self, planes, target, sigma, lr: float) -> float:
        n, _, h, w = planes.shape
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma, np.float32), (n,)))
        loss = ctypes.c_float()
        _check(_lib().cpipe_bm3d_pilot_train_step(self._h, _planes3(planes), _planes3(target), w, h,
                                                  n, s.ctypes.data, float(lr), ctypes.byref(loss)))
        return loss.value
