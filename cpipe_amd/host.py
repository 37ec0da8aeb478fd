"""Drive the BM3D node through the cpipe plugin ABI with the minimal host.

Mirrors the reference unit-test driver process_bm3d
(tests/unit/test_node_denoise_bm3d.cpp:89-125): registry lookup by node id,
create -> process -> freeze metadata -> destroy, host buffers in and out, or
device-resident buffers (HIP device memory, offered to the nodes through the
host's "hip_buffer" suite) that chained nodes read and write with no host copy.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import engine

MINIHOST_SO = os.path.join(engine.PKG, "lib", "libcpipe_minihost.so")
NODE_ID = "com.cpipe.denoise.bm3d"

FORMAT_R16U = 1
FORMAT_RGBA8 = 3
FORMAT_RGBA16F = 6
FORMAT_R32F = 10
FORMAT_RGBA16U = 16

# element type and channels per pixel of each format
_LAYOUT = {FORMAT_R16U: (np.uint16, 1), FORMAT_RGBA8: (np.uint8, 4), FORMAT_RGBA16F: (np.uint16, 4),
           FORMAT_R32F: (np.float32, 1), FORMAT_RGBA16U: (np.uint16, 4)}

CPU_READ, CPU_WRITE = 0, 1  # cpipe_cpu_access_t

CPIPE_OK = 0
CPIPE_REPLY_DEFAULT = 2
CPIPE_BAD_PRECISION = 4
CPIPE_BAD_INDEX = 5
CPIPE_INTERNAL_ERROR = 7
CPIPE_UNSUPPORTED = 8


class PluginDesc(ctypes.Structure):
    _fields_ = [("abi_major", ctypes.c_uint32), ("abi_minor", ctypes.c_uint32),
                ("node_id", ctypes.c_char_p), ("node_version", ctypes.c_char_p),
                ("manifest_json", ctypes.c_char_p), ("main_entry", ctypes.c_void_p)]


_mh = None


def mh() -> ctypes.CDLL:
    global _mh
    if _mh is None:
        if not os.path.exists(MINIHOST_SO):
            raise RuntimeError(f"minimal host missing: {MINIHOST_SO} (run cpipe_amd.build.build())")
        L = ctypes.CDLL(MINIHOST_SO)
        P, I, S = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
        L.mh_log.restype = ctypes.c_char_p
        L.mh_find_node.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.mh_find_node.restype = ctypes.POINTER(PluginDesc)
        L.mh_registry_count.argtypes = [ctypes.c_char_p]
        L.mh_buffer_new.argtypes = [I, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64]
        L.mh_buffer_new.restype = P
        L.mh_buffer_new_device.argtypes = [I, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, I]
        L.mh_buffer_new_device.restype = P
        L.mh_buffer_is_device.argtypes = [P]
        L.mh_buffer_lock.argtypes = [P, I]
        L.mh_buffer_lock.restype = P
        L.mh_buffer_unlock.argtypes = [P]
        L.mh_offer_hip_buffer.argtypes = [I]
        L.mh_buffer_free.argtypes = [P]
        L.mh_buffer_data.argtypes = [P]
        L.mh_buffer_data.restype = P
        L.mh_buffer_row_bytes.argtypes = [P]
        L.mh_buffer_row_bytes.restype = ctypes.c_uint64
        L.mh_buffer_set_noise_profile.argtypes = [P, P, P, S]
        L.mh_buffer_set_cs_role.argtypes = [P, ctypes.c_char_p]
        L.mh_buffer_add_step.argtypes = [P, ctypes.c_char_p]
        L.mh_buffer_cs_role.argtypes = [P]
        L.mh_buffer_cs_role.restype = ctypes.c_char_p
        L.mh_buffer_step_count.argtypes = [P]
        L.mh_buffer_step.argtypes = [P, I]
        L.mh_buffer_step.restype = ctypes.c_char_p
        L.mh_params_new.restype = P
        L.mh_params_free.argtypes = [P]
        L.mh_params_set_double.argtypes = [P, ctypes.c_char_p, ctypes.c_double]
        L.mh_params_set_enum.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p]
        L.mh_run_node.argtypes = [ctypes.POINTER(PluginDesc), P, P, S, P, S]
        L.mh_call_action.argtypes = [ctypes.POINTER(PluginDesc), ctypes.c_char_p]
        _mh = L
    return _mh


def find_node(node_id: str = NODE_ID, library: str = engine.ENGINE_SO):
    d = mh().mh_find_node(library.encode(), node_id.encode())
    if not d:
        raise LookupError(f"{node_id} not registered in {library}: {mh().mh_log().decode()}")
    return d


def offer_hip_buffer(on: bool) -> None:
    """Whether the host offers the "hip_buffer" suite (default on); off, the nodes
    see a host without it and reach every buffer through lock_cpu."""
    mh().mh_offer_hip_buffer(1 if on else 0)


class Buffer:
    """An image buffer with cpipe metadata: host memory (CpuBuffer analogue) or,
    with `device` set, HIP device memory on that device ordinal (zero-filled;
    write/read go through lock_cpu, which stages a host copy)."""

    def __init__(self, fmt: int, width: int, height: int, row_bytes: int = 0, device: int | None = None):
        self.fmt, self.width, self.height = fmt, width, height
        if device is None:
            self.h = mh().mh_buffer_new(fmt, width, height, row_bytes)
        else:
            self.h = mh().mh_buffer_new_device(fmt, width, height, row_bytes, device)
        if not self.h:
            raise ValueError(f"bad buffer layout or device allocation failed: {host_log()}")
        self.row_bytes = int(mh().mh_buffer_row_bytes(self.h))

    @property
    def on_device(self) -> bool:
        return bool(mh().mh_buffer_is_device(self.h))

    def __del__(self):  # pragma: no cover
        try:
            mh().mh_buffer_free(self.h)
        except Exception:
            pass

    def _view(self, ptr) -> np.ndarray:
        dtype, per_px = _LAYOUT[self.fmt]
        n = self.row_bytes * self.height // np.dtype(dtype).itemsize
        ctype = np.ctypeslib.as_ctypes_type(dtype)
        base = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(n,))
        rows = base.reshape(self.height, -1)
        return rows[:, : self.width * per_px]

    def _locked(self, access: int):
        ptr = mh().mh_buffer_lock(self.h, access)
        if not ptr:
            raise RuntimeError(f"lock_cpu failed: {host_log()}")
        return ptr

    def write(self, arr: np.ndarray) -> None:
        v = self._view(self._locked(CPU_WRITE))
        try:
            v[...] = np.asarray(arr).reshape(v.shape)
        finally:
            if mh().mh_buffer_unlock(self.h) != CPIPE_OK:
                raise RuntimeError(f"unlock_cpu failed: {host_log()}")

    def read(self) -> np.ndarray:
        try:
            v = self._view(self._locked(CPU_READ)).copy()
        finally:
            mh().mh_buffer_unlock(self.h)
        return v.reshape(self.height, self.width, 4) if _LAYOUT[self.fmt][1] == 4 else v

    def set_noise_profile(self, pairs) -> None:
        beta = np.array([p[0] for p in pairs], np.float32)
        lam = np.array([p[1] for p in pairs], np.float32)
        mh().mh_buffer_set_noise_profile(self.h, beta.ctypes.data, lam.ctypes.data, len(pairs))

    def set_cs_role(self, role: str) -> None:
        mh().mh_buffer_set_cs_role(self.h, role.encode())

    def add_step(self, step: str) -> None:
        mh().mh_buffer_add_step(self.h, step.encode())

    @property
    def cs_role(self) -> str:
        return mh().mh_buffer_cs_role(self.h).decode()

    @property
    def applied_steps(self):
        return [mh().mh_buffer_step(self.h, i).decode() for i in range(mh().mh_buffer_step_count(self.h))]


def run_node(desc, inputs, outputs, params: dict | None = None) -> int:
    """create -> process -> destroy; returns the process status (create/destroy errors + 1000)."""
    p = mh().mh_params_new()
    try:
        for k, v in (params or {}).items():
            if isinstance(v, str):
                mh().mh_params_set_enum(p, k.encode(), v.encode())
            else:
                mh().mh_params_set_double(p, k.encode(), float(v))
        ins = (ctypes.c_void_p * len(inputs))(*[b.h for b in inputs])
        outs = (ctypes.c_void_p * len(outputs))(*[b.h for b in outputs])
        return int(mh().mh_run_node(desc, p, ins, len(inputs), outs, len(outputs)))
    finally:
        mh().mh_params_free(p)


def host_log() -> str:
    return mh().mh_log().decode()
