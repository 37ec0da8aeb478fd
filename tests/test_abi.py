"""The C ABI library and the node's plugin contract, checked without a GPU.

No compute call reaches a kernel here: without a HIP device the engine must
refuse with defined statuses (CPIPE_UNSUPPORTED), never fall back to the CPU.
"""
import ctypes
import json
import os
import re
import subprocess

import pytest

from cpipe_amd import engine, host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared_functions(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(cpipe_\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if not n.endswith("_t")))


def test_library_exports_every_declared_symbol():
    lib = engine.lib()
    headers = sorted(f for f in os.listdir(INCLUDE) if f.endswith(".h"))
    declared = [n for h in headers for n in _declared_functions(h)]
    assert "cpipe_dng_noise_profile" in declared
    assert "cpipe_convert" in declared and "cpipe_convert_node_descriptor" in declared
    assert "cpipe_bm3d_plan_run" in declared and "cpipe_bm3d_node_descriptor" in declared
    assert "cpipe_wavelet_rgba16f" in declared and "cpipe_wavelet_node_descriptor" in declared
    for name in declared:
        assert hasattr(lib, name), f"missing export {name}"
    for name in engine.EXPORTS:
        assert hasattr(lib, name), f"missing export {name}"


def test_abi_header_layout_static_asserts(tmp_path):
    src = tmp_path / "abi.cpp"
    src.write_text('#include "cpipe_node_abi.h"\n#include "cpipe_bm3d.h"\n#include "cpipe_wavelet.h"\n'
                   '#include "cpipe_convert.h"\n#include "cpipe_dng.h"\nint main(){return 0;}\n')
    subprocess.run(["g++", "-std=c++17", f"-I{INCLUDE}", "-fsyntax-only", str(src)], check=True)
    csrc = tmp_path / "abi.c"
    csrc.write_text('#include "cpipe_node_abi.h"\n#include "cpipe_bm3d.h"\n#include "cpipe_wavelet.h"\n'
                    '#include "cpipe_convert.h"\n#include "cpipe_dng.h"\nint main(void){return 0;}\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{INCLUDE}", "-fsyntax-only", str(csrc)],
                   check=True)


def test_registry_section_holds_the_nodes():
    assert host.mh().mh_registry_count(engine.ENGINE_SO.encode()) == 3
    d = host.find_node()
    assert d.contents.abi_major == 0 and d.contents.abi_minor == 4
    assert d.contents.node_id == b"com.cpipe.denoise.bm3d"
    direct = engine.lib().cpipe_bm3d_node_descriptor()
    assert direct == ctypes.addressof(d.contents)
    wd = host.find_node(WAVELET_ID)
    assert wd.contents.abi_major == 0 and wd.contents.abi_minor == 4
    assert engine.lib().cpipe_wavelet_node_descriptor() == ctypes.addressof(wd.contents)
    cd = host.find_node(CONVERT_ID)
    assert engine.lib().cpipe_convert_node_descriptor() == ctypes.addressof(cd.contents)


CONVERT_ID = "com.cpipe.precision_convert"


def test_convert_manifest_and_status_contract():
    """Manifest ports and applied step (precision_convert.json), and the node's statuses
    (precision_convert.cpp:49-72): arity -> BAD_INDEX, no target_format -> its param
    status, target/output mismatch and unsupported pairs -> BAD_PRECISION; without a
    device CPIPE_UNSUPPORTED, never a CPU fallback."""
    d = host.find_node(CONVERT_ID)
    m = json.loads(d.contents.manifest_json.decode())
    assert m["id"] == CONVERT_ID and m["params"] == []
    assert m["ports"][1]["metadata"]["sets_steps_applied"] == ["precision_convert"]
    assert m["compute"]["device"] == "GPU" and m["compute"]["engine"] == "HIP"
    a = host.Buffer(host.FORMAT_RGBA16F, 2, 1)
    b8 = host.Buffer(host.FORMAT_RGBA8, 2, 1)
    c = host.Buffer(host.FORMAT_RGBA16F, 2, 1)
    assert host.run_node(d, [a, c], [b8], {"target_format": "R8G8B8A8_UNORM"}) == host.CPIPE_BAD_INDEX
    assert host.run_node(d, [a], [b8], {}) == 6  # CPIPE_NEED_PARAM
    assert host.run_node(d, [a], [b8], {"target_format": "R16G16B16A16_UNORM"}) == host.CPIPE_BAD_PRECISION
    f = host.Buffer(host.FORMAT_R32F, 2, 1)
    assert host.run_node(d, [f], [b8], {"target_format": "R8G8B8A8_UNORM"}) == host.CPIPE_BAD_PRECISION
    L = engine.lib()
    assert L.cpipe_convert_supported(6, 3) == 1 and L.cpipe_convert_supported(3, 6) == 0
    assert L.cpipe_convert(6, None, 16, 3, None, 8, 2, 1, None) == host.CPIPE_BAD_INDEX
    assert L.cpipe_convert(3, None, 16, 6, None, 8, 2, 1, None) == host.CPIPE_BAD_PRECISION
    import torch
    if torch.cuda.is_available():
        return
    assert host.run_node(d, [a], [b8], {"target_format": "R8G8B8A8_UNORM"}) == host.CPIPE_UNSUPPORTED


WAVELET_ID = "com.cpipe.denoise.wavelet_bayes_shrink"


def test_wavelet_manifest_contract():
    """Assertions of test_node_denoise_wavelet_bayes_shrink.cpp:46-52 (ports, applied
    step) and the manifest's params (wavelet_bayes_shrink.json:37-46), with the
    engine changed to HIP."""
    m = json.loads(host.find_node(WAVELET_ID).contents.manifest_json.decode())
    assert m["id"] == WAVELET_ID
    assert m["ports"][1]["metadata"]["sets_steps_applied"] == ["denoise.wavelet_bayes_shrink"]
    assert m["ports"][0]["metadata"]["requires_steps_applied"] == ["color_matrix"]
    assert m["ports"][0]["caps"]["precision"] == ["f16"] and m["ports"][0]["caps"]["channels"] == ["rgba"]
    p = m["params"][0]
    assert p["name"] == "chroma_strength" and p["range"] == {"min": 0, "max": 2} and p["default"] == 1
    assert m["compute"]["device"] == "GPU" and m["compute"]["engine"] == "HIP"


def test_wavelet_status_contract():
    """Arity -> BAD_INDEX (wavelet_bayes_shrink.cpp:31-35), format gate -> BAD_PRECISION,
    and without a device CPIPE_UNSUPPORTED, never a CPU fallback."""
    d = host.find_node(WAVELET_ID)
    a, b = _rgba_buffers()
    c = host.Buffer(host.FORMAT_RGBA16F, 4, 4)
    assert host.run_node(d, [a, c], [b]) == host.CPIPE_BAD_INDEX
    f1 = host.Buffer(host.FORMAT_R32F, 4, 4)
    assert host.run_node(d, [f1], [b]) == host.CPIPE_BAD_PRECISION
    big = host.Buffer(host.FORMAT_RGBA16F, 8, 4)
    assert host.run_node(d, [a], [big]) == host.CPIPE_BAD_INDEX
    assert host.mh().mh_call_action(d, b"describe") == host.CPIPE_REPLY_DEFAULT
    L = engine.lib()
    assert L.cpipe_wavelet_rgba16f(None, 4, None, 4, 4, 4, ctypes.c_float(1.0), None) == host.CPIPE_BAD_INDEX
    assert b"null" in L.cpipe_wavelet_last_error()
    import torch
    if torch.cuda.is_available():
        return
    assert host.run_node(d, [a], [b]) == host.CPIPE_UNSUPPORTED
    assert "no HIP device" in host.host_log()
    import numpy as np
    with pytest.raises(engine.EngineError) as e:
        engine.wavelet_host_rgba16f(np.zeros((4, 4, 4), np.uint16))
    assert e.value.status == host.CPIPE_UNSUPPORTED


def test_manifest_contract():
    """Assertions of test_node_denoise_bm3d.cpp:137-144, with the engine changed to HIP."""
    d = host.find_node()
    m = json.loads(d.contents.manifest_json.decode())
    assert m["id"] == "com.cpipe.denoise.bm3d"
    assert m["params"][0]["name"] == "sigma"
    assert m["params"][1]["name"] == "sigma_override"
    assert m["params"][0]["range"] == {"min": 0, "max": 0.2}
    assert m["ports"][1]["metadata"]["sets_steps_applied"] == ["denoise.bm3d"]
    assert m["ports"][0]["metadata"]["requires_steps_applied"] == ["color_matrix"]
    assert m["compute"]["device"] == "GPU" and m["compute"]["engine"] == "HIP"
    assert "f16" in m["ports"][0]["caps"]["precision"] and "f32" in m["ports"][0]["caps"]["precision"]
    modes = [p for p in m["params"] if p["name"] == "mode"][0]
    assert modes["enum_values"] == ["compat", "bm3d"] and modes["default"] == "compat"
    pil = [p for p in m["params"] if p["name"] == "pilot"][0]
    assert pil["enum_values"] == ["off", "learned"] and pil["default"] == "off"


def test_embedded_pilot_weights_match_the_package():
    """The node's learned-pilot weights are the packaged blob, byte for byte."""
    import ctypes
    from cpipe_amd import engine, pilot
    n = ctypes.c_size_t(0)
    ptr = engine.lib().cpipe_bm3d_pilot_embedded(ctypes.byref(n))
    with open(pilot.DEFAULT_WEIGHTS, "rb") as f:
        want = f.read()
    assert n.value == len(want)
    assert ctypes.string_at(ptr, n.value) == want


def test_actions_without_work():
    d = host.find_node()
    assert host.mh().mh_call_action(d, b"describe") == host.CPIPE_REPLY_DEFAULT
    assert host.mh().mh_call_action(d, b"prepare") == host.CPIPE_REPLY_DEFAULT
    assert host.mh().mh_call_action(d, b"no-such-action") == host.CPIPE_REPLY_DEFAULT


def _rgba_buffers(w=4, h=4):
    a = host.Buffer(host.FORMAT_RGBA16F, w, h)
    b = host.Buffer(host.FORMAT_RGBA16F, w, h)
    return a, b


def test_status_contract_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("this checks the no-device status")
    d = host.find_node()
    a, b = _rgba_buffers()
    host.mh().mh_clear_log()
    assert host.run_node(d, [a], [b]) == host.CPIPE_UNSUPPORTED
    assert "no HIP device" in host.host_log()
    assert engine.device_count() == 0
    with pytest.raises(engine.EngineError) as e:
        engine.Plan(64, 64, 3)
    assert e.value.status == host.CPIPE_UNSUPPORTED


def test_format_gate_and_arity():
    d = host.find_node()
    # compat mode takes RGBA16F only (bm3d.cpp:57-68 -> CPIPE_BAD_PRECISION)
    f1 = host.Buffer(host.FORMAT_R32F, 8, 8)
    f2 = host.Buffer(host.FORMAT_R32F, 8, 8)
    assert host.run_node(d, [f1], [f2]) == host.CPIPE_BAD_PRECISION
    a, _ = _rgba_buffers()
    # mixed formats
    assert host.run_node(d, [a], [f2], {"mode": "bm3d"}) == host.CPIPE_BAD_PRECISION
    # two inputs, two outputs is not an accepted arity
    g1 = host.Buffer(host.FORMAT_R32F, 8, 8)
    g2 = host.Buffer(host.FORMAT_R32F, 8, 8)
    assert host.run_node(d, [f1, g1], [f2, g2], {"mode": "bm3d"}) == host.CPIPE_BAD_INDEX
    # size mismatch
    big = host.Buffer(host.FORMAT_RGBA16F, 8, 4)
    assert host.run_node(d, [a], [big]) == host.CPIPE_BAD_INDEX
    # unknown enum values
    b = host.Buffer(host.FORMAT_RGBA16F, 4, 4)
    assert host.run_node(d, [a], [b], {"mode": "fast"}) == host.CPIPE_BAD_INDEX
    assert host.run_node(d, [a], [b], {"precision": "int4"}) == host.CPIPE_BAD_INDEX


def test_engine_argument_checks_without_device():
    L = engine.lib()
    desc = engine.PlanDesc(ctypes.sizeof(engine.PlanDesc), 64, 64, 2, 0, 0)
    h = ctypes.c_void_p()
    assert L.cpipe_bm3d_plan_create(ctypes.byref(desc), ctypes.byref(h)) == host.CPIPE_BAD_INDEX
    desc = engine.PlanDesc(4, 64, 64, 3, 0, 0)
    assert L.cpipe_bm3d_plan_create(ctypes.byref(desc), ctypes.byref(h)) == host.CPIPE_BAD_INDEX
    assert L.cpipe_bm3d_plan_create(None, ctypes.byref(h)) == host.CPIPE_BAD_INDEX
    assert L.cpipe_bm3d_compat_rgba16f(None, 4, None, 4, 4, 4, ctypes.c_float(0.03), 1, None) == host.CPIPE_BAD_INDEX
    assert b"null" in L.cpipe_bm3d_last_error()


def test_device_buffer_path_without_device():
    """The device-resident path: the engine's device entry points refuse with
    defined statuses, and the minimal host refuses a device buffer, when no HIP
    device is visible; host buffers keep the lock_cpu path whether or not the
    host offers the "hip_buffer" suite."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("this checks the no-device status")
    L = engine.lib()
    F = ctypes.c_float
    assert L.cpipe_bm3d_device_rgba16f(None, 4, None, 4, 4, 4, F(0.03), 1, 0, 0, 0, None) == host.CPIPE_BAD_INDEX
    assert L.cpipe_bm3d_device_planes(None, 4, None, 4, 4, 4, 3, F(0.03), 0, 0, 0, None) == host.CPIPE_BAD_INDEX
    dummy = (ctypes.c_uint16 * 64)()
    assert L.cpipe_bm3d_device_rgba16f(dummy, 4, dummy, 4, 4, 4, F(0.03), 1, 0, 0, 0, None) == host.CPIPE_UNSUPPORTED
    with pytest.raises(ValueError):
        host.Buffer(host.FORMAT_RGBA16F, 4, 4, device=0)
    d = host.find_node()
    for offer in (True, False):
        host.offer_hip_buffer(offer)
        a, b = _rgba_buffers()
        assert host.run_node(d, [a], [b]) == host.CPIPE_UNSUPPORTED
    host.offer_hip_buffer(True)


def test_host_buffer_lock_round_trip():
    """write/read go through lock_cpu/unlock_cpu; a second lock while one is held fails."""
    import numpy as np
    b = host.Buffer(host.FORMAT_R32F, 5, 3, row_bytes=32)
    v = np.arange(15, dtype=np.float32).reshape(3, 5)
    b.write(v)
    assert np.array_equal(b.read(), v)
    assert not b.on_device
    p = host.mh().mh_buffer_lock(b.h, host.CPU_READ)
    assert p
    assert not host.mh().mh_buffer_lock(b.h, host.CPU_READ)
    assert host.mh().mh_buffer_unlock(b.h) == host.CPIPE_OK


def test_noise_profile_argument_checks_without_device():
    """The noise-profile entry points check their arguments before any device work."""
    L = engine.lib()
    F3 = ctypes.c_float * 3
    b, lam = F3(1e-4, 1e-4, 1e-4), F3(0.01, 0.01, 0.01)
    assert L.cpipe_bm3d_host_noise_profile(b, lam, 4) == host.CPIPE_BAD_INDEX
    assert L.cpipe_bm3d_host_noise_profile(None, None, 2) == host.CPIPE_BAD_INDEX
    assert L.cpipe_bm3d_host_noise_profile(F3(float("nan"), 0, 0), lam, 1) == host.CPIPE_BAD_INDEX
    assert L.cpipe_bm3d_host_noise_profile(b, lam, 3) == host.CPIPE_OK
    assert L.cpipe_bm3d_host_noise_profile(None, None, 0) == host.CPIPE_OK
    assert L.cpipe_bm3d_plan_set_noise_profile(None, b, lam, 3) == host.CPIPE_BAD_INDEX
    assert L.cpipe_bm3d_batch_set_noise_profiles(None, b, lam, 3) == host.CPIPE_BAD_INDEX
    # the profile flag needs mode bm3d (checked before the device)
    px = (ctypes.c_uint16 * 64)()
    assert L.cpipe_bm3d_host_rgba16f(px, 4, px, 4, 4, 4, ctypes.c_float(0.03), 0, 0, 0, 2) == host.CPIPE_BAD_INDEX
