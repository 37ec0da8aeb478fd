"""In-tree build of the native libraries (hipcc for gfx950, g++ for host-only code).

Outputs (git-ignored, shipped with the tree):
  cpipe_amd/lib/libcpipe_bm3d.so      HIP engine + C ABI (include/cpipe_bm3d.h, include/cpipe_wavelet.h,
                                      include/cpipe_convert.h, include/cpipe_dng.h)
                                      + the nodes' descriptors in the `cpipe_registry` section
  cpipe_amd/lib/libcpipe_minihost.so  minimal cpipe host (suites, registry walk, one-node runner,
                                      host and HIP device-resident buffers)
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
ENGINE_SO = os.path.join(LIBDIR, "libcpipe_bm3d.so")
MINIHOST_SO = os.path.join(LIBDIR, "libcpipe_minihost.so")

ARCH = "gfx950"
# Exact FP32 order: no contraction, IEEE division/sqrt, denormals kept.
# Packed FP32 instructions are disabled for the device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{INCLUDE}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
             f"-I{INCLUDE}"]
# Host C++ that calls the HIP runtime API (device selection, streams, copies).
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIP_HOST_FLAGS = ["-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]
HIP_HOST_LIBS = [f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the BM3D engine needs ROCm's hipcc to build")


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {cmd[0]} ... {cmd[-1]}")
    return r


MANIFESTS = {"CPIPE_BM3D_MANIFEST_JSON": "bm3d.json",
             "CPIPE_WAVELET_MANIFEST_JSON": "wavelet_bayes_shrink.json",
             "CPIPE_CONVERT_MANIFEST_JSON": "precision_convert.json"}


def _embed_manifest() -> str:
    """Turn the nodes' csrc/*.json manifests into C++ string constants (like the
    reference's EmbedJson.cmake)."""
    out = os.path.join(BUILD, "manifest_gen.cpp")
    lines = ["// generated from " + ", ".join(f"cpipe_amd/csrc/{f}" for f in MANIFESTS.values())]
    for sym, name in MANIFESTS.items():
        with open(os.path.join(CSRC, name), encoding="utf-8") as f:
            text = f.read()
        json.loads(text)  # fail early on a malformed manifest
        body = ", ".join(str(b) for b in text.encode("utf-8")) + ", 0"
        lines.append(f'extern "C" const char {sym}[];')
        lines.append(f'extern "C" const char {sym}[] = {{{body}}};')
    with open(out, "w", encoding="utf-8") as f:
        f.write("\n".join(lines) + "\n")
    return out


def _embed_pilot() -> str:
    """Turn the packaged learned-pilot weights into a C++ byte array linked into the engine."""
    src = os.path.join(PKG, "weights", "pilot_v1.bin")
    with open(src, "rb") as f:
        blob = f.read()
    out = os.path.join(BUILD, "pilot_blob_gen.cpp")
    body = ",".join(str(b) for b in blob)
    with open(out, "w", encoding="utf-8") as f:
        f.write("// generated from cpipe_amd/weights/pilot_v1.bin\n#include <cstddef>\n")
        f.write('extern "C" const unsigned char CPIPE_BM3D_PILOT_BLOB[];\n')
        f.write('extern "C" const size_t CPIPE_BM3D_PILOT_BLOB_SIZE;\n')
        f.write(f'extern "C" const unsigned char CPIPE_BM3D_PILOT_BLOB[] = {{{body}}};\n')
        f.write(f'extern "C" const size_t CPIPE_BM3D_PILOT_BLOB_SIZE = {len(blob)};\n')
    return out


def build(verbose: bool = False) -> dict:
    """Compile every native library of the package; returns their paths."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hipcc = _hipcc()
    manifest_cpp = _embed_manifest()
    blob_cpp = _embed_pilot()
    dev_o = os.path.join(BUILD, "bm3d_device.o")
    pilot_o = os.path.join(BUILD, "pilot.o")
    mx_o = os.path.join(BUILD, "group_mx.o")
    bmx_o = os.path.join(BUILD, "bm_mx.o")
    wav_o = os.path.join(BUILD, "wavelet.o")
    cvt_o = os.path.join(BUILD, "convert.o")
    node_o = os.path.join(BUILD, "node.o")
    dng_o = os.path.join(BUILD, "dng_noise.o")
    man_o = os.path.join(BUILD, "manifest_gen.o")
    blob_o = os.path.join(BUILD, "pilot_blob_gen.o")
    _run([hipcc, *HIP_FLAGS, f"-I{CSRC}", "-c", os.path.join(CSRC, "bm3d_device.hip"), "-o", dev_o],
         verbose)
    _run([hipcc, *HIP_FLAGS, f"-I{CSRC}", "-c", os.path.join(CSRC, "pilot.hip"), "-o", pilot_o], verbose)
    _run([hipcc, *HIP_FLAGS, f"-I{CSRC}", "-c", os.path.join(CSRC, "group_mx.hip"), "-o", mx_o], verbose)
    _run([hipcc, *HIP_FLAGS, f"-I{CSRC}", "-c", os.path.join(CSRC, "bm_mx.hip"), "-o", bmx_o], verbose)
    _run([hipcc, *HIP_FLAGS, f"-I{CSRC}", "-c", os.path.join(CSRC, "wavelet.hip"), "-o", wav_o], verbose)
    _run([hipcc, *HIP_FLAGS, f"-I{CSRC}", "-c", os.path.join(CSRC, "convert.hip"), "-o", cvt_o], verbose)
    _run(["g++", *CXX_FLAGS, *HIP_HOST_FLAGS, "-c", os.path.join(CSRC, "node.cpp"), "-o", node_o], verbose)
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "dng_noise.cpp"), "-o", dng_o], verbose)
    _run(["g++", *CXX_FLAGS, "-c", manifest_cpp, "-o", man_o], verbose)
    _run(["g++", *CXX_FLAGS, "-c", blob_cpp, "-o", blob_o], verbose)
    tmp = ENGINE_SO + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, dev_o, pilot_o, mx_o, bmx_o, wav_o, cvt_o, node_o,
          dng_o, man_o, blob_o], verbose)
    os.replace(tmp, ENGINE_SO)
    tmp = MINIHOST_SO + ".tmp"
    _run(["g++", *CXX_FLAGS, *HIP_HOST_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "minihost.cpp"), "-ldl",
          *HIP_HOST_LIBS], verbose)
    os.replace(tmp, MINIHOST_SO)
    from . import schema  # the pipeline-schema node blocks follow the manifests
    return {"engine": ENGINE_SO, "minihost": MINIHOST_SO, "schema": schema.write()}


if __name__ == "__main__":
    print(build(verbose=True))
