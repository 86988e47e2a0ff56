"""Build libnqpu_hip.so (gfx950) in-tree with hipcc.

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
generation disabled, and linked into ``nqpu_amd/lib/libnqpu_hip.so``.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libnqpu_hip.so")
# host-only C++ (x86 code the device pass must not see), compiled with the host compiler
HOST_SOURCES = ["blob.cpp"]
SOURCES = ["kernels.hip", "engine.hip", "capi.hip", "server.hip", "jit.hip", "reduce.hip", "wire.hip", "tune.hip", "graph.hip"]
# device code that run-time compiled passes embed (jit.hip), in this order
JIT_HEADERS = ["devtypes.h", "regtile.h"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core results go straight to the VGPRs that hold the amplitudes
    # (no accumulator-register copies after every fused block)
    "-mllvm", "-amdgpu-mfma-vgpr-form",
]


def write_jit_source(objdir: str) -> None:
    """jit_src.inc: the text of JIT_HEADERS as one C++ raw string, after
    fixed-width integer typedefs (hipRTC provides no <cstdint>)."""
    parts = ["typedef signed char int8_t;\ntypedef unsigned char uint8_t;\ntypedef int int32_t;\n"
             "typedef unsigned int uint32_t;\ntypedef long long int64_t;\ntypedef unsigned long long uint64_t;\n"]
    for h in JIT_HEADERS:
        with open(os.path.join(CSRC, h)) as f:
            parts.append("".join(l for l in f if not l.startswith("#pragma once")))
    text = "".join(parts)
    assert ")NQPU_JIT\"" not in text
    out = os.path.join(objdir, "jit_src.inc")
    body = "static const char kJitPrelude[] = R\"NQPU_JIT(" + text + ")NQPU_JIT\";\n"
    old = open(out).read() if os.path.exists(out) else None
    if old != body:
        with open(out, "w") as f:
            f.write(body)


def build(verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    write_jit_source(objdir)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-I", objdir, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, src.replace(".cpp", ".o"))
        objs.append(obj)
        cmd = [CXX, "-O3", "-std=c++17", "-fPIC", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature; that note is harmless
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines and (verbose or p.returncode):
            print("\n".join(lines), file=sys.stderr)
        if p.returncode:
            failed.append(src)
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp, "-lhiprtc"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
