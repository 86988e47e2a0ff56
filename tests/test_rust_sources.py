"""The Rust integration sources under rust/ (no Rust toolchain here, so they
are checked as text): the FFI crate declares exactly the functions of
include/nqpu_hip.h, its #[repr(C)] structs list the header's fields in
order, and the reference-side patch applies cleanly to the reference's files
(copied to a temporary directory; skipped where the reference is absent)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nqpu_hip.h")
SYS = os.path.join(ROOT, "rust", "nqpu-hip-sys", "src", "lib.rs")
PATCH = os.path.join(ROOT, "rust", "nqpu-metal", "rocm_backend.patch")
REF = "/root/reference"


def header_functions():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return set(re.findall(r"\b(nqpu_[a-z0-9_]+)\s*\(", text))


def rust_functions():
    text = open(SYS).read()
    block = text[text.index('extern "C" {'):]
    block = block[: block.index("\n}\n")]
    return set(re.findall(r"pub fn (nqpu_[a-z0-9_]+)\s*\(", block))


def test_crate_declares_every_header_function():
    h, r = header_functions(), rust_functions()
    assert h - r == set(), f"missing in the crate: {sorted(h - r)}"
    assert r - h == set(), f"not in the header: {sorted(r - h)}"


def _c_fields(name):
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    m = re.search(r"typedef struct \{([^}]*)\}\s*" + name + ";", text)
    fields = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = decl.split(None, 1)[1] if not decl.startswith("uint64_t submitted") else decl[len("uint64_t "):]
        for part in names.split(","):
            fields.append(re.sub(r"\[.*\]", "", part).strip().split()[-1])
    return fields


def _rust_fields(name):
    text = open(SYS).read()
    m = re.search(r"pub struct " + name + r" \{([^}]*)\}", text)
    return re.findall(r"pub (\w+):", m.group(1))


@pytest.mark.parametrize("name", ["NqpuComplex", "NqpuMeasurement", "NqpuVersion", "NqpuGateOp", "NqpuRunOptions",
                                  "NqpuRunStats", "NqpuFusionCounts", "NqpuPauliTerm", "NqpuBatchStats",
                                  "NqpuServerConfig", "NqpuRequestStatus", "NqpuServerStats",
                                  "NqpuTunedPlan", "NqpuGraphInfo"])
def test_struct_fields_match_the_header(name):
    assert _c_fields(name) == _rust_fields(name)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_reference_patch_applies(tmp_path):
    files = ["sdk/rust/Cargo.toml", "sdk/rust/src/backends/mod.rs", "sdk/rust/src/backends/runtime/auto_simulator.rs"]
    for f in files:
        dst = tmp_path / f
        dst.parent.mkdir(parents=True, exist_ok=True)
        shutil.copy(os.path.join(REF, f), dst)
    r = subprocess.run(["patch", "-p1", "--forward", "-i", PATCH], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    auto = (tmp_path / files[2]).read_text()
    assert auto.count("SimBackend::RocmGPU") >= 4 and "fn execute_rocm_strict" in auto
