"""C-ABI library tests that need no GPU: the library loads and exports every
symbol include/nqpu_hip.h declares, argument validation and error codes
follow the reference's C ABI (sdk/rust/src/infra/bindings/c_ffi.rs tests
1646-2656), circuit bookkeeping, the reference fusion counts
(gate_fusion.rs tests 1205-1700) and the tile planner.  No compute call is
made on a device."""
import ctypes as ct
import math
import os
import re
import subprocess
import sys

import pytest

from nqpu_amd import _lib as L
from nqpu_amd.engine import Circuit
from nqpu_amd.workloads import random_u3_brickwork, qft_trotter

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    return L.load()


def header_functions():
    text = open(os.path.join(ROOT, "include", "nqpu_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nqpu_\w+)\s*\(", text)))


def test_every_header_symbol_exported(lib):
    names = header_functions()
    assert len(names) >= 55
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and every binding the Python side declares is in the header
    assert set(L.SIGNATURES) <= set(names)


def test_version_and_messages(lib):
    v = lib.nqpu_version()
    assert (v.major, v.minor, v.patch) == (0, 1, 0)  # c_ffi.rs:2207
    for code in range(10):
        assert lib.nqpu_error_message(code)  # c_ffi.rs:2218
    assert lib.nqpu_error_message(1) == b"Invalid qubit index"


def test_no_device_is_reported_not_faked():
    code = ("import ctypes as ct; from nqpu_amd import _lib as L; lib=L.load();"
            "p=lib.nqpu_state_create(3); assert not p;"
            "assert b'no usable' in lib.nqpu_last_error_detail();"
            "assert not lib.nqpu_server_create(None);"
            "assert b'no usable' in lib.nqpu_last_error_detail();"
            "assert not lib.nqpu_batch_create();"
            "assert lib.nqpu_device_count()==0; print('ok')")
    env = dict(os.environ, NQPU_HIP_FORCE_NO_DEVICE="1")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr


def test_invalid_create_and_null_handles(lib):
    assert not lib.nqpu_state_create(0)            # c_ffi.rs:2426
    assert not lib.nqpu_state_create(31)
    assert not lib.nqpu_state_create_ex(35, L.PRECISION_C64)
    assert not lib.nqpu_state_create_ex(34, L.PRECISION_C128)
    assert not lib.nqpu_state_create_ex(10, 7)
    assert lib.nqpu_state_norm(None) == -1.0       # c_ffi.rs:2563
    assert lib.nqpu_state_num_qubits(None) == 0
    lib.nqpu_state_destroy(None)
    lib.nqpu_circuit_destroy(None)
    NP = 5
    assert lib.nqpu_gate_h(None, 0) == NP          # c_ffi.rs:2068
    assert lib.nqpu_gate_rx(None, 0, 0.1) == NP
    assert lib.nqpu_gate_cx(None, 0, 1) == NP
    assert lib.nqpu_state_reset(None) == NP
    assert lib.nqpu_sample(None, 4, None) == NP
    assert lib.nqpu_circuit_execute(None, None) == NP
    assert lib.nqpu_circuit_gate_count(None) == 0
    assert lib.nqpu_circuit_depth(None) == 0
    assert lib.nqpu_expectation_pauli(None, 0, 0, None) == NP
    rid = ct.c_uint64()
    assert lib.nqpu_server_submit(None, None, None, 0, 0, 0, ct.byref(rid)) == NP
    assert lib.nqpu_server_poll(None, 1, None) == NP
    assert lib.nqpu_server_wait(None, 1, 0, None) == NP
    assert lib.nqpu_server_take(None, 1, None, 0, None, 0) == NP
    assert lib.nqpu_server_stats(None, None) == NP
    lib.nqpu_server_destroy(None)


def test_circuit_api(lib):
    c = lib.nqpu_circuit_create(3)                 # c_ffi.rs:1991-2066
    assert c
    assert lib.nqpu_circuit_add_h(c, 0) == 0
    assert lib.nqpu_circuit_add_cx(c, 0, 1) == 0
    assert lib.nqpu_circuit_add_rx(c, 2, 0.5) == 0
    assert lib.nqpu_circuit_gate_count(c) == 3
    assert lib.nqpu_circuit_depth(c) == 2
    assert lib.nqpu_circuit_add_h(c, 3) == 1       # INVALID_QUBIT
    assert lib.nqpu_circuit_add_cx(c, 1, 1) == 2   # same qubit: INVALID_PARAMETER
    assert lib.nqpu_circuit_add_cx(c, 0, 9) == 1
    assert lib.nqpu_circuit_add_rx(c, 0, float("nan")) == 2  # c_ffi.rs:2586
    assert lib.nqpu_circuit_add_rx(c, 0, float("inf")) == 2
    lib.nqpu_circuit_destroy(c)
    c = lib.nqpu_circuit_create(4)                 # c_ffi.rs:2542
    for i in range(4):
        lib.nqpu_circuit_add_h(c, i)
    assert lib.nqpu_circuit_depth(c) == 1
    lib.nqpu_circuit_destroy(c)
    c = lib.nqpu_circuit_create(2)                 # c_ffi.rs:2622
    assert lib.nqpu_circuit_depth(c) == 0
    lib.nqpu_circuit_destroy(c)
    assert not lib.nqpu_circuit_create(0)
    assert not lib.nqpu_circuit_create(31)
    assert lib.nqpu_circuit_create_ex(34)


def _op(kind, t, c=(), p=()):
    op = L.NqpuGateOp()
    op.kind = L.GATE_ID[kind]
    op.num_targets, op.num_controls = len(t), len(c)
    for i, q in enumerate(t):
        op.targets[i] = q
    for i, q in enumerate(c):
        op.controls[i] = q
    for i, v in enumerate(p):
        op.params[i] = v
    return op


def test_add_gate_validation(lib):
    c = lib.nqpu_circuit_create_ex(5)
    ok = [_op("u", (0,), (), (1, 2, 3)), _op("toffoli", (2,), (0, 1)), _op("cswap", (3, 4), (0,)),
          _op("rzz", (1, 4), (), (0.3,)), _op("cu", (1,), (2,), (1, 2, 3, 4))]
    for op in ok:
        assert lib.nqpu_circuit_add_gate(c, ct.byref(op)) == 0
    assert lib.nqpu_circuit_add_gate(c, ct.byref(_op("cnot", (1,), (1,)))) == 2     # repeated qubit
    assert lib.nqpu_circuit_add_gate(c, ct.byref(_op("cnot", (1,), ()))) == 2       # missing control
    assert lib.nqpu_circuit_add_gate(c, ct.byref(_op("h", (5,)))) == 1              # out of range
    assert lib.nqpu_circuit_add_gate(c, ct.byref(_op("rz", (0,), (), (math.nan,)))) == 2
    bad = _op("h", (0,))
    bad.kind = 99
    assert lib.nqpu_circuit_add_gate(c, ct.byref(bad)) == 2
    assert lib.nqpu_circuit_gate_count(c) == len(ok)
    lib.nqpu_circuit_destroy(c)


def counts(circ):
    return circ.native().fusion_counts()


def test_reference_fusion_known_counts():
    c = Circuit(1).h(0).h(0)                       # test_h_h_cancels_to_identity
    r = counts(c)
    assert r["gates_eliminated"] == 2 and r["fused_count"] == 0 and r["items"] == 0
    r = counts(Circuit(1).rx(0, 0.3).rx(0, 0.5))   # test_rx_merge
    assert r["fused_count"] == 1 and r["items"] == 1
    r = counts(Circuit(2).h(0).h(1).cnot(0, 1).h(0))  # test_multi_qubit_flushes_correctly
    assert r["items"] == 2 and r["fused_count"] == 1
    r = counts(Circuit(1).x(0).x(0))               # test_x_x_cancels
    assert r["gates_eliminated"] == 2
    r = counts(Circuit(2).h(0).rx(1, 0.5).h(0).rx(1, 0.3))  # test_mixed_qubits_fuse_independently
    assert r["gates_eliminated"] >= 2 and r["fused_count"] >= 1
    r = counts(Circuit(1).h(0))                    # test_single_gate_preserved
    assert r["items"] == 1 and r["fused_count"] == 0
    r = counts(Circuit(1))                         # test_empty_circuit
    assert r == {"original_count": 0, "fused_count": 0, "gates_eliminated": 0, "items": 0}
    r = counts(Circuit(2).h(0).cnot(0, 1))         # test_h_absorbed_into_cnot
    assert r["items"] == 1 and r["fused_count"] == 1
    r = counts(Circuit(2).h(0).h(1).cnot(0, 1))    # test_both_qubits_absorbed
    assert r["items"] == 1
    r = counts(Circuit(3).h(0).toffoli(0, 1, 2))   # test_3qubit_gate_not_absorbed
    assert r["items"] == 2 and r["fused_count"] == 0
    r = counts(Circuit(2).h(0).swap(0, 1))         # test_swap_absorption
    assert r["items"] == 1 and r["fused_count"] == 1
    # diagonal runs (fuse_diagonal_pass, gate_fusion.rs:757): z, s, cz, t merge into one item
    r = counts(Circuit(3).cnot(0, 1).z(0).cnot(1, 2).s(2).cz(0, 2).t(1).h(0))
    assert r["original_count"] == 7


def test_tiled_planner_reduces_sweeps():
    c = random_u3_brickwork(30, 40, 102).native()
    none = c.plan("c64", "none")
    ref = c.plan("c64", "reference")
    tiled = c.plan("c64", "tiled")
    assert none["passes"] == none["gates"] == 30 * 40 + 20 * 14 + 20 * 15
    assert ref["passes"] < none["passes"]
    assert tiled["passes"] * 5 < ref["passes"]
    # U3s stay separate from the CNOTs; qubits 0 and 29 idle in odd layers, so
    # their U3 merges with the next layer's (19 odd layers before the last)
    assert tiled["device_ops"] == 1200 + 580 - 2 * 19
    assert tiled["bytes"] == tiled["passes"] * 2 * 8 * 2 ** 30
    # a larger tile never needs more passes here
    p12 = c.plan("c64", "tiled", tile_qubits=12)["passes"]
    p10 = c.plan("c64", "tiled", tile_qubits=10)["passes"]
    assert p12 <= p10


def test_planner_on_diagonal_heavy_circuit():
    c = qft_trotter(20, 4).native()
    tiled = c.plan("c128", "tiled")
    none = c.plan("c128", "none")
    assert tiled["passes"] < none["passes"] / 5
