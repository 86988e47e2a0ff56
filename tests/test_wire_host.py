"""Request formats on the host (no device): JSON in the web API's
CircuitRequest/SampleRequest form (sdk/rust/src/web/mod.rs:34-160) and in
circuit_serde's SerializableCircuit form (circuit_serde.rs:329-340, 435-558)
parse into the circuit the reference's constructors build.  Checked by
running the parsed circuit through the FP64 host emulation of the tiled
execution and the same gates through the oracle, to 1e-12, and by the
rejection of malformed requests."""
import json
import math

import numpy as np
import pytest

from nqpu_amd import _lib as L
from nqpu_amd import engine as E
from nqpu_amd.engine import Gate

WEB = {
    "num_qubits": 4,
    "backend": "auto",
    "shots": 1000,
    "gates": [
        {"type": "h", "target": 0}, {"type": "x", "target": 1}, {"type": "y", "target": 2},
        {"type": "z", "target": 3}, {"type": "s", "target": 0}, {"type": "t", "target": 1},
        {"type": "sx", "target": 2}, {"type": "rx", "target": 3, "angle": 0.3},
        {"type": "ry", "target": 0, "angle": -1.1}, {"type": "rz", "target": 1, "angle": 2.5},
        {"type": "phase", "target": 2, "angle": 0.7}, {"type": "cnot", "control": 0, "target": 3},
        {"type": "cz", "control": 2, "target": 1}, {"type": "swap", "q0": 1, "q1": 3},
        {"type": "cr", "control": 3, "target": 0, "angle": 1.3}, {"type": "iswap", "q0": 0, "q1": 2},
        {"type": "toffoli", "control1": 2, "control2": 3, "target": 1}, {"type": "h", "target": 3},
    ],
}

# the reference's Gate constructors for the same list (gates.rs:669-760,
# web/mod.rs:86-120)
WEB_GATES = [
    Gate("h", (0,)), Gate("x", (1,)), Gate("y", (2,)), Gate("z", (3,)), Gate("s", (0,)), Gate("t", (1,)),
    Gate("sx", (2,)), Gate("rx", (3,), (), (0.3,)), Gate("ry", (0,), (), (-1.1,)), Gate("rz", (1,), (), (2.5,)),
    Gate("phase", (2,), (), (0.7,)), Gate("cnot", (3,), (0,)), Gate("cz", (1,), (2,)), Gate("swap", (1, 3)),
    Gate("cr", (0,), (3,), (1.3,)), Gate("iswap", (0, 2)), Gate("toffoli", (1,), (2, 3)), Gate("h", (3,)),
]

# circuit_serde's pretty-printed layout (circuit_to_json, gate_to_json,
# metadata_to_json; integral params as "1.0")
SERDE = """{
  "version": 1,
  "num_qubits": 5,
  "gates": [
    {"gate_type": "H", "targets": [0], "controls": [], "params": []},
    {"gate_type": "U", "targets": [4], "controls": [], "params": [0.4, 1.0, -2.2]},
    {"gate_type": "CNOT", "targets": [1], "controls": [0], "params": []},
    {"gate_type": "CRz", "targets": [2], "controls": [1], "params": [0.9]},
    {"gate_type": "CRy", "targets": [3], "controls": [4], "params": [-0.6]},
    {"gate_type": "Rzz", "targets": [0, 3], "controls": [], "params": [0.25]},
    {"gate_type": "CCZ", "targets": [2], "controls": [0, 4], "params": []},
    {"gate_type": "CSWAP", "targets": [1, 4], "controls": [0], "params": []},
    {"gate_type": "SX", "targets": [3], "controls": [], "params": []},
    {"gate_type": "CU", "targets": [0], "controls": [2], "params": [0.1, 0.2, 0.3, 0.4]}
  ],
  "metadata": {"name": "t\\u00e9st", "description": "", "author": "", "creation_date": "", "qubit_count": 5, "gate_count": 10, "depth": 6, "tags": ["a","b"]}
}"""

SERDE_GATES = [
    Gate("h", (0,)), Gate("u", (4,), (), (0.4, 1.0, -2.2)), Gate("cnot", (1,), (0,)),
    Gate("crz", (2,), (1,), (0.9,)), Gate("cry", (3,), (4,), (-0.6,)), Gate("rzz", (0, 3), (), (0.25,)),
    Gate("ccz", (2,), (0, 4)), Gate("cswap", (1, 4), (0,)), Gate("sx", (3,)),
    Gate("cu", (0,), (2,), (0.1, 0.2, 0.3, 0.4)),
]


def _check(text, gates, n, oracle):
    nc, shots = E.circuit_from_json(text)
    a0 = np.zeros(1 << n, complex)
    a0[0] = 1
    got = nc.emulate(a0, precision="c128")
    want = oracle.run(n, gates)
    assert np.max(np.abs(got - want)) < 1e-12
    return shots


def test_web_request_every_gate_type(oracle):
    assert _check(json.dumps(WEB), WEB_GATES, 4, oracle) == 1000


def test_circuit_serde_layout(oracle):
    assert _check(SERDE, SERDE_GATES, 5, oracle) == 0
    compact = json.dumps(json.loads(SERDE), separators=(",", ":"))
    assert _check(compact, SERDE_GATES, 5, oracle) == 0


@pytest.mark.parametrize("text,code", [
    ('{"num_qubits": 2, "gates": [{"type": "hadamard", "target": 0}]}', 2),      # unknown type
    ('{"num_qubits": 2, "gates": [{"type": "cnot", "control": 0}]}', 2),         # missing field
    ('{"num_qubits": 2, "gates": [{"type": "h", "target": 1.0}]}', 2),           # index not an integer
    ('{"num_qubits": 2, "gates": [{"type": "h", "target": 2}]}', 1),             # qubit out of range
    ('{"num_qubits": 2, "gates": [{"type": "cnot", "control": 1, "target": 1}]}', 2),
    ('{"num_qubits": 2, "gates": [{"type": "rx", "target": 0}]}', 2),            # missing angle
    ('{"num_qubits": 2, "gates": []} x', 2),                                       # trailing characters
    ('{"num_qubits": 2, "gates": [', 2),                                           # truncated
    ('{"num_qubits": 0, "gates": []}', 1),
    ('{"num_qubits": 40, "gates": []}', 1),
    ('{"gates": []}', 2),
    ('{"num_qubits": 2, "gates": [], "backend": "quantum"}', 2),
    ('{"num_qubits": 2, "gates": [{"gate_type": "Rx", "targets": [0], "controls": [], "params": []}]}', 2),
    ('{"num_qubits": 2, "gates": [{"gate_type": "Custom", "targets": [0]}]}', 2),
    ('{"version": 2, "num_qubits": 2, "gates": []}', 2),
    ('[1, 2]', 2),
])
def test_malformed_requests_are_rejected(text, code):
    with pytest.raises(L.NqpuError) as ei:
        E.circuit_from_json(text)
    assert ei.value.code == code


def test_empty_and_unicode_requests():
    nc, shots = E.circuit_from_json('{"num_qubits": 3, "gates": [], "shots": 5, "extra": {"k": [null, true]}}')
    assert shots == 5 and nc.num_qubits == 3
    nc, _ = E.circuit_from_json('{"num_qubits": 1, "gates": [{"type": "h", "target": 0, "label": "\\ud83d\\ude00"}]}')
    assert nc.plan("c64")["gates"] == 1
