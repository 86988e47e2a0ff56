"""Request formats on the GPU: nqpu_execute_json answers a CircuitRequest as
ExecuteResponse::execute does (web/mod.rs:288-375): probabilities against the
FP64 oracle (complex128, 1e-12) and the fidelity of the fused run against the
one-pass-per-gate run; the request server answers a SampleRequest as
SampleResponse::sample (web/mod.rs:378-419): MSB-first bitstring counts equal
to the histogram of the same request submitted as a native circuit with the
same seed, and observables against the oracle."""
import json
from collections import Counter

import numpy as np
import pytest

from nqpu_amd import _lib as L
from nqpu_amd import engine as E
from nqpu_amd.engine import Server
from tests.test_wire_host import SERDE, SERDE_GATES, WEB, WEB_GATES

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if E.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X")


def test_execute_json_matches_oracle(oracle):
    for text, gates, n in ((json.dumps(WEB), WEB_GATES, 4), (SERDE, SERDE_GATES, 5)):
        r = E.execute_json(text)
        want = oracle.probabilities(oracle.run(n, gates), n)
        assert np.max(np.abs(np.array(r["probabilities"]) - want)) < 1e-12
        assert abs(r["fidelity"] - 1.0) < 1e-12
        assert r["num_gates"] == len(gates) and r["backend_used"] and r["time_ms"] >= 0


def test_execute_json_buffer_protocol():
    lib = L.load()
    import ctypes as ct
    need = ct.c_uint64()
    text = json.dumps(WEB).encode()
    assert lib.nqpu_execute_json(text, 2, None, 0, ct.byref(need)) == 2
    buf = ct.create_string_buffer(int(need.value))
    assert lib.nqpu_execute_json(text, 2, buf, len(buf), ct.byref(need)) == 0
    assert json.loads(buf.value.decode())["num_gates"] == len(WEB_GATES)


def test_sample_request_through_the_server(oracle):
    from nqpu_amd.workloads import random_u3_brickwork
    n = 12
    c = random_u3_brickwork(n, 6, 77)
    # U3 has no web-format form, so the brickwork goes in circuit_serde's form
    obs = [{"pauli": "ZZ" + "I" * (n - 2), "coeff": -1.0}, {"pauli": "X" + "I" * (n - 1)}]
    serde = {"version": 1, "num_qubits": n, "shots": 2048, "observables": obs,
             "gates": [{"gate_type": "U" if g.kind == "u" else "CNOT", "targets": list(g.targets),
                        "controls": list(g.controls), "params": list(g.params)} for g in c.gates]}
    with Server(slots=4, precision="c128", memory_bytes=1 << 30) as srv:
        a = srv.submit_json(json.dumps(serde), seed=99)
        b = srv.submit(c, [(-1.0, "ZZ" + "I" * (n - 2)), (1.0, "X" + "I" * (n - 1))], 2048, 99)
        got = srv.take_json(a, timeout=120)
        ref = srv.result(b, timeout=120)
    want = Counter(format(int(i), f"0{n}b") for i in ref["samples"])
    assert got["outcomes"] == dict(want) and got["shots"] == 2048
    assert all(len(k) == n for k in got["outcomes"])
    st = oracle.run(n, c.gates)
    zz = oracle.expectation_pauli(st, n, *E.pauli_masks("ZZ" + "I" * (n - 2)))
    xx = oracle.expectation_pauli(st, n, *E.pauli_masks("X" + "I" * (n - 1)))
    assert abs(got["expectations"][0] + zz) < 1e-12 and abs(got["expectations"][1] - xx) < 1e-12


def test_server_take_json_keeps_the_response_for_a_larger_buffer():
    import ctypes as ct
    lib = L.load()
    text = json.dumps(dict(WEB, shots=500))
    with Server(slots=2, precision="c64", memory_bytes=1 << 28) as srv:
        rid = srv.submit_json(text, seed=5)
        need = ct.c_uint64()
        small = ct.create_string_buffer(8)
        assert lib.nqpu_server_take_json(srv.ptr, rid, L.WAIT_FOREVER, small, 8, ct.byref(need)) == 2
        assert need.value > 8
        full = srv.take_json(rid)
        assert sum(full["outcomes"].values()) == 500
        with pytest.raises(L.NqpuError):
            srv.submit_json('{"num_qubits": 2, "gates": []}')  # neither shots nor observables
