"""The continuous-batching request server (nqpu_server_*) on the GPU.

* Alone vs mixed: every request of a stream gives bit-identical expectation
  values and sampled indices whether it runs alone (one slot, one request at
  a time) or concurrently with the others (8 slots, all submitted at once).
* Parity: a complex128 server's expectations match the FP64 oracle to 1e-12,
  and its shots equal the oracle's restatement of the C ABI sampler applied
  to the same state with the request's seed, index for index.
* Admission: with an arena that holds two 22-qubit states, no more than two
  run at once, the arena is never exceeded, and every request completes.
"""
import numpy as np
import pytest

from nqpu_amd import _lib as L
from nqpu_amd import engine as E
from nqpu_amd.engine import Server, StateVector
from nqpu_amd.workloads import served_stream

pytestmark = pytest.mark.gpu

GiB = 1 << 30


@pytest.fixture(scope="module", autouse=True)
def device():
    if E.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X")


@pytest.fixture(scope="module")
def stream():
    return served_stream(24, 100.0, seed=104, shots=256, qubits=(10, 14, 18, 22),
                         weights=(0.3, 0.3, 0.25, 0.15), dmin=10, dmax=30)


def run_alone(reqs, precision="c64"):
    out = []
    with Server(slots=1, precision=precision, memory_bytes=GiB) as srv:
        for r in reqs:
            rid = srv.submit(r.circuit, r.terms, r.shots, r.seed)
            res = srv.result(rid, timeout=120)
            assert res["state"] == "done"
            out.append((res["expectations"], res["samples"]))
    return out


def test_alone_equals_mixed(stream):
    alone = run_alone(stream)
    with Server(slots=8, precision="c64", memory_bytes=GiB) as srv:
        ids = [srv.submit(r.circuit, r.terms, r.shots, r.seed) for r in stream]
        mixed = [srv.result(i, timeout=120) for i in ids]
        st = srv.stats()
    assert st["completed"] == len(stream) and st["failed"] == 0
    assert st["peak_running"] > 1  # requests really shared the device
    for (e, s), m in zip(alone, mixed):
        assert m["state"] == "done"
        assert np.array_equal(e, m["expectations"])
        assert np.array_equal(s, m["samples"])


def test_complex128_server_matches_oracle(oracle, stream):
    reqs = [r for r in stream if r.circuit.num_qubits <= 14][:6]
    with Server(slots=4, precision="c128", memory_bytes=GiB) as srv:
        ids = [srv.submit(r.circuit, r.terms, r.shots, r.seed) for r in reqs]
        res = [srv.result(i, timeout=120) for i in ids]
    for r, got in zip(reqs, res):
        n = r.circuit.num_qubits
        a = oracle.run(n, r.circuit.gates)
        want = [c * oracle.expectation_pauli(a, n, *E.pauli_masks(p)) for c, p in r.terms]
        assert np.max(np.abs(np.array(want) - got["expectations"])) < 1e-12
        with StateVector(n, "c128") as sv:
            sv.run(r.circuit)
            amps = sv.amplitudes()
        assert np.array_equal(got["samples"], oracle.sample_cabi(np.ascontiguousarray(amps), n, r.shots,
                                                                   seed=r.seed))


def test_admission_by_memory():
    reqs = served_stream(6, 100.0, seed=7, shots=64, qubits=(22,), weights=(1.0,), dmin=10, dmax=20)
    arena = 2 * (8 << 22)  # two 22-qubit complex64 states
    with Server(slots=8, precision="c64", memory_bytes=arena) as srv:
        ids = [srv.submit(r.circuit, r.terms, r.shots, r.seed) for r in reqs]
        res = [srv.result(i, timeout=120) for i in ids]
        st = srv.stats()
    assert all(x["state"] == "done" for x in res)
    assert st["peak_running"] <= 2
    assert st["peak_bytes"] <= arena == st["memory_bytes"]
    ref = run_alone(reqs)
    for (e, s), m in zip(ref, res):
        assert np.array_equal(e, m["expectations"]) and np.array_equal(s, m["samples"])


def test_state_larger_than_arena_is_refused():
    from nqpu_amd.workloads import random_u3_brickwork
    with Server(slots=2, precision="c64", memory_bytes=64 << 20) as srv:
        with pytest.raises(L.NqpuError) as ei:
            srv.submit(random_u3_brickwork(24, 2, 1), [], 0, 0)  # 128 MiB > 64 MiB
        assert ei.value.code == 8


def test_destroy_with_queued_requests():
    reqs = served_stream(8, 100.0, seed=9, shots=16, qubits=(18,), weights=(1.0,), dmin=10, dmax=20)
    srv = Server(slots=1, precision="c64", memory_bytes=GiB)
    for r in reqs:
        srv.submit(r.circuit, r.terms, r.shots, r.seed)
    srv.close()  # finishes the running request, fails the queued ones, returns


def test_26_qubit_1024_shot_requests_match_oracle(oracle):
    """The served workload's larger requests: 26 qubits, 1024 shots and
    8 random Pauli strings each, complex64, against the FP64 oracle
    (expectations) and the C ABI sampler on the request's own final state."""
    reqs = served_stream(2, 100.0, seed=26, shots=1024, qubits=(26,), weights=(1.0,), dmin=10, dmax=14)
    with Server(slots=2, precision="c64", memory_bytes=4 * GiB) as srv:
        ids = [srv.submit(r.circuit, r.terms, r.shots, r.seed) for r in reqs]
        res = [srv.result(i, timeout=120) for i in ids]
    for r, got in zip(reqs, res):
        assert got["state"] == "done"
        n = r.circuit.num_qubits
        assert n == 26 and r.shots == 1024 and len(r.terms) == 8
        a = oracle.run(n, r.circuit.gates)
        want = [c * oracle.expectation_pauli(a, n, *E.pauli_masks(p)) for c, p in r.terms]
        assert np.max(np.abs(np.array(want) - got["expectations"])) < 1e-5
        with StateVector(n, "c64") as sv:
            sv.run(r.circuit)
            amps = sv.amplitudes()
        assert np.array_equal(got["samples"], oracle.sample_cabi(np.ascontiguousarray(amps), n, r.shots,
                                                                   seed=r.seed))
