"""The reference's own C-ABI tests (sdk/rust/src/infra/bindings/c_ffi.rs
1646-2656), ported one-to-one onto libnqpu_hip.so through ctypes."""
import ctypes as ct
import math

import numpy as np
import pytest

from nqpu_amd import _lib as L

pytestmark = pytest.mark.gpu
EPS = 1e-10


@pytest.fixture(scope="module")
def lib():
    lib = L.load()
    if lib.nqpu_device_count() == 0:
        pytest.fail("no HIP device visible")
    return lib


def amp(lib, s, i):
    out = L.NqpuComplex()
    assert lib.nqpu_state_get_amplitude(s, i, ct.byref(out)) == 0
    return complex(out.re, out.im)


def test_create_destroy_and_initial(lib):
    s = lib.nqpu_state_create(3)
    assert s and lib.nqpu_state_num_qubits(s) == 3
    assert abs(amp(lib, s, 0) - 1) < EPS
    for i in range(1, 8):
        assert abs(amp(lib, s, i)) < EPS
    lib.nqpu_state_destroy(s)


def test_get_set_amplitude(lib):
    s = lib.nqpu_state_create(2)
    v = L.NqpuComplex(0.6, -0.8)
    assert lib.nqpu_state_set_amplitude(s, 3, v) == 0
    assert abs(amp(lib, s, 3) - (0.6 - 0.8j)) < EPS
    out = L.NqpuComplex()
    assert lib.nqpu_state_get_amplitude(s, 4, ct.byref(out)) == 2   # c_ffi.rs:2572
    assert lib.nqpu_state_set_amplitude(s, 4, v) == 2
    lib.nqpu_state_destroy(s)


def test_h_x_cx(lib):
    s = lib.nqpu_state_create(1)
    lib.nqpu_gate_h(s, 0)
    assert abs(amp(lib, s, 0) - 1 / math.sqrt(2)) < EPS and abs(amp(lib, s, 1) - 1 / math.sqrt(2)) < EPS
    lib.nqpu_state_destroy(s)
    s = lib.nqpu_state_create(1)
    lib.nqpu_gate_x(s, 0)
    assert abs(amp(lib, s, 1) - 1) < EPS
    lib.nqpu_state_destroy(s)
    s = lib.nqpu_state_create(2)                       # c_ffi.rs:1752
    lib.nqpu_gate_h(s, 0)
    lib.nqpu_gate_cx(s, 0, 1)
    p = np.empty(4)
    assert lib.nqpu_state_probabilities(s, p.ctypes.data, 4) == 0
    assert np.allclose(p, [0.5, 0, 0, 0.5], atol=EPS)
    assert lib.nqpu_state_probabilities(s, p.ctypes.data, 3) == 2  # c_ffi.rs:2466
    lib.nqpu_state_destroy(s)


def test_rotations(lib):
    s = lib.nqpu_state_create(1)                       # c_ffi.rs:1775
    lib.nqpu_gate_rx(s, 0, math.pi)
    assert abs(amp(lib, s, 1) + 1j) < EPS
    lib.nqpu_state_destroy(s)
    s = lib.nqpu_state_create(1)
    lib.nqpu_gate_ry(s, 0, math.pi)
    assert abs(amp(lib, s, 1) - 1) < EPS
    lib.nqpu_state_destroy(s)
    s = lib.nqpu_state_create(1)
    lib.nqpu_gate_h(s, 0)
    lib.nqpu_gate_rz(s, 0, math.pi / 2)
    assert abs(amp(lib, s, 0) - np.exp(-1j * math.pi / 4) / math.sqrt(2)) < EPS
    assert abs(amp(lib, s, 1) - np.exp(1j * math.pi / 4) / math.sqrt(2)) < EPS
    assert lib.nqpu_gate_rx(s, 0, float("nan")) == 2   # c_ffi.rs:2586
    lib.nqpu_state_destroy(s)


def test_y_z_s_t(lib):
    s = lib.nqpu_state_create(1)                       # c_ffi.rs:2296-2370
    lib.nqpu_gate_y(s, 0)
    assert abs(amp(lib, s, 1) - 1j) < EPS
    lib.nqpu_gate_z(s, 0)
    assert abs(amp(lib, s, 1) + 1j) < EPS
    lib.nqpu_gate_s(s, 0)
    assert abs(amp(lib, s, 1) - 1) < EPS
    lib.nqpu_gate_t(s, 0)
    assert abs(amp(lib, s, 1) - np.exp(1j * math.pi / 4)) < EPS
    lib.nqpu_state_destroy(s)


def test_measure(lib):
    s = lib.nqpu_state_create(1)                       # c_ffi.rs:1852-1889
    m = L.NqpuMeasurement()
    assert lib.nqpu_measure(s, 0, ct.byref(m)) == 0
    assert m.outcome == 0 and abs(m.probability - 1) < EPS
    lib.nqpu_gate_x(s, 0)
    assert lib.nqpu_measure(s, 0, ct.byref(m)) == 0
    assert m.outcome == 1 and abs(m.probability - 1) < EPS
    out = (ct.c_uint32 * 1)()
    assert lib.nqpu_measure_all(s, out, 1) == 0 and out[0] == 1   # c_ffi.rs:2607
    assert lib.nqpu_measure_all(s, out, 2) == 2
    lib.nqpu_state_destroy(s)


def test_expectation_z(lib):
    s = lib.nqpu_state_create(1)
    v = ct.c_double()
    assert lib.nqpu_expectation_z(s, 0, ct.byref(v)) == 0 and abs(v.value - 1) < EPS
    lib.nqpu_gate_x(s, 0)
    assert lib.nqpu_expectation_z(s, 0, ct.byref(v)) == 0 and abs(v.value + 1) < EPS
    assert lib.nqpu_expectation_z(s, 1, ct.byref(v)) == 1
    lib.nqpu_state_destroy(s)


def test_clone_independent_and_reset(lib):
    s = lib.nqpu_state_create(2)                       # c_ffi.rs:1937-1990
    lib.nqpu_gate_h(s, 0)
    c = lib.nqpu_state_clone(s)
    lib.nqpu_gate_x(s, 1)
    assert abs(amp(lib, c, 1) - 1 / math.sqrt(2)) < EPS and abs(amp(lib, c, 3)) < EPS
    assert lib.nqpu_state_reset(s) == 0
    assert abs(amp(lib, s, 0) - 1) < EPS
    lib.nqpu_state_destroy(c)
    lib.nqpu_state_destroy(s)


def test_circuit_execute_and_ghz(lib):
    c = lib.nqpu_circuit_create(3)                     # c_ffi.rs:2007, 2634
    lib.nqpu_circuit_add_h(c, 0)
    lib.nqpu_circuit_add_cx(c, 0, 1)
    lib.nqpu_circuit_add_cx(c, 1, 2)
    s = lib.nqpu_state_create(3)
    assert lib.nqpu_circuit_execute(c, s) == 0
    assert abs(amp(lib, s, 0) - 1 / math.sqrt(2)) < EPS and abs(amp(lib, s, 7) - 1 / math.sqrt(2)) < EPS
    small = lib.nqpu_state_create(2)
    assert lib.nqpu_circuit_execute(c, small) == 2
    lib.nqpu_state_destroy(small)
    lib.nqpu_state_destroy(s)
    lib.nqpu_circuit_destroy(c)


def test_invalid_qubit_and_same_qubit(lib):
    s = lib.nqpu_state_create(2)                       # c_ffi.rs:2107, 2414
    assert lib.nqpu_gate_h(s, 2) == 1
    assert lib.nqpu_gate_cx(s, 0, 5) == 1
    assert lib.nqpu_gate_cx(s, 1, 1) == 2
    assert lib.nqpu_gate_swap(s, 0, 0) == 2
    assert lib.nqpu_gate_custom_1q(s, 0, None) == 5
    lib.nqpu_state_destroy(s)


def test_custom_gates(lib):
    s = lib.nqpu_state_create(1)                       # c_ffi.rs:2135
    m = (L.NqpuComplex * 4)(L.NqpuComplex(1, 0), L.NqpuComplex(0, 0), L.NqpuComplex(0, 0), L.NqpuComplex(1, 0))
    assert lib.nqpu_gate_custom_1q(s, 0, m) == 0
    assert abs(amp(lib, s, 0) - 1) < EPS
    lib.nqpu_state_destroy(s)
    s = lib.nqpu_state_create(2)                       # c_ffi.rs:2168
    lib.nqpu_gate_x(s, 1)
    z, o = L.NqpuComplex(0, 0), L.NqpuComplex(1, 0)
    sw = (L.NqpuComplex * 16)(o, z, z, z, z, z, o, z, z, o, z, z, z, z, z, o)
    assert lib.nqpu_gate_custom_2q(s, 1, 0, sw) == 0
    assert abs(amp(lib, s, 1) - 1) < EPS and abs(amp(lib, s, 2)) < EPS
    lib.nqpu_state_destroy(s)


def test_custom_2q_every_group_once(lib, oracle):
    """c_ffi.rs:319-399 re-applies U to some groups when min(q1,q2) >= 1; the
    ABI here transforms every group exactly once (checked against a dense
    application of the 4x4 on qubits (3, 1) of a 5-qubit state)."""
    from tests.golden import make_golden as mg
    rng = np.random.default_rng(8)
    u = mg.random_unitary(rng, 4)
    a0 = rng.normal(size=32) + 1j * rng.normal(size=32)
    a0 /= np.linalg.norm(a0)
    s = lib.nqpu_state_create(5)
    buf = np.ascontiguousarray(a0)
    assert lib.nqpu_state_set_amplitudes(s, buf.ctypes.data, 32) == 0
    m = (L.NqpuComplex * 16)(*[L.NqpuComplex(v.real, v.imag) for v in u.reshape(-1)])
    assert lib.nqpu_gate_custom_2q(s, 3, 1, m) == 0
    out = np.empty(32, np.complex128)
    assert lib.nqpu_state_get_amplitudes(s, out.ctypes.data, 32) == 0
    assert np.max(np.abs(out - mg.apply(a0, 5, u, [3, 1]))) < 1e-12
    lib.nqpu_state_destroy(s)


def test_sample_bell_and_seed_determinism(lib):
    lib.nqpu_set_seed(42)                              # c_ffi.rs:2240
    s = lib.nqpu_state_create(2)
    lib.nqpu_gate_h(s, 0)
    lib.nqpu_gate_cx(s, 0, 1)
    out = np.zeros(100, np.uint64)
    assert lib.nqpu_sample(s, 100, out.ctypes.data) == 0
    assert set(out.tolist()) == {0, 3}
    lib.nqpu_state_destroy(s)

    def run(seed):                                     # c_ffi.rs:2478
        lib.nqpu_set_seed(seed)
        s = lib.nqpu_state_create(2)
        lib.nqpu_gate_h(s, 0)
        lib.nqpu_gate_h(s, 1)
        r = np.zeros(50, np.uint64)
        lib.nqpu_sample(s, 50, r.ctypes.data)
        lib.nqpu_state_destroy(s)
        return r
    r1, r2, r3 = run(12345), run(12345), run(99999)
    assert np.array_equal(r1, r2) and not np.array_equal(r1, r3)


def test_norm_preserved_and_self_inverse(lib):
    s = lib.nqpu_state_create(3)                       # c_ffi.rs:2274
    lib.nqpu_gate_h(s, 0)
    lib.nqpu_gate_cx(s, 0, 1)
    lib.nqpu_gate_ry(s, 2, 1.234)
    lib.nqpu_gate_rz(s, 0, 0.567)
    lib.nqpu_gate_t(s, 1)
    lib.nqpu_gate_s(s, 2)
    assert abs(lib.nqpu_state_norm(s) - 1) < 1e-12
    lib.nqpu_state_destroy(s)
    s = lib.nqpu_state_create(3)                       # c_ffi.rs:2436-2465, 2521
    lib.nqpu_gate_h(s, 1)
    assert abs(amp(lib, s, 2) - 1 / math.sqrt(2)) < EPS
    lib.nqpu_gate_h(s, 1)
    lib.nqpu_gate_x(s, 2)
    lib.nqpu_gate_x(s, 2)
    assert abs(amp(lib, s, 0) - 1) < EPS
    lib.nqpu_state_destroy(s)


def test_swap_and_cz(lib):
    s = lib.nqpu_state_create(2)                       # c_ffi.rs:2371-2413
    lib.nqpu_gate_x(s, 0)
    lib.nqpu_gate_swap(s, 0, 1)
    assert abs(amp(lib, s, 2) - 1) < EPS
    lib.nqpu_gate_x(s, 0)
    lib.nqpu_gate_cz(s, 0, 1)
    assert abs(amp(lib, s, 3) + 1) < EPS
    lib.nqpu_state_destroy(s)


def test_extended_precision_and_cap(lib):
    s = lib.nqpu_state_create_ex(31, L.PRECISION_C64)  # above the reference's 30-qubit cap
    assert s and lib.nqpu_state_precision(s) == L.PRECISION_C64
    assert abs(lib.nqpu_state_norm(s) - 1) < 1e-12
    lib.nqpu_state_destroy(s)
