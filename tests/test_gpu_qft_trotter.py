"""QFT followed by first-order Trotter steps of the 1D transverse-field Ising
chain (dt = 0.05, g = 1), the memory-ceiling circuit family, on the GPU.

* At 18 qubits the whole final state is compared with the FP64 oracle
  (complex128 within 1e-12 per amplitude, complex64 infidelity <= 1e-6).
* At 28 qubits and at the full 34-qubit complex64 / 33-qubit complex128
  sizes, <Z0 Z1> is compared with the oracle's value at 22 qubits.  That
  value is exact at every size >= 22: QFT|0...0> is the product state |+>^n,
  and each Trotter step widens the support of Z0 Z1 (Heisenberg picture) by
  one qubit, so after 20 steps it covers qubits 0..21 only
  (test_oracle.test_trotter_light_cone pins this on the oracle itself).
  Tolerances: complex128 1e-12, complex64 1e-5; norms 1e-11 / 1e-4.
"""
import numpy as np
import pytest

from nqpu_amd import engine as E
from nqpu_amd import _lib as L
from nqpu_amd.engine import StateVector
from nqpu_amd.workloads import qft_trotter

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if E.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X")


@pytest.fixture(scope="module")
def z0z1_exact(oracle):
    a = oracle.run(22, qft_trotter(22, 20).gates)
    return oracle.expectation_pauli(a, 22, 0, 3)


def z0z1(sv, n):
    return sv.expectation_pauli("ZZ" + "I" * (n - 2))


@pytest.mark.parametrize("prec,lds", [("c128", False), ("c64", False), ("c64", True)])
def test_state_matches_oracle(oracle, prec, lds):
    n = 18
    c = qft_trotter(n, 20)
    exp = oracle.run(n, c.gates)
    with StateVector(n, prec) as sv:
        sv.run(c, lds_tiles=lds)
        out = sv.amplitudes()
    if prec == "c128":
        assert np.max(np.abs(out - exp)) < 1e-12
    else:
        inf = 1.0 - abs(np.vdot(out, exp)) ** 2 / (np.vdot(out, out).real * np.vdot(exp, exp).real)
        assert inf < 1e-6


def check_z0z1(sv, n, prec, exact):
    """complex128: the raw state.  complex64: every RX (and RZZ) of the
    circuit scales the norm by the same FP32-rounded cos^2 + sin^2, so the
    raw norm drifts by ~1e-5 over the 560 Trotter rotations; the FP64
    renormalisation on the GPU removes that global factor."""
    nrm = sv.norm()
    if prec == "c128":
        assert abs(nrm - 1.0) < 1e-11
        assert abs(z0z1(sv, n) - exact) < 1e-12, (z0z1(sv, n), exact)
    else:
        assert abs(nrm - 1.0) < 1e-4
        sv.normalize()
        assert abs(sv.norm() - 1.0) < 1e-6
        got = z0z1(sv, n)
        assert abs(got - exact) < 1e-5, (got, exact)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_light_cone_value_at_28_qubits(z0z1_exact, prec):
    n = 28
    with StateVector(n, prec) as sv:
        sv.run(qft_trotter(n, 20))
        check_z0z1(sv, n, prec, z0z1_exact)


def _big_state(n, prec):
    try:
        return StateVector(n, prec)
    except L.NqpuError as e:
        if e.code != 8:  # only "out of device memory" is a reason to skip
            raise
        pytest.skip(f"{n}-qubit {prec} state does not fit the free device memory: {e}")


@pytest.mark.parametrize("n,prec", [(34, "c64"), (33, "c128")])
def test_memory_ceiling_state(z0z1_exact, n, prec):
    """128 GiB states: the full-size circuit, checked through the exact
    light-cone value and the norm."""
    sv = _big_state(n, prec)
    try:
        st = sv.run(qft_trotter(n, 20))
        assert st["passes"] <= 60
        check_z0z1(sv, n, prec, z0z1_exact)
    finally:
        sv.close()
