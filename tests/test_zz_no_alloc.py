"""nqpu_expectation_pauli_sum_noalloc: the Pauli-sum entry point that
allocates nothing (its device buffers are a fixed array of the library),
called through the C ABI and checked against nqpu_expectation_pauli_sum on
the same state: two calls in a row, the second with a smaller state and
fewer terms; FP64 sums in the same order, so complex128 results agree to
1e-12 (in fact bit for bit).  More than 1024 terms are refused."""
import ctypes as ct

import numpy as np
import pytest

from nqpu_amd import _lib as L
from nqpu_amd import engine as E
from nqpu_amd.engine import StateVector
from nqpu_amd.workloads import random_u3_brickwork

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if E.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X")


def _terms(n, count, seed):
    rng = np.random.default_rng(seed)
    xs, zs, cs = [], [], []
    for k in range(count):
        if k < n - 1:   # diagonal ZZ chain first: one shared pass
            x, z = 0, 3 << k
        else:
            s = "".join("IXYZ"[j] for j in rng.integers(0, 4, n))
            x, z = E.pauli_masks(s)
            if k % 5 == 0:  # repeat a mask so groups hold several terms
                x = xs[-1]
        xs.append(x)
        zs.append(z)
        cs.append(float(rng.normal()))
    return (np.array(xs, dtype=np.uint64), np.array(zs, dtype=np.uint64), np.array(cs, dtype=np.float64))


def _both(sv, xs, zs, cs):
    lib = L.load()
    a, b = ct.c_double(), ct.c_double()
    L.check(lib.nqpu_expectation_pauli_sum(sv.ptr, xs.ctypes.data, zs.ctypes.data, cs.ctypes.data, len(xs),
                                           ct.byref(a)), "pauli_sum")
    L.check(lib.nqpu_expectation_pauli_sum_noalloc(sv.ptr, xs.ctypes.data, zs.ctypes.data, cs.ctypes.data,
                                                   len(xs), ct.byref(b)), "pauli_sum_noalloc")
    return a.value, b.value


@pytest.mark.parametrize("prec,tol", [("c128", 1e-12), ("c64", 1e-12)])
def test_noalloc_matches_pauli_sum_twice_shrinking(prec, tol):
    with StateVector(20, prec) as big:
        big.run(random_u3_brickwork(20, 8, 5))
        a, b = _both(big, *_terms(20, 300, 1))
    assert abs(a - b) <= tol * max(1.0, abs(a)), (a, b)
    with StateVector(11, prec) as small:
        small.run(random_u3_brickwork(11, 6, 6))
        c, d = _both(small, *_terms(11, 7, 2))
    assert abs(c - d) <= tol * max(1.0, abs(c)), (c, d)
    assert a != c  # the second call did not return the first call's leftovers


def test_noalloc_limits():
    lib = L.load()
    with StateVector(8, "c64") as sv:
        xs, zs, cs = _terms(8, 1025, 3)
        out = ct.c_double()
        code = lib.nqpu_expectation_pauli_sum_noalloc(sv.ptr, xs.ctypes.data, zs.ctypes.data, cs.ctypes.data,
                                                      1025, ct.byref(out))
        assert code == 2
        assert b"1024" in lib.nqpu_last_error_detail()
        code = lib.nqpu_expectation_pauli_sum_noalloc(sv.ptr, xs.ctypes.data, zs.ctypes.data, cs.ctypes.data,
                                                      0, ct.byref(out))
        assert code == 0 and out.value == 0.0
        bad = np.array([1 << 8], dtype=np.uint64)
        code = lib.nqpu_expectation_pauli_sum_noalloc(sv.ptr, bad.ctypes.data, bad.ctypes.data, cs.ctypes.data,
                                                      1, ct.byref(out))
        assert code == 1
