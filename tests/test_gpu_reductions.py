"""GPU reductions and state mutations through the C ABI against the FP64
oracle and numpy.

* Overlap <a|b> for all four precision pairings against np.vdot (and
  |<a|b>|^2 against orc_fidelity), device copies bit for bit.
* Grouped Pauli sums (one pass per distinct X mask, or one tile pass for
  strings whose X bits fit a tile) term by term against
  orc_expectation_pauli: complex128 within 1e-12, complex64 within 1e-6.
* Pauli strings in both qubit orders and the Estimator's MSB terms against
  orc_pauli_expectation_msb (primitives.rs:683-728), repeated qubits too.
* reset_qubit (lib.rs:1705-1730) and initialize_from_amplitudes
  (lib.rs:1918-1935) against the oracle, with the rejection cases.
* Probabilities of a state above 2^24 amplitudes, sampling at 26 qubits and
  on tiny states index for index against the oracle's C ABI sampler.
* A measurement whose drawn outcome has no measured mass (complex64 drift,
  here forced through an unnormalised state) collapses onto the other one.
"""
import ctypes as ct

import numpy as np
import pytest

from nqpu_amd import _lib as L
from nqpu_amd import engine as E
from nqpu_amd.engine import Circuit, StateVector
from nqpu_amd.workloads import random_u3_brickwork

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if E.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X")


@pytest.fixture(scope="module")
def pair(oracle):
    """Two different 16-qubit states (oracle amplitudes)."""
    a = random_u3_brickwork(16, 6, 31)
    b = random_u3_brickwork(16, 6, 32)
    return a, b, oracle.run(16, a.gates), oracle.run(16, b.gates)


@pytest.mark.parametrize("pa,pb", [("c128", "c128"), ("c64", "c64"), ("c64", "c128"), ("c128", "c64")])
def test_overlap_all_precision_pairs(oracle, pair, pa, pb):
    ca, cb, ea, eb = pair
    with StateVector(16, pa) as sa, StateVector(16, pb) as sb:
        sa.run(ca)
        sb.run(cb)
        got = sa.overlap(sb)
        amps_a, amps_b = sa.amplitudes(), sb.amplitudes()
        self_ov = sa.overlap(sa)
    # against numpy on the device's own amplitudes: only summation order differs
    assert abs(got - np.vdot(amps_a, amps_b)) < 1e-12
    assert abs(self_ov - 1.0) < (1e-12 if pa == "c128" else 1e-6)
    tol = 1e-12 if pa == pb == "c128" else 2e-6
    assert abs(got - np.vdot(ea, eb)) < tol
    assert abs(abs(got) ** 2 - oracle.fidelity(ea, eb, 16)) < tol


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_copy_is_exact(pair, prec):
    ca, _, _, _ = pair
    with StateVector(16, prec) as s, StateVector(16, prec) as d:
        s.run(ca)
        d.copy_from(s)
        assert np.array_equal(d.amplitudes(), s.amplitudes())
        assert d.fidelity(s) == pytest.approx(1.0, abs=1e-6)
    with StateVector(16, "c64") as s, StateVector(15, "c64") as d:
        with pytest.raises(L.NqpuError):
            d.copy_from(s)


@pytest.fixture(scope="module")
def c20(oracle):
    c = random_u3_brickwork(20, 12, 77)
    return c, oracle.run(20, c.gates)


def _terms20(rng):
    n = 20
    t = []
    for i in range(n - 1):                      # the TFIM ZZ chain: one pass
        s = ["I"] * n
        s[i] = s[i + 1] = "Z"
        t.append((-1.0, "".join(s)))
    for i in range(0, n, 3):                    # X terms: one pass per mask
        s = ["I"] * n
        s[i] = "X"
        t.append((-0.7, "".join(s)))
    for _ in range(12):                         # random strings
        t.append((float(rng.normal()), "".join("IXYZ"[k] for k in rng.integers(0, 4, n))))
    t.append((0.25, "Y" + "I" * 18 + "Y"))        # shared mask with different signs
    t.append((0.5, "X" + "I" * 18 + "X"))
    t.append((0.125, "Y" + "I" * 18 + "X"))
    return t


@pytest.mark.parametrize("prec,tol", [("c128", 1e-12), ("c64", 1e-6)])
def test_grouped_pauli_sum_matches_oracle_term_by_term(oracle, c20, prec, tol):
    c, exp = c20
    terms = _terms20(np.random.default_rng(3))
    with StateVector(20, prec) as sv:
        sv.run(c)
        per_term = [sv.expectation_pauli(p) for _, p in terms]
        grouped = sv.expectation_pauli_values([p for _, p in terms])
        total = sv.expectation_pauli_sum(terms)
    want = [oracle.expectation_pauli(exp, 20, *E.pauli_masks(p)) for _, p in terms]
    assert np.max(np.abs(np.array(per_term) - want)) < tol
    assert np.max(np.abs(grouped - want)) < tol
    assert abs(total - sum(cf * w for (cf, _), w in zip(terms, want))) < 50 * tol


def test_pauli_strings_both_orders_and_estimator(oracle):
    n = 9
    c = random_u3_brickwork(n, 5, 8)
    exp = oracle.run(n, c.gates)
    rng = np.random.default_rng(4)
    with StateVector(n, "c128") as sv:
        sv.run(c)
        for _ in range(10):
            k = int(rng.integers(1, n + 1))
            s = "".join("IXYZ"[j] for j in rng.integers(0, 4, k))
            lsb = sv.expectation_pauli_string(s, "lsb")
            assert abs(lsb - oracle.expectation_pauli(exp, n, *E.pauli_masks(s))) < 1e-12, s
            msb = sv.expectation_pauli_string(s, "msb")
            term = [(q, p) for q, p in enumerate(s)]
            assert abs(msb - oracle.pauli_expectation_msb(exp, n, term)) < 1e-12, s
        # an observable of several terms, one with a repeated qubit (applied
        # factor by factor against the original index, as the reference does)
        obs = [(0.7, [(0, "X"), (3, "Y")]), (-1.2, [(2, "Z"), (8, "Z")]), (0.4, [(5, "Y"), (5, "Z")]),
               (0.9, [(1, "Y"), (1, "Y")]), (0.3, [(4, "I")])]
        got = sv.estimator_expectation(obs)
        want = sum(cf * oracle.pauli_expectation_msb(exp, n, t) for cf, t in obs)
        assert abs(got - want) < 1e-12
        with pytest.raises(L.NqpuError) as ei:
            sv.expectation_pauli_string("XQ", "lsb")
        assert ei.value.code == 2
        with pytest.raises(L.NqpuError) as ei:
            sv.expectation_pauli_string("X" * (n + 1), "msb")
        assert ei.value.code == 1
        with pytest.raises(L.NqpuError) as ei:
            sv.estimator_expectation([(1.0, [(n, "Z")])])
        assert ei.value.code == 1


@pytest.mark.parametrize("prec,tol", [("c128", 1e-12), ("c64", 2e-6)])
def test_reset_qubit_matches_oracle(oracle, prec, tol):
    n = 14
    c = random_u3_brickwork(n, 6, 41)
    exp = oracle.run(n, c.gates)
    with StateVector(n, prec) as sv:
        sv.run(c)
        for q in (0, 6, 13):
            sv.reset_qubit(q)
            oracle.reset_qubit(exp, n, q)
            assert np.max(np.abs(sv.amplitudes() - exp)) < tol
        assert abs(sv.norm() - 1) < 10 * tol
    # |1> on a qubit has nothing left: zero state, no renormalisation
    with StateVector(3, prec) as sv:
        sv.run(Circuit(3).x(2))
        sv.reset_qubit(2)
        assert np.all(sv.amplitudes() == 0)
    with StateVector(3, prec) as sv:
        with pytest.raises(L.NqpuError) as ei:
            sv.reset_qubit(3)
        assert ei.value.code == 1


@pytest.mark.parametrize("prec,tol", [("c128", 1e-15), ("c64", 1e-7)])
def test_initialize_matches_oracle_and_rejects(oracle, prec, tol):
    n = 12
    rng = np.random.default_rng(6)
    src = rng.normal(size=1 << n) * 3 + 1j * rng.normal(size=1 << n)
    want = oracle.initialize_from_amplitudes(n, src)
    with StateVector(n, prec) as sv:
        assert sv.initialize(src)
        got = sv.amplitudes()
        assert np.max(np.abs(got - want)) < tol
        assert sv.initialize(src[:100]) is False            # wrong length
        assert sv.initialize(np.zeros(1 << n)) is False     # zero norm
        tiny = np.zeros(1 << n, complex)
        tiny[5] = 5e-11
        assert sv.initialize(tiny) is False                 # below 1e-10
        assert np.array_equal(sv.amplitudes(), got)         # rejected calls leave the state
        tiny[5] = 3e-10
        assert sv.initialize(tiny)
        assert abs(sv.amplitudes()[5] - 1) < tol


def test_initialize_above_one_slice(oracle):
    """25 qubits: the upload is split into 2^24-amplitude slices (norm pass,
    then import pass); probabilities come from the device in slices too."""
    n = 25
    rng = np.random.default_rng(12)
    src = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex128)
    want = src / np.linalg.norm(src)
    with StateVector(n, "c64") as sv:
        assert sv.initialize(src)
        got = sv.amplitudes()
        assert np.max(np.abs(got - want)) < 1e-7
        p = sv.probabilities()
        assert np.max(np.abs(p - np.abs(got) ** 2)) < 1e-15
        assert abs(p.sum() - 1) < 1e-5


def test_sampling_26_qubits_and_tiny_states(oracle):
    c = random_u3_brickwork(26, 6, 9)
    with StateVector(26, "c64") as sv:
        sv.run(c)
        got = sv.sample(4096, seed=555)
        amps = sv.amplitudes()
    want = oracle.sample_cabi(np.ascontiguousarray(amps), 26, 4096, seed=555)
    assert np.array_equal(got, want)
    for n in (1, 3, 5, 7, 11, 17):
        cc = random_u3_brickwork(n, 3, n)
        for prec in ("c128", "c64"):
            with StateVector(n, prec) as sv:
                sv.run(cc)
                got = sv.sample(3000, seed=n)
                amps = sv.amplitudes()
            assert np.array_equal(got, oracle.sample_cabi(np.ascontiguousarray(amps), n, 3000, seed=n)), (n, prec)


def test_measure_never_collapses_onto_a_massless_outcome(oracle):
    """|0> with amplitude 1/2 (mass 1/4): a draw r >= 1/4 picks outcome 1,
    whose measured mass is zero; the collapse keeps outcome 0 instead and
    renormalises by its measured mass."""
    hits = 0
    for seed in range(1, 40):
        with StateVector(2, "c64") as sv:
            a = np.zeros(4, complex)
            a[0] = 0.5
            sv.set_amplitudes(a)
            E.set_seed(seed)
            o, p = sv.measure(1)
            out = sv.amplitudes()
        oracle.rng_seed(seed)
        r = oracle.rng_next()
        assert o == 0
        assert abs(out[0] - 1) < 1e-7 and np.all(out[1:] == 0)
        # the reported probability keeps the reference's formula
        assert abs(p - 0.25) < 1e-7
        hits += r >= 0.25
    assert hits > 10  # the massless outcome was drawn, and avoided, many times


@pytest.mark.parametrize("n,prec,tol", [(6, "c128", 1e-12), (11, "c64", 1e-6), (14, "c128", 1e-12),
                                        (20, "c64", 1e-6), (20, "c128", 1e-12)])
def test_strings_with_different_x_masks_in_one_tile(oracle, n, prec, tol):
    """Strings whose X/Y bits fit one tile with the lowest qubits (every X_i of
    a chain, neighbouring XX / XY / YZ pairs, Z factors inside and outside the
    tile) share tile passes; each value must still be the oracle's, term by
    term, and the no-allocation sum must agree."""
    c = random_u3_brickwork(n, 8, 900 + n)
    exp = oracle.run(n, c.gates)
    terms = []
    for i in range(n):
        s = ["I"] * n
        s[i] = "X"
        terms.append((-0.7, "".join(s)))
    for i in range(n - 1):
        for a, b in (("X", "X"), ("X", "Y"), ("Y", "Z"), ("Y", "Y")):
            s = ["I"] * n
            s[i], s[i + 1] = a, b
            if i + 3 < n:
                s[n - 1 - (i % 2)] = "Z" if s[n - 1 - (i % 2)] == "I" else s[n - 1 - (i % 2)]
            terms.append((0.3, "".join(s)))
    with StateVector(n, prec) as sv:
        sv.run(c)
        got = sv.expectation_pauli_values([p for _, p in terms])
        total = sv.expectation_pauli_sum(terms)
        xs = np.array([E.pauli_masks(p)[0] for _, p in terms], dtype=np.uint64)
        zs = np.array([E.pauli_masks(p)[1] for _, p in terms], dtype=np.uint64)
        cs = np.array([cf for cf, _ in terms], dtype=np.float64)
        v = ct.c_double()
        lib = L.load()
        L.check(lib.nqpu_expectation_pauli_sum_noalloc(sv.ptr, xs.ctypes.data, zs.ctypes.data, cs.ctypes.data,
                                                       len(terms), ct.byref(v)), "noalloc")
    want = np.array([oracle.expectation_pauli(exp, n, *E.pauli_masks(p)) for _, p in terms])
    assert np.max(np.abs(got - want)) < tol
    ref_total = float(np.dot(cs, want))
    assert abs(total - ref_total) < 50 * tol
    assert abs(v.value - ref_total) < 50 * tol


@pytest.mark.parametrize("prec,tol", [("c128", 1e-12), ("c64", 1e-6)])
def test_group_with_more_accumulators_than_one_pass(oracle, prec, tol):
    """32 strings sharing one X mask (X or Y on qubit 0) with Z factors on every
    subset of qubits 16..19: at 20 qubits those are the pass's iteration bits,
    so the strings need 32 distinct (iteration bits, real/imaginary)
    accumulators, more than one pass holds, and the group is split over
    passes.  Each value must be the oracle's; diagonal strings on the same
    qubits (16 accumulators) ride along."""
    n = 20
    c = random_u3_brickwork(n, 8, 77)
    exp = oracle.run(n, c.gates)
    terms = []
    for first in ("X", "Y"):
        for sub in range(16):
            s = ["I"] * n
            s[0] = first
            for b in range(4):
                if (sub >> b) & 1:
                    s[16 + b] = "Z"
            terms.append("".join(s))
    for sub in range(1, 16):
        s = ["I"] * n
        s[1] = "Z"
        for b in range(4):
            if (sub >> b) & 1:
                s[16 + b] = "Z"
        terms.append("".join(s))
    with StateVector(n, prec) as sv:
        sv.run(c)
        got = sv.expectation_pauli_values(terms)
    want = np.array([oracle.expectation_pauli(exp, n, *E.pauli_masks(p)) for p in terms])
    assert np.max(np.abs(got - want)) < tol
