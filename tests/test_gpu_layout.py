"""Runs from |0...0> and the state's qubit layout.

A tiled complex64 run may leave the state's qubits where its last pass
stored them (no closing permutation passes); the state records that layout.
Every read or write by index first restores the canonical layout with
permutation-only passes, and the Pauli sums, <Z_i>, measurement and collapse
map qubits to their positions instead.  Each accessor must therefore give on a
state left in a moved layout what it gives on the same state after the layout
was restored: amplitudes, probabilities and shots bit for bit, FP64 reductions
to their summation rounding.  NQPU_RUN_FROM_ZERO runs start from |0...0>: the
first register pass generates that state instead of reading it, which must
equal a reset followed by a run bit for bit."""
import ctypes as ct

import numpy as np
import pytest

from nqpu_amd import _lib as L
from nqpu_amd import engine as E
from nqpu_amd.engine import Circuit
from nqpu_amd.workloads import random_u3_brickwork

pytestmark = pytest.mark.gpu

N = 20
IDENT = list(range(N))


def layout(sv):
    arr = (ct.c_int32 * sv.num_qubits)()
    L.check(L.load().nqpu_debug_state_layout(sv.ptr, arr, sv.num_qubits), "layout")
    return list(arr)


@pytest.fixture(scope="module")
def circ():
    return random_u3_brickwork(N, 16, 7)


@pytest.fixture(scope="module")
def native(circ):
    return circ.native()


@pytest.fixture(scope="module")
def expected(circ, oracle):
    return oracle.run(N, circ.gates)


def pair(native, **kw):
    """(a, b): the same run from |0...0> twice; a keeps the moved layout, b
    has had its layout restored by a read."""
    a, b = E.StateVector(N, "c64"), E.StateVector(N, "c64")
    a.run(native, from_zero=True, **kw)
    b.run(native, from_zero=True, **kw)
    assert layout(a) != IDENT, "the plan should leave a moved layout"
    b.amplitudes()
    assert layout(b) == IDENT
    return a, b


def infidelity(x, y):
    return 1.0 - abs(np.vdot(x, y)) ** 2 / (np.vdot(x, x).real * np.vdot(y, y).real)


@pytest.mark.parametrize("compile", [False, True])
def test_run_from_zero_matches_oracle_and_reset_then_run(native, expected, compile):
    kw = dict(compile=compile, tile_qubits=12) if compile else {}
    a = E.StateVector(N, "c64")
    a.set_amplitudes(np.full(1 << N, 0.3 + 0.1j))  # garbage a run from |0...0> must ignore
    st = a.run(native, from_zero=True, **kw)
    b = E.StateVector(N, "c64")
    b.set_amplitudes(np.full(1 << N, 0.3 + 0.1j))
    b.reset()
    st2 = b.run(native, **kw)
    assert st["passes"] == st2["passes"]
    assert layout(a) == layout(b)
    if not compile:  # (this compiled plan's last pass happens to end in the canonical layout)
        assert layout(a) != IDENT
    x, y = a.amplitudes(), b.amplitudes()
    assert np.array_equal(x, y)
    assert infidelity(x, expected) < 1e-6


def test_reads_restore_the_layout(native, expected):
    a, b = pair(native)
    assert np.array_equal(a.probabilities(), b.probabilities())
    assert layout(a) == IDENT
    a, b = pair(native)
    assert np.array_equal(a.sample(500, seed=11), b.sample(500, seed=11))
    a, b = pair(native)
    x = np.random.default_rng(3).standard_normal(1 << N) * (1 + 0j)
    a.set_amplitudes(x)
    b.set_amplitudes(x)
    assert layout(a) == IDENT
    assert np.array_equal(a.amplitudes(), b.amplitudes())


def test_mapped_reductions_keep_the_layout(native):
    a, b = pair(native)
    moved = layout(a)
    rng = np.random.default_rng(5)
    strings = ["".join(rng.choice(list("IXYZ"), N)) for _ in range(24)]
    strings += ["Z" * N, "X" + "I" * (N - 2) + "Y", "I" * 14 + "XYZXY" + "Z"]
    va, vb = a.expectation_pauli_values(strings), b.expectation_pauli_values(strings)
    assert np.max(np.abs(va - vb)) < 1e-12
    terms = [(0.5 + k, s) for k, s in enumerate(strings)]
    assert abs(a.expectation_pauli_sum(terms) - b.expectation_pauli_sum(terms)) < 1e-10
    assert np.max(np.abs(a.expectation_z_all() - b.expectation_z_all())) < 1e-12
    for q in (0, 7, N - 1):
        assert abs(a.expectation_z(q) - b.expectation_z(q)) < 1e-12
    assert abs(a.norm() - b.norm()) < 1e-12
    assert layout(a) == moved  # none of these restored it


def test_measure_and_reset_qubit_in_a_moved_layout(native):
    for q in (0, 9, N - 1):
        a, b = pair(native)
        E.set_seed(41 + q)
        oa, pa = a.measure(q)
        E.set_seed(41 + q)
        ob, pb = b.measure(q)
        assert oa == ob and abs(pa - pb) < 1e-12
        assert np.max(np.abs(a.amplitudes() - b.amplitudes())) < 1e-6
    a, b = pair(native)
    a.reset_qubit(4)
    b.reset_qubit(4)
    assert np.max(np.abs(a.amplitudes() - b.amplitudes())) < 1e-6


def test_overlap_copy_and_further_runs(native):
    a, b = pair(native)
    a2 = E.StateVector(N, "c64")
    a2.run(native, from_zero=True)
    # same layout on both sides: no restore needed
    assert abs(a.overlap(a2) - b.overlap(b)) < 1e-12
    assert layout(a) != IDENT and layout(a2) != IDENT
    # different layouts: both restored
    assert abs(a.overlap(b) - b.overlap(b)) < 1e-12
    assert layout(a) == IDENT
    a, b = pair(native)
    c = E.StateVector(N, "c64")
    c.copy_from(a)
    assert layout(c) == layout(a)
    assert np.array_equal(c.amplitudes(), b.amplitudes())
    # a run that does not start from |0...0> continues from the state
    a, b = pair(native)
    a.run(native)
    b.run(native)
    assert np.array_equal(a.amplitudes(), b.amplitudes())
    a, b = pair(native)
    h = Circuit(N).h(3).cnot(3, 11)
    a.run(h, fusion="none")
    b.run(h, fusion="none")
    assert np.array_equal(a.amplitudes(), b.amplitudes())
    a, b = pair(native)
    a.initialize(np.ones(1 << N))
    assert layout(a) == IDENT
    assert np.allclose(a.amplitudes(), 2.0 ** (-N / 2))


def test_graph_replays_and_layout(native, circ):
    ref = E.StateVector(N, "c64")
    ref.run(native, from_zero=True)
    want = ref.amplitudes()
    sv = E.StateVector(N, "c64")
    with E.RunGraph(sv, native, from_zero=True) as g:
        for _ in range(3):
            g.launch()
            assert layout(sv) != IDENT
            assert np.array_equal(sv.amplitudes(), want)
    # a graph that continues from the state restores the layout first
    sv.run(native, from_zero=True)
    ref.run(native, from_zero=True)
    ref.run(native)
    with E.RunGraph(sv, native, from_zero=False) as g:
        g.launch()
        assert np.array_equal(sv.amplitudes(), ref.amplitudes())


@pytest.mark.parametrize("kw", [dict(lds_tiles=True), dict(fusion="none"), dict(fusion="reference")])
def test_from_zero_on_other_paths(native, kw):
    a, b = E.StateVector(N, "c64"), E.StateVector(N, "c64")
    a.set_amplitudes(np.full(1 << N, 0.5))
    b.set_amplitudes(np.full(1 << N, 0.5))
    a.run(native, from_zero=True, **kw)
    b.reset()
    b.run(native, **kw)
    assert np.array_equal(a.amplitudes(), b.amplitudes())


@pytest.mark.parametrize("compile", [False, True])
def test_complex128_from_zero(native, expected, compile):
    kw = dict(compile=True, tile_qubits=11) if compile else {}
    a, b = E.StateVector(N, "c128"), E.StateVector(N, "c128")
    a.set_amplitudes(np.full(1 << N, 0.25j))
    a.run(native, from_zero=True, **kw)
    b.run(native, **kw)
    x = a.amplitudes()
    assert np.array_equal(x, b.amplitudes())
    assert np.max(np.abs(x - expected)) < 1e-12


def test_26_qubit_compiled_run_in_13_qubit_tiles(oracle):
    """Compiled complex64 runs of 26 qubits and more use 13-qubit tiles by
    default; the layout such a run leaves is restored by 13-qubit closing
    passes (amplitudes against the oracle) or mapped (Pauli values equal to
    the restored state's)."""
    n = 26
    c = random_u3_brickwork(n, 8, 31)
    exp = oracle.run(n, c.gates)
    nc = c.native()
    a, b = E.StateVector(n, "c64"), E.StateVector(n, "c64")
    st = a.run(nc, compile=True, from_zero=True)
    b.run(nc, compile=True, from_zero=True)
    assert st["passes"] > 0
    moved = layout(a)
    assert moved != list(range(n))
    zz = [(0, 3 << i) for i in range(n - 1)] + ["X" + "I" * (n - 2) + "Z"]
    va = a.expectation_pauli_values(zz)
    assert layout(a) == moved
    x = b.amplitudes()  # restores b's layout
    assert layout(b) == list(range(n))
    assert infidelity(x, exp) < 1e-6
    vb = b.expectation_pauli_values(zz)
    assert np.max(np.abs(va - vb)) < 1e-12
