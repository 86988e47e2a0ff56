"""GPU parity: the HIP path through the C ABI against the FP64 oracle and the
golden vectors.  Tolerances: complex128 amplitudes within 1e-12 of the
oracle (and <Z_i> within 1e-12); complex64 amplitudes within 2e-6 per
small circuit and infidelity <= 1e-5 on the large workloads.  Sampling is
compared index-for-index with the oracle's restatement of the C ABI's seeded
stream; measurement outcome, probability and collapsed state likewise."""
import math

import numpy as np
import pytest

from nqpu_amd import engine as E
from nqpu_amd.engine import Circuit, Gate, StateVector
from nqpu_amd.workloads import qft, random_u3_brickwork, tfim_trotter
from tests.conftest import decode_gates

pytestmark = pytest.mark.gpu

FUSIONS = ["none", "reference", "tiled"]
TOL = {"c128": 1e-12, "c64": 2e-6}


@pytest.fixture(scope="module", autouse=True)
def device():
    n = E.device_count()
    if n == 0:
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X")
    return n


def run_gpu(n, gates, prec="c128", fusion="tiled", semantics="standard", psi0=None, **kw):
    c = Circuit(n)
    c.extend(gates)
    with StateVector(n, prec) as sv:
        if psi0 is not None:
            sv.set_amplitudes(psi0)
        sv.run(c, fusion=fusion, semantics=semantics, **kw)
        return sv.amplitudes()


def infidelity(a, b):
    return 1.0 - abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("fusion", FUSIONS)
def test_every_kind_against_golden(golden, prec, fusion):
    n = int(golden["single_n"])
    gates = decode_gates(golden, "single")
    for i, g in enumerate(gates):
        out = run_gpu(n, [g], prec, fusion, psi0=golden["single_in"][i])
        err = np.max(np.abs(out - golden["single_out"][i]))
        assert err < TOL[prec], (g.kind, prec, fusion, err)


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("fusion", FUSIONS)
@pytest.mark.parametrize("idx", [0, 1, 2])
def test_random_circuits_against_golden(golden, prec, fusion, idx):
    n = int(golden[f"circ{idx}_n"])
    gates = decode_gates(golden, f"circ{idx}")
    out = run_gpu(n, gates, prec, fusion)
    tol = 1e-11 if prec == "c128" else 1e-5
    assert np.max(np.abs(out - golden[f"circ{idx}_out"])) < tol


@pytest.mark.parametrize("fusion", FUSIONS)
def test_reference_semantics_against_oracle(oracle, golden, fusion):
    """The dispatcher's quirks (CRx as CRz, CRz parallel branch, dropped
    Rxx/Ryy/Rzz/CSWAP/CU/Custom) reproduced in reference mode."""
    for idx in (0, 1, 2):
        n = int(golden[f"circ{idx}_n"])
        gates = decode_gates(golden, f"circ{idx}")
        exp = oracle.run(n, gates, quirk=True)
        out = run_gpu(n, gates, "c128", fusion, semantics="reference")
        assert np.max(np.abs(out - exp)) < 1e-11


@pytest.mark.parametrize("tile,low", [(5, 2), (8, 3), (10, 5), (12, 5), (13, 6)])
def test_tile_shapes(oracle, tile, low):
    c = random_u3_brickwork(16, 6, 7)
    c.extend(qft(16).gates[:60])
    c.extend(tfim_trotter(16, 2).gates)
    exp = oracle.run(16, c.gates)
    out = run_gpu(16, c.gates, "c128", "tiled", tile_qubits=tile, low_qubits=low)
    assert np.max(np.abs(out - exp)) < 1e-12


@pytest.mark.parametrize("tile,low", [(11, 3), (12, 5), (13, 6)])
@pytest.mark.parametrize("lds", [False, True])
def test_complex64_tile_kernels(oracle, tile, low, lds):
    """Register-resident (default) and LDS tile kernels in complex64."""
    from tests.test_planner_emulation import random_circuit
    c = random_circuit(15, 300, 11)
    c.extend(qft(15).gates)
    exp = oracle.run(15, c.gates)
    with StateVector(15, "c64") as sv:
        sv.run(c, tile_qubits=tile, low_qubits=low, lds_tiles=lds)
        out = sv.amplitudes()
    assert np.max(np.abs(out - exp)) < 2e-5
    assert infidelity(out, exp) < 1e-6


@pytest.mark.parametrize("tile", [12, 13])
@pytest.mark.parametrize("name", ["mixed16", "qft_trotter15", "brick14"])
def test_compiled_passes_match_oracle(oracle, name, tile):
    """complex64 register tiles run as straight-line kernels compiled at run
    time (compile=True): the same register programs as the interpreting
    kernel, every op kind, controls and diagonal factors outside the tile.
    13-qubit tiles hold 32 amplitudes per thread: matrix-core blocks run over
    four groups of 8 registers."""
    from tests.test_planner_emulation import CIRCUITS, random_state
    c = CIRCUITS[name]()
    n = c.num_qubits
    a0 = random_state(n, 5)
    exp = a0.copy()
    oracle.apply_circuit(exp, n, c.gates)
    outs = []
    for compiled in (True, False):
        with StateVector(n, "c64") as sv:
            sv.set_amplitudes(a0)
            st = sv.run(c, compile=compiled, tile_qubits=tile)
            out = sv.amplitudes()
        assert st["passes"] > 0
        if tile == 13 and name != "mixed16":
            assert st["mfma_blocks"] > 0
        assert np.max(np.abs(out - exp)) < 2e-5, compiled
        assert infidelity(out, exp) < 1e-6, compiled
        outs.append(out)
    assert np.array_equal(outs[0], outs[1])  # the same program, compiled or interpreted


@pytest.mark.parametrize("n,tile,low,count", [(17, 12, 5, 200), (20, 11, 3, 90), (20, 13, 6, 150)])
def test_quadratic_phase_runs_match_oracle(oracle, n, tile, low, count):
    """Runs of diagonal gates merged into one quadratic phase per run, with
    its angle tables over thread slots, register slots and chunks of qubits
    outside the tile (hardware sin/cos in revolutions): interpreted and, for
    12-qubit tiles, compiled; the same circuit with the gates applied one by
    one agrees too."""
    from tests.test_planner_emulation import diagonal_heavy_circuit, random_state
    c = diagonal_heavy_circuit(n, count, 40 + n + tile)
    a0 = random_state(n, 3)
    exp = a0.copy()
    oracle.apply_circuit(exp, n, c.gates)
    for compiled in ((False, True) if tile == 12 else (False,)):
        for qphase in (True, False):
            with StateVector(n, "c64") as sv:
                sv.set_amplitudes(a0)
                st = sv.run(c, tile_qubits=tile, low_qubits=low, compile=compiled, qphase=qphase)
                out = sv.amplitudes()
            assert (st["qphase_gates"] > 0) == qphase
            assert np.max(np.abs(out - exp)) < 2e-5, (compiled, qphase)
            assert infidelity(out, exp) < 1e-6, (compiled, qphase)


@pytest.mark.parametrize("n,tile,depth", [(20, 12, 16), (18, 11, 12), (19, 13, 14)])
def test_moving_low_positions_match_oracle(oracle, n, tile, depth):
    """Passes that hand the physical low positions to the next window's qubits
    (permuted stores, closing permutation-only passes): interpreted and, for
    12-qubit tiles, compiled, against the oracle and the fixed-low plan."""
    from tests.test_planner_emulation import random_state
    c = random_u3_brickwork(n, depth, 50 + n)
    a0 = random_state(n, 4)
    exp = a0.copy()
    oracle.apply_circuit(exp, n, c.gates)
    for compiled in ((False, True) if tile == 12 else (False,)):
        for fixed in (False, True):
            with StateVector(n, "c64") as sv:
                sv.set_amplitudes(a0)
                sv.run(c, tile_qubits=tile, low_qubits=5, compile=compiled, fixed_low=fixed)
                out = sv.amplitudes()
            assert np.max(np.abs(out - exp)) < 2e-5, (compiled, fixed)
            assert infidelity(out, exp) < 1e-6, (compiled, fixed)


def test_permuted_store_across_wavefronts_is_ordered():
    """A 22-qubit brickwork whose closing permutation pass moves the positions
    held by the wave slots without any LDS remap: wavefronts store amplitudes
    other wavefronts load, so the pass must order all loads first (without
    that, a few hundred amplitudes differed from run to run).  Repeated runs,
    interpreted and compiled, equal the fixed-low plan's result."""
    c = random_u3_brickwork(22, 12, 7)
    nc = c.native()
    assert nc.plan("c64", tile_qubits=12)["passes"] < nc.plan("c64", tile_qubits=12, fixed_low=True)["passes"]
    with StateVector(22, "c64") as sv:
        sv.run(nc, tile_qubits=12, fixed_low=True)
        ref = sv.amplitudes()
    for compiled in (False, True):
        for _ in range(3):
            with StateVector(22, "c64") as sv:
                sv.run(nc, tile_qubits=12, compile=compiled)
                out = sv.amplitudes()
            assert np.max(np.abs(out - ref)) < 1e-6, compiled


def test_compiled_passes_c1_bitwise_equal_interpreted(c1):
    """The compiled and interpreted kernels execute the same operations in the
    same order on the same registers: identical complex64 results."""
    c, _ = c1
    outs = []
    for compiled in (False, True, True):  # the second compiled run reuses the cached kernels
        with StateVector(20, "c64") as sv:
            sv.run(c, compile=compiled, tile_qubits=12)  # the same plan for both kernels
            outs.append(sv.amplitudes())
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[1], outs[2])


@pytest.fixture(scope="module")
def c1(oracle):
    """20 qubits, 20 U3 + CNOT-brickwork layers, seed 101; the oracle state."""
    c = random_u3_brickwork(20, 20, 101)
    return c, oracle.run(20, c.gates)


@pytest.mark.parametrize("fusion", FUSIONS)
def test_c1_complex128_matches_oracle(oracle, c1, fusion):
    c, exp = c1
    with StateVector(20, "c128") as sv:
        sv.run(c, fusion=fusion)
        out = sv.amplitudes()
        z = sv.expectation_z_all()
    assert np.max(np.abs(out - exp)) < 1e-12
    assert infidelity(out, exp) < 1e-12
    zexp = np.array([oracle.expectation_z(exp, 20, q) for q in range(20)])
    assert np.max(np.abs(z - zexp)) < 1e-12


def test_c1_complex64_infidelity(c1):
    c, exp = c1
    out = run_gpu(20, c.gates, "c64", "tiled")
    assert infidelity(out, exp) < 1e-5


def test_reductions_against_oracle(oracle, c1):
    c, exp = c1
    rng = np.random.default_rng(11)
    with StateVector(20, "c128") as sv:
        sv.run(c)
        assert abs(sv.norm() - oracle.norm(exp, 20)) < 1e-12
        p = sv.probabilities()
        assert np.max(np.abs(p - oracle.probabilities(exp, 20))) < 1e-15
        for q in (0, 7, 19):
            assert abs(sv.expectation_z(q) - oracle.expectation_z(exp, 20, q)) < 1e-12
        for _ in range(12):
            s = "".join("IXYZ"[k] for k in rng.integers(0, 4, 20))
            x, z = E.pauli_masks(s)
            assert abs(sv.expectation_pauli(s) - oracle.expectation_pauli(exp, 20, x, z)) < 1e-12, s
        terms = [(-1.0, "ZZ" + "I" * 18), (-0.7, "X" + "I" * 19), (0.3, "I" * 19 + "Y")]
        ref = sum(cf * oracle.expectation_pauli(exp, 20, *E.pauli_masks(s)) for cf, s in terms)
        assert abs(sv.expectation_pauli_sum(terms) - ref) < 1e-12


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_sampling_matches_oracle_stream(oracle, c1, prec):
    c, exp = c1
    with StateVector(20, prec) as sv:
        sv.run(c)
        got = sv.sample(20000, seed=2024)
        amps = sv.amplitudes()
    want = oracle.sample_cabi(np.ascontiguousarray(amps), 20, 20000, seed=2024)
    assert np.array_equal(got, want)


def test_sampling_edge_cases(oracle):
    with StateVector(3, "c128") as sv:
        sv.run(Circuit(3).x(1))
        assert set(sv.sample(100, seed=1).tolist()) == {2}          # lib.rs:2585 basis state
        assert len(sv.sample(0)) == 0
    with StateVector(1, "c64") as sv:
        sv.run(Circuit(1).h(0))
        s = sv.sample(10000, seed=9)                               # lib.rs:2568
        assert abs((s == 0).mean() - 0.5) < 0.05


def test_measure_matches_oracle(oracle):
    c = random_u3_brickwork(10, 4, 3)
    exp = oracle.run(10, c.gates)
    with StateVector(10, "c128") as sv:
        sv.run(c)
        E.set_seed(77)
        oracle.rng_seed(77)
        for q in (0, 4, 9):
            o, p = sv.measure(q)
            oo, pp = oracle.measure_qubit(exp, 10, q)
            assert o == oo and abs(p - pp) < 1e-12
        assert np.max(np.abs(sv.amplitudes() - exp)) < 1e-12


def test_large_state_c64_vs_c128_and_fusion_modes():
    """28 qubits, depth 24: tiled c64 against tiled c128 and the per-gate
    c128 path; norm preserved."""
    c = random_u3_brickwork(28, 24, 102)
    with StateVector(28, "c128") as a:
        a.run(c, fusion="tiled")
        za = a.expectation_z_all()
        amp_a = a.amplitudes()
    with StateVector(28, "c128") as b:
        b.run(c, fusion="none")
        amp_b = b.amplitudes()
    assert np.max(np.abs(amp_a - amp_b)) < 1e-12
    del amp_b
    with StateVector(28, "c64") as d:
        d.run(c, fusion="tiled")
        assert abs(d.norm() - 1) < 1e-4
        zd = d.expectation_z_all()
        amp_d = d.amplitudes()
    assert infidelity(amp_d, amp_a) < 1e-5
    assert np.max(np.abs(zd - za)) < 1e-5


@pytest.fixture(scope="module")
def c2_26(oracle):
    """The single-large-state family (U3 + CNOT brickwork, depth 40, seed
    102) at 26 qubits, where the FP64 oracle finishes in seconds."""
    c = random_u3_brickwork(26, 40, 102)
    return c, oracle.run(26, c.gates)


def _zz(oracle, a, n):
    return np.array([oracle.expectation_pauli(a, n, 0, (1 << i) | (1 << (i + 1))) for i in range(n - 1)])


def test_c2_family_26_qubits_complex128_matches_oracle(oracle, c2_26):
    c, exp = c2_26
    with StateVector(26, "c128") as sv:
        sv.run(c)
        out = sv.amplitudes()
        zz = [sv.expectation_pauli(xmask=0, zmask=(1 << i) | (1 << (i + 1))) for i in range(25)]
    assert np.max(np.abs(out - exp)) < 1e-12
    assert infidelity(out, exp) < 1e-12
    assert np.max(np.abs(np.array(zz) - _zz(oracle, exp, 26))) < 1e-12


@pytest.mark.parametrize("compile_", [False, True])
def test_c2_family_26_qubits_complex64_matches_oracle(oracle, c2_26, compile_):
    c, exp = c2_26
    with StateVector(26, "c64") as sv:
        sv.run(c, compile=compile_)
        out = sv.amplitudes()
        zz = [sv.expectation_pauli(xmask=0, zmask=(1 << i) | (1 << (i + 1))) for i in range(25)]
    assert infidelity(out, exp) < 1e-5
    assert np.max(np.abs(np.array(zz) - _zz(oracle, exp, 26))) < 1e-5


def test_async_runs_queue_and_match_sync_runs(oracle):
    """Runs queued with async_=True (the host plans the next while the device
    works) give the synchronous result; reads are ordered after them and
    pass_ms covers every queued run."""
    c = random_u3_brickwork(18, 10, 5)
    exp = oracle.run(18, c.gates)
    nc = c.native()
    with StateVector(18, "c128") as sv:
        for _ in range(3):
            sv.reset()
            sv.run(nc, async_=True, time_passes=True)
        st = sv.synchronize()
        out = sv.amplitudes()
        per_run = sv.run(nc, time_passes=True)["timed_passes"]
    assert np.max(np.abs(out - exp)) < 1e-12
    assert st["timed_passes"] == 3 * per_run > 0
    with StateVector(18, "c64") as sv:
        sv.run(nc, async_=True, compile=True)
        z = sv.expectation_z_all()  # ordered after the queued run
    zexp = np.array([oracle.expectation_z(exp, 18, q) for q in range(18)])
    assert np.max(np.abs(z - zexp)) < 1e-5


@pytest.mark.parametrize("tile", [11, 12])
def test_compiled_complex128_passes_match_oracle(oracle, c1, tile):
    """complex128 register tile passes as compiled straight-line kernels
    (which move bits with register/lane exchanges; the interpreting kernel
    uses LDS remaps) against the oracle to 1e-12, and against the
    interpreting kernel's result."""
    c, exp = c1
    with StateVector(20, "c128") as sv:
        st = sv.run(c, tile_qubits=tile, compile=True)
        out = sv.amplitudes()
        sv.reset()
        sv.run(c, tile_qubits=tile, compile=False)
        interp = sv.amplitudes()
    assert st["exchanges"] > 0
    assert np.max(np.abs(out - exp)) < 1e-12
    assert np.max(np.abs(out - interp)) < 1e-13
