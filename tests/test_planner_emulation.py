"""The tiled execution plan on the CPU: nqpu_debug_emulate interprets the very
plan and per-pass register programs nqpu_run launches (tile partition, slot
remaps, relabelled SWAPs, control/diagonal handling) in FP64, and the result
must equal the oracle's gate-by-gate state to 1e-12.  No GPU needed."""
import numpy as np
import pytest

from nqpu_amd.engine import Circuit, Gate
from nqpu_amd.workloads import qft, random_u3_brickwork, tfim_trotter
from tests.golden import make_golden as mg


def random_circuit(n, count, seed):
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for _ in range(count):
        k, t, cq, p, cust = mg.random_gate(rng, n)
        c.gates.append(Gate(k, tuple(t), tuple(cq), tuple(p), cust))
    return c


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return a / np.linalg.norm(a)


CIRCUITS = {
    "mixed13": lambda: random_circuit(13, 300, 1),
    "mixed14": lambda: random_circuit(14, 200, 2),
    "brick14": lambda: random_u3_brickwork(14, 12, 3),
    "qft_trotter13": lambda: tfim_trotter(13, 3, circuit=qft(13)),
    # 4-5 qubits outside every tile: controls and diagonal factors there are
    # resolved from the tile's base index
    "mixed16": lambda: random_circuit(16, 160, 9),
    "qft_trotter15": lambda: tfim_trotter(15, 4, circuit=qft(15)),
}


@pytest.mark.parametrize("name", sorted(CIRCUITS))
@pytest.mark.parametrize("tile,low", [(11, 3), (12, 5), (13, 6), (13, 5), (12, 0)])
@pytest.mark.parametrize("lds", [False, True])
def test_emulated_tiles_match_oracle(oracle, name, tile, low, lds):
    c = CIRCUITS[name]()
    n = c.num_qubits
    a0 = random_state(n, 7)
    exp = a0.copy()
    oracle.apply_circuit(exp, n, c.gates)
    got = c.native().emulate(a0, "c64", tile_qubits=tile, low_qubits=low, lds_tiles=lds)
    assert np.max(np.abs(got - exp)) < 1e-12


@pytest.mark.parametrize("name", ["mixed16", "qft_trotter15"])
def test_emulated_c128_tiles_match_oracle(oracle, name):
    """complex128 runs the LDS tile kernel with its own op encoding."""
    c = CIRCUITS[name]()
    n = c.num_qubits
    a0 = random_state(n, 11)
    exp = a0.copy()
    oracle.apply_circuit(exp, n, c.gates)
    got = c.native().emulate(a0, "c128", tile_qubits=11, low_qubits=3)
    assert np.max(np.abs(got - exp)) < 1e-12


@pytest.mark.parametrize("name", ["mixed13", "brick14", "mixed16", "qft_trotter15"])
@pytest.mark.parametrize("tile", [11, 12])
def test_emulated_c128_compiled_exchanges_match_oracle(oracle, name, tile):
    """complex128 compiled passes move bits with register/lane exchanges
    instead of LDS remaps (and store in the layout the program ends in)."""
    c = CIRCUITS[name]()
    n = c.num_qubits
    a0 = random_state(n, 13)
    exp = a0.copy()
    oracle.apply_circuit(exp, n, c.gates)
    nc = c.native()
    got = nc.emulate(a0, "c128", tile_qubits=tile, compile=True)
    assert np.max(np.abs(got - exp)) < 1e-12
    st = nc.plan("c128", tile_qubits=tile, compile=True)
    assert st["exchanges"] > 0 and st["remaps"] < nc.plan("c128", tile_qubits=tile)["remaps"]


def test_diagonal_gates_do_not_need_tile_bits():
    """A QFT's controlled phases and a Trotter step's RZZ are diagonal, so a
    pass needs only the Hadamard/RX targets as tile bits: a 30-qubit QFT fits
    in ceil((30-12)/7)+1 passes of 12-qubit tiles with 5 fixed low qubits,
    and 20 Trotter steps cost far fewer than one pass per step and tile."""
    st = qft(30).native().plan("c64", tile_qubits=12, low_qubits=5)
    assert st["passes"] <= 6
    st = tfim_trotter(30, 20).native().plan("c64", tile_qubits=12, low_qubits=5)
    assert st["passes"] <= 40


def test_emulated_reference_semantics(oracle):
    c = random_circuit(12, 250, 5)
    a0 = random_state(12, 8)
    exp = a0.copy()
    oracle.apply_circuit(exp, 12, c.gates, quirk=True)
    got = c.native().emulate(a0, "c64", semantics="reference", tile_qubits=12)
    assert np.max(np.abs(got - exp)) < 1e-12


DIAGONAL_KINDS = ["z", "s", "t", "rz", "phase", "cz", "crz", "cr", "rzz", "ccz"]
MIXING_KINDS = ["h", "rx", "u", "cnot", "swap", "cswap"]


def diagonal_heavy_circuit(n, count, seed):
    """Long runs of one- and two-qubit diagonal gates (merged into quadratic
    phase ops), some with three qubits (CCZ: left as single ops), broken up by
    a few mixing gates; many touch only qubits outside the tile."""
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for _ in range(count):
        pool = DIAGONAL_KINDS if rng.random() < 0.8 else MIXING_KINDS
        k, t, cq, p, cust = mg.random_gate(rng, n, pool[rng.integers(len(pool))])
        c.gates.append(Gate(k, tuple(t), tuple(cq), tuple(p), cust))
    return c


@pytest.mark.parametrize("n,tile,low,count", [(14, 12, 5, 300), (17, 12, 5, 200), (17, 13, 6, 200),
                                               (20, 11, 3, 90)])
@pytest.mark.parametrize("qphase", [True, False])
def test_emulated_quadratic_phase_runs(oracle, n, tile, low, count, qphase):
    """Runs of diagonal gates merged into one quadratic phase per run (tables
    over thread slots, register slots and chunks of qubits outside the tile,
    up to 9 such qubits at n = 20) give the oracle's state; with qphase=False
    the same gates run one by one."""
    c = diagonal_heavy_circuit(n, count, 40 + n + tile)
    a0 = random_state(n, 3)
    exp = a0.copy()
    oracle.apply_circuit(exp, n, c.gates)
    nc = c.native()
    got = nc.emulate(a0, "c64", tile_qubits=tile, low_qubits=low, qphase=qphase)
    assert np.max(np.abs(got - exp)) < 1e-12
    merged = nc.plan("c64", tile_qubits=tile, low_qubits=low, qphase=qphase)["qphase_gates"]
    assert (merged > count // 4) if qphase else merged == 0


@pytest.mark.parametrize("n,tile,depth", [(20, 12, 16), (18, 11, 12), (19, 13, 14)])
def test_emulated_moving_low_positions(oracle, n, tile, depth):
    """A brickwork needs far fewer passes when each pass hands the physical
    low positions to qubits of the next window (the planner tracks the
    qubit-to-position map, permutes inside each tile's store and ends with
    passes that only permute); the emulated state equals the oracle's."""
    c = random_u3_brickwork(n, depth, 50 + n)
    nc = c.native()
    moving = nc.plan("c64", tile_qubits=tile, low_qubits=5)["passes"]
    fixed = nc.plan("c64", tile_qubits=tile, low_qubits=5, fixed_low=True)["passes"]
    assert moving < fixed
    a0 = random_state(n, 4)
    exp = a0.copy()
    oracle.apply_circuit(exp, n, c.gates)
    for fl in (False, True):
        got = nc.emulate(a0, "c64", tile_qubits=tile, low_qubits=5, fixed_low=fl)
        assert np.max(np.abs(got - exp)) < 1e-12, fl


@pytest.mark.parametrize("seed", range(6))
def test_emulated_moving_low_random_mixes(oracle, seed):
    """Random gate mixes (every kind, controls and diagonals outside the tile)
    on 15-17 qubits with tiles of 11-13: whichever plan wins (moving or fixed
    low positions), the emulated state equals the oracle's."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(15, 18))
    tile = int(rng.integers(11, 14))
    c = random_circuit(n, int(rng.integers(80, 200)), 100 + seed)
    a0 = random_state(n, seed)
    exp = a0.copy()
    oracle.apply_circuit(exp, n, c.gates)
    got = c.native().emulate(a0, "c64", tile_qubits=tile, low_qubits=5)
    assert np.max(np.abs(got - exp)) < 1e-12


def test_qft_controlled_phases_merge():
    """A 30-qubit QFT's controlled phases merge into one quadratic phase per
    Hadamard, so far fewer matrix-core blocks remain."""
    nc = qft(30).native()
    on, off = nc.plan("c64", compile=True), nc.plan("c64", compile=True, qphase=False)
    assert on["qphase_gates"] >= 400 and off["qphase_gates"] == 0
    assert on["mfma_blocks"] * 4 < off["mfma_blocks"]


@pytest.mark.parametrize("n", [4, 5, 8, 11, 12, 13, 15])
def test_emulated_batch_layout_matches_oracle(oracle, n):
    """The register programs the batched kernel runs: whole-state tiles with
    fewer than 8 thread bits (n <= 12) and 12-qubit HBM passes (n >= 13)."""
    c = random_circuit(n, 150, 20 + n)
    a0 = random_state(n, n)
    exp = a0.copy()
    oracle.apply_circuit(exp, n, c.gates)
    got = c.native().emulate(a0, "c64", batch_layout=True)
    assert np.max(np.abs(got - exp)) < 1e-12


@pytest.mark.parametrize("n", [13, 16])
def test_emulated_batch_layout_quadratic_phases(oracle, n):
    """The batched kernel's HBM passes with matrix-core blocks and quadratic
    phase ops (diagonal runs over qubits inside and outside the tile)."""
    c = diagonal_heavy_circuit(n, 300, 40 + n)
    nc = c.native()
    assert nc.plan("c64", tile_qubits=12)["qphase_gates"] > 0
    a0 = random_state(n, 3 * n)
    exp = a0.copy()
    oracle.apply_circuit(exp, n, c.gates)
    got = nc.emulate(a0, "c64", batch_layout=True)
    assert np.max(np.abs(got - exp)) < 1e-12
