"""Batched execution (one persistent launch over many complex64 circuits)
against the FP64 oracle, and the alone-vs-mixed identity of every result."""
import numpy as np
import pytest

from nqpu_amd import engine as E
from nqpu_amd.engine import Batch, Circuit
from nqpu_amd.workloads import variational_batch
from tests.test_planner_emulation import diagonal_heavy_circuit, random_circuit

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    if E.device_count() == 0:
        pytest.fail("no HIP device visible")
    b = Batch()
    yield b
    b.close()


def oracle_energy(oracle, circuit, terms):
    n = circuit.num_qubits
    a = oracle.run(n, circuit.gates)
    tot = 0.0
    for coeff, p in terms:
        x, z = E.pauli_masks(p)
        tot += coeff * oracle.expectation_pauli(a, n, x, z)
    return tot


def test_variational_energies_match_oracle(batch, oracle):
    """The full qubit range of the batched workload (8-20)."""
    vcs = variational_batch(40, seed=103, nmin=8, nmax=20, lmin=4, lmax=12)
    assert max(v.circuit.num_qubits for v in vcs) >= 19
    e = batch.expectation([v.circuit for v in vcs], [v.hamiltonian() for v in vcs])
    ref = np.array([oracle_energy(oracle, v.circuit, v.hamiltonian()) for v in vcs])
    assert np.max(np.abs(e - ref)) < 2e-5
    st = batch.last_stats()
    assert st["circuits"] == 40 and 0 < st["resident"] < 40


def test_alone_equals_mixed_bitwise(batch):
    vcs = variational_batch(24, seed=7, nmin=8, nmax=15)
    circs = [v.circuit for v in vcs]
    obs = [v.hamiltonian() for v in vcs]
    mixed = batch.expectation(circs, obs)
    again = batch.expectation(circs[::-1], obs[::-1])[::-1]
    assert np.array_equal(mixed, again)
    for k in (0, 5, 11, 23):
        alone = batch.expectation([circs[k]], [obs[k]])
        assert alone[0] == mixed[k]


def test_general_pauli_terms(batch, oracle):
    rng = np.random.default_rng(5)
    circs, obs = [], []
    for n in (4, 6, 9, 12, 13, 15):
        c = random_circuit(n, 120, n)
        terms = [(float(rng.normal()), "".join("IXYZ"[k] for k in rng.integers(0, 4, n))) for _ in range(6)]
        if n == 15:  # X on both ends: no standard tile set holds it
            terms.append((0.5, "X" + "I" * 13 + "Y"))
        circs.append(c)
        obs.append(terms)
    e = batch.expectation(circs, obs)
    ref = np.array([oracle_energy(oracle, c, t) for c, t in zip(circs, obs)])
    assert np.max(np.abs(e - ref)) < 2e-5


def test_quadratic_phase_passes_match_oracle(batch, oracle):
    """Circuits above 12 qubits whose HBM passes hold matrix-core blocks and
    quadratic phase ops, beside resident ones, in one launch."""
    rng = np.random.default_rng(11)
    circs, obs = [], []
    for k, n in enumerate((13, 16, 10, 14)):
        c = diagonal_heavy_circuit(n, 300, 60 + k)
        terms = [(float(rng.normal()), "".join("IXYZ"[j] for j in rng.integers(0, 4, n))) for _ in range(5)]
        circs.append(c)
        obs.append(terms)
    e = batch.expectation(circs, obs)
    ref = np.array([oracle_energy(oracle, c, t) for c, t in zip(circs, obs)])
    assert np.max(np.abs(e - ref)) < 2e-5


def test_circuits_without_gate_passes(batch):
    """Circuits above 12 qubits with no gate left after merging (empty, or
    gates that cancel) start from |0...0>: <Z_0> = 1 and <X_0> = 0, however
    the workgroup's state slot was left by an earlier circuit."""
    noisy = [random_circuit(14, 200, 3), random_circuit(13, 200, 4)]
    empty = [Circuit(14), Circuit(13).x(0).x(0).rz(5, 0.3).rz(5, -0.3), Circuit(14).h(2).h(2)]
    circs = noisy + empty + noisy + empty
    obs = [[(1.0, "Z" + "I" * (c.num_qubits - 1)), (1.0, "X" + "I" * (c.num_qubits - 1))] for c in circs]
    for _ in range(2):
        e = batch.expectation(circs, obs)
        for k in (2, 3, 4, 7, 8, 9):
            assert abs(e[k] - 1.0) < 1e-6, k


def test_batch_argument_errors(batch):
    with pytest.raises(RuntimeError):
        batch.expectation([Circuit(3).h(0)], [[(1.0, "ZII")]])
    with pytest.raises(RuntimeError):
        batch.expectation([Circuit(5).h(0)], [[(1.0, 0, 1 << 7)]])


def test_queued_batches_equal_synchronous_batches(batch):
    """Three batches queued with async_=True (the third waits for the first's
    slot) give, after synchronize(), exactly the energies of the same batches
    run one at a time; a larger batch in between grows the device buffers."""
    sets = [variational_batch(k, seed=s, nmin=8, nmax=nmax) for k, s, nmax in ((30, 11, 16), (200, 12, 20), (17, 13, 14))]
    want = [batch.expectation([v.circuit for v in vs], [v.hamiltonian() for v in vs]) for vs in sets]
    got = [batch.expectation([v.circuit for v in vs], [v.hamiltonian() for v in vs], async_=True) for vs in sets]
    batch.synchronize()
    for k in range(3):
        assert np.array_equal(got[k], want[k]), f"batch {k}"
    # and again, now with every buffer already large enough
    got = [batch.expectation([v.circuit for v in vs], [v.hamiltonian() for v in vs], async_=True) for vs in sets]
    batch.synchronize()
    for k in range(3):
        assert np.array_equal(got[k], want[k]), f"batch {k} (second round)"
    assert batch.last_stats()["circuits"] == 17


def test_split_circuits_alone_equal_mixed_and_oracle(batch, oracle):
    """Circuits of 18 qubits and more run their passes as grids of work items
    (several workgroups per circuit) next to the persistent launch: each
    energy is the same alone and in a mix, and the oracle's."""
    vcs = variational_batch(40, seed=21, nmin=16, nmax=20, lmin=2, lmax=5)
    vcs += [v for v in variational_batch(60, seed=22, nmin=8, nmax=12)]
    assert sum(1 for v in vcs if v.circuit.num_qubits >= 18) >= 10
    circs = [v.circuit for v in vcs]
    obs = [v.hamiltonian() for v in vcs]
    mixed = batch.expectation(circs, obs)
    for k in range(0, len(vcs), 7):
        alone = batch.expectation([circs[k]], [obs[k]])
        assert alone[0] == mixed[k], k
    big = [k for k, v in enumerate(vcs) if v.circuit.num_qubits >= 18][:4]
    for k in big:
        assert abs(mixed[k] - oracle_energy(oracle, vcs[k].circuit, obs[k])) < 2e-5
    # a split circuit without observables gets 0
    e = batch.expectation([circs[big[0]], circs[0]], [[], obs[0]])
    assert e[0] == 0.0 and e[1] == mixed[0]


def test_split_circuits_in_two_waves(batch):
    """1100 circuits of 20 qubits exceed one wave's state memory (1024 of
    them): the batch runs them in two waves over the same states, and every
    energy equals the one from a batch of one wave."""
    vcs = variational_batch(1100, seed=23, nmin=20, nmax=20, lmin=1, lmax=1)
    circs = [v.circuit for v in vcs]
    obs = [v.hamiltonian() for v in vcs]
    both = batch.expectation(circs, obs)
    assert batch.last_stats()["state_bytes"] == 8.0 * (1 << 30)
    first = batch.expectation(circs[:600], obs[:600])
    last = batch.expectation(circs[600:], obs[600:])
    assert np.array_equal(both, np.concatenate([first, last]))
