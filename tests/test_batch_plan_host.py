"""Host half of batched execution (no device): circuits are planned on a pool
of host threads and their tables concatenated in circuit order, so the plan
is the same on every call; errors of any circuit reach the caller."""
import pytest

from nqpu_amd import _lib as L
from nqpu_amd.engine import Batch, Circuit
from nqpu_amd.workloads import variational_batch


def test_plan_is_deterministic_and_counts_passes():
    vcs = variational_batch(300, seed=103)
    circs = [v.circuit.native() for v in vcs]
    obs = [v.hamiltonian() for v in vcs]
    a = Batch.plan(circs, obs)
    b = Batch.plan(circs, obs)
    for k in ("circuits", "resident", "gate_passes", "observable_passes", "state_bytes", "program_ops"):
        assert a[k] == b[k], k
    assert a["circuits"] == 300
    small = sum(1 for v in vcs if v.circuit.num_qubits <= 12)
    assert a["resident"] == small
    assert a["gate_passes"] > 0 and a["observable_passes"] > 0
    # the plan's state bytes are one workgroup slot (the largest HBM-resident
    # circuit below 18 qubits) plus a state of its own for each circuit of 18
    # qubits or more, which runs as grids of work items
    slot = max(1 << v.circuit.num_qubits for v in vcs if 12 < v.circuit.num_qubits < 18)
    split = sum(1 << v.circuit.num_qubits for v in vcs if v.circuit.num_qubits >= 18)
    assert split > 0
    assert a["state_bytes"] == 8.0 * (slot + split)


def test_circuits_without_gate_passes_still_get_one():
    """A 14-qubit circuit whose gates all cancel (x.x) or that has none still
    gets one (empty) gate pass: that pass writes the |0...0> start state its
    observable passes read."""
    empty = Circuit(14)
    cancel = Circuit(14).x(0).x(0).rz(3, 0.4).rz(3, -0.4)
    for c in (empty, cancel):
        p = Batch.plan([c.native()], [[(1.0, "Z" + "I" * 13)]])
        assert p["gate_passes"] == 1 and p["observable_passes"] == 1


def test_error_of_one_circuit_is_reported():
    vcs = variational_batch(100, seed=5)
    circs = [v.circuit.native() for v in vcs]
    obs = [v.hamiltonian() for v in vcs]
    tiny = Circuit(3)
    tiny.h(0)
    circs[57] = tiny.native()
    obs[57] = [(1.0, "ZII")]
    with pytest.raises(L.NqpuError) as e:
        Batch.plan(circs, obs)
    assert e.value.code == 2  # INVALID_PARAMETER: batched circuits need at least 4 qubits


def test_split_circuits_run_in_waves_of_bounded_memory():
    """Circuits of 18 qubits and more get states of their own, in waves that
    fit 2^30 complex64 amplitudes (8 GiB): 1100 circuits of 20 qubits need two
    waves, and the plan's state memory is one wave's."""
    vcs = variational_batch(1100, seed=23, nmin=20, nmax=20, lmin=1, lmax=1)
    circs = [v.circuit.native() for v in vcs]
    obs = [v.hamiltonian() for v in vcs]
    p = Batch.plan(circs, obs)
    assert p["state_bytes"] == 8.0 * (1 << 30)
    few = Batch.plan(circs[:100], obs[:100])
    assert few["state_bytes"] == 8.0 * 100 * (1 << 20)
