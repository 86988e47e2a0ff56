"""The nqpu_metal binding surface that NQPUBackend._run_with_rust uses
(sdk/python/nqpu/core/simulator.py:145-191), served by the GPU engine."""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "nqpu_amd", "compat"))

import nqpu_metal as bindings  # noqa: E402


def test_surface_matches_the_backend_contract():
    c = bindings.QuantumCircuit(3)
    for name in ("h", "x", "y", "z", "s", "t", "rx", "ry", "rz", "cx", "cz", "swap"):
        assert callable(getattr(c, name))
    for name in ("CPU", "GPU", "AUTO", "Auto"):
        assert getattr(bindings.Backend, name, None) is not None
    assert callable(bindings.Simulator(bindings.Backend.GPU).run)
    with pytest.raises(ValueError):
        c.h(3)
    with pytest.raises(ValueError):
        c.cx(1, 1)
    with pytest.raises(ValueError):
        bindings.QuantumCircuit(35)


def _run_like_nqpu_backend(gates, n, shots, backend="GPU"):
    """The call sequence of NQPUBackend._run_with_rust, step by step."""
    rust_circuit = bindings.QuantumCircuit(n)
    for name, qubits, params in gates:
        getattr(rust_circuit, name)(*qubits, *params)
    rust_backend = getattr(bindings.Backend, backend, None)
    simulator = bindings.Simulator(rust_backend) if rust_backend is not None else bindings.Simulator()
    raw = simulator.run(rust_circuit, shots)
    return raw.counts if hasattr(raw, "counts") else raw.results()


@pytest.mark.gpu
def test_bell_state_through_the_backend_path():
    bindings.set_seed(7)
    counts = _run_like_nqpu_backend([("h", (0,), ()), ("cx", (0, 1), ())], 2, 512)
    dominant = counts.get("00", 0) + counts.get("11", 0)   # test_model_qpu.py:472-483
    assert dominant == 512 and min(counts.values()) > 200
    assert all(len(k) == 2 for k in counts)


@pytest.mark.gpu
def test_counts_follow_the_oracle_distribution(oracle):
    rng = np.random.default_rng(3)
    n, gates = 6, []
    for _ in range(60):
        k = ["h", "x", "y", "z", "s", "t", "rx", "ry", "rz", "cx", "cz", "swap"][rng.integers(12)]
        if k in ("cx", "cz", "swap"):
            a, b = rng.choice(n, 2, replace=False)
            gates.append((k, (int(a), int(b)), ()))
        elif k in ("rx", "ry", "rz"):
            gates.append((k, (int(rng.integers(n)),), (float(rng.uniform(-3, 3)),)))
        else:
            gates.append((k, (int(rng.integers(n)),), ()))
    bindings.set_seed(11)
    shots = 40000
    counts = _run_like_nqpu_backend(gates, n, shots)
    from nqpu_amd.engine import Gate
    kinds = {"cx": "cnot"}
    og = []
    for name, qs, ps in gates:
        if name in ("cx", "cz"):
            og.append(Gate(kinds.get(name, name), (qs[1],), (qs[0],), ()))
        else:
            og.append(Gate(name, tuple(qs), (), tuple(ps)))
    p = oracle.probabilities(oracle.run(n, og), n)
    obs = np.zeros(1 << n)
    for bits, c in counts.items():
        obs[int(bits, 2)] = c  # MSB-first bitstrings (lib.rs:553-564)
    keep = p * shots > 5
    chi2 = np.sum((obs[keep] - shots * p[keep]) ** 2 / (shots * p[keep]))
    dof = keep.sum() - 1
    assert chi2 < dof + 6 * math.sqrt(2 * dof)
