"""``nqpu_metal`` — the Python binding surface nQPU's NumPy-free backend path
expects, served by the MI355X engine.

nQPU's ``NQPUBackend.run(circuit, shots)`` (sdk/python/nqpu/core/simulator.py:135-191)
imports ``nqpu_metal`` (sdk/python/nqpu/_compat.py:20-36) and uses exactly:
``QuantumCircuit(n)`` with ``h x y z s t rx ry rz cx cz swap``,
``Backend.<CPU|GPU|AUTO>``, ``Simulator([backend])`` and
``Simulator.run(circuit, shots)`` returning an object with ``counts`` (MSB-first
bitstring -> count) and ``results()``; the reference's own implementation of
that surface is sdk/rust/src/infra/bindings/python_api_v2.rs:19-289.  Put this
directory on ``sys.path`` (``nqpu_amd/compat``) and ``NQPUBackend.run``
executes on the GPU unchanged.

Differences from python_api_v2.rs, all deliberate: the qubit cap is 34 instead
of 28 (python_api_v2.rs:45-47); every Backend value runs on the GPU engine;
shots are drawn with the engine's seeded sampler in one pass instead of one
state scan per shot (python_api_v2.rs:270-276).
"""
from __future__ import annotations

import enum
import os
import struct

from nqpu_amd import engine as _E

__all__ = ["Backend", "QuantumCircuit", "Simulator", "Result", "PI", "VERSION", "set_seed"]

PI = 3.141592653589793
VERSION = "1.0.0"
MAX_QUBITS = 34


class Backend(enum.Enum):
    """python_api_v2.rs:19-25 (CPU, GPU, Auto) plus ROCM; all run on the GPU."""

    CPU = "cpu"
    GPU = "gpu"
    Auto = "auto"
    AUTO = "auto"  # alias: the nqpu Python enum spells it AUTO
    ROCM = "rocm"


class QuantumCircuit:
    """python_api_v2.rs:32-176: an ordered gate list; qubit q = bit q."""

    def __init__(self, num_qubits: int):
        if num_qubits > MAX_QUBITS or num_qubits < 1:
            raise ValueError(f"1 to {MAX_QUBITS} qubits supported")
        self._c = _E.Circuit(num_qubits)

    @property
    def num_qubits(self) -> int:
        return self._c.num_qubits

    def _q(self, *qs):
        for q in qs:
            if q < 0 or q >= self.num_qubits:
                raise ValueError("Qubit index out of range")

    def h(self, q): self._q(q); self._c.h(q)
    def x(self, q): self._q(q); self._c.x(q)
    def y(self, q): self._q(q); self._c.y(q)
    def z(self, q): self._q(q); self._c.z(q)
    def s(self, q): self._q(q); self._c.s(q)
    def t(self, q): self._q(q); self._c.t(q)
    def rx(self, q, theta): self._q(q); self._c.rx(q, theta)
    def ry(self, q, theta): self._q(q); self._c.ry(q, theta)
    def rz(self, q, theta): self._q(q); self._c.rz(q, theta)

    def cx(self, control, target):
        self._q(control, target)
        if control == target:
            raise ValueError("control and target must differ")
        self._c.cnot(control, target)

    cnot = cx

    def cz(self, control, target):
        self._q(control, target)
        if control == target:
            raise ValueError("control and target must differ")
        self._c.cz(control, target)

    def swap(self, a, b):
        self._q(a, b)
        if a == b:
            raise ValueError("swap targets must differ")
        self._c.swap(a, b)

    def gate_count(self) -> int:
        return len(self._c.gates)

    def __repr__(self):
        return f"<QuantumCircuit {self.num_qubits} qubits, {len(self._c.gates)} gates>"


class Result:
    """python_api_v2.rs:180-208."""

    def __init__(self, counts: dict, shots: int):
        self._counts = counts
        self.shots = shots

    @property
    def counts(self) -> dict:
        return dict(self._counts)

    def results(self) -> dict:
        return dict(self._counts)

    def __repr__(self):
        return f"<Result shots={self.shots}, outcomes={len(self._counts)}>"


_seed = None


def set_seed(seed: int | None) -> None:
    """Seed the shot sampler (the C ABI's xorshift64* stream)."""
    global _seed
    _seed = seed


class Simulator:
    """python_api_v2.rs:214-289: runs a circuit from |0...0> and samples shots."""

    def __init__(self, backend: Backend | None = None):
        self.backend = backend or Backend.Auto

    def run(self, circuit: QuantumCircuit, shots: int | None = None) -> Result:
        shots = 1024 if shots is None else int(shots)
        n = circuit.num_qubits
        precision = "c128" if n <= 30 else "c64"
        seed = _seed if _seed is not None else struct.unpack("<Q", os.urandom(8))[0]
        with _E.StateVector(n, precision) as sv:
            sv.run(circuit._c)
            counts = sv.sample_counts(shots, seed=seed) if shots > 0 else {}
        return Result(counts, shots)

    def __repr__(self):
        return f"<Simulator backend={self.backend}>"
