"""Seeded synthetic workloads of the statevector benchmarks.

Every random draw comes from a SplitMix64 stream (uniform double =
(x >> 11) * 2^-53), consumed in the order written here, so the same seed
gives the same circuit everywhere.

* ``random_u3_brickwork``: per layer a U3 on every qubit (theta, phi, lambda
  uniform in [0, 2pi)), then CNOT(q, q+1) for q = l%2, l%2+2, ... < n-1.
  20 qubits / 20 layers / seed 101 is the complex128 parity workload;
  30 qubits / 40 layers / seed 102 is the single-large-state workload.
* ``variational_batch``: 4096 hardware-efficient-ansatz circuits (RY.RZ per
  qubit per layer plus a CZ ladder) with transverse-field Ising energies.
* ``qft`` (benchmark_suite.rs:25-40) and ``tfim_trotter``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

from .engine import Circuit

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + min(int(self.uniform() * (hi - lo + 1)), hi - lo)


def random_u3_brickwork(n: int, layers: int, seed: int) -> Circuit:
    rng = SplitMix64(seed)
    c = Circuit(n)
    two_pi = 2.0 * math.pi
    for layer in range(layers):
        for q in range(n):
            th, ph, la = two_pi * rng.uniform(), two_pi * rng.uniform(), two_pi * rng.uniform()
            c.u(q, th, ph, la)
        for q in range(layer % 2, n - 1, 2):
            c.cnot(q, q + 1)
    return c


@dataclass
class VariationalCircuit:
    circuit: Circuit
    g: float

    def hamiltonian(self):
        """TFIM H = -sum Z_i Z_{i+1} - g sum X_i on an open chain
        (sdk/python/nqpu/simulation/hamiltonians.py:435-470)."""
        n = self.circuit.num_qubits
        terms = []
        for i in range(n - 1):
            s = ["I"] * n
            s[i] = s[i + 1] = "Z"
            terms.append((-1.0, "".join(s)))
        for i in range(n):
            s = ["I"] * n
            s[i] = "X"
            terms.append((-self.g, "".join(s)))
        return terms


def variational_batch(count: int = 4096, seed: int = 103, nmin: int = 8, nmax: int = 20,
                      lmin: int = 4, lmax: int = 12):
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = rng.randint(nmin, nmax)
        layers = rng.randint(lmin, lmax)
        g = 0.2 + 1.6 * rng.uniform()
        c = Circuit(n)
        for _layer in range(layers):
            for q in range(n):
                a = -math.pi + 2 * math.pi * rng.uniform()
                b = -math.pi + 2 * math.pi * rng.uniform()
                c.ry(q, a)
                c.rz(q, b)
            for q in range(n - 1):
                c.cz(q, q + 1)
        out.append(VariationalCircuit(c, g))
    return out


def qft(n: int, circuit: Circuit | None = None) -> Circuit:
    """QFT as benchmark_suite.rs:25-40: H(i), CR(pi/2^k) control i target j, swaps."""
    c = circuit or Circuit(n)
    for i in range(n):
        c.h(i)
        for j in range(i + 1, n):
            c.cr(i, j, math.pi / (1 << (j - i)))
    for i in range(n // 2):
        c.swap(i, n - 1 - i)
    return c


def tfim_trotter(n: int, steps: int = 20, dt: float = 0.05, g: float = 1.0,
                 circuit: Circuit | None = None) -> Circuit:
    """First-order Trotter steps of H = -sum ZZ - g sum X: each step applies
    RZZ(-2 dt) on every neighbour pair, then RX(-2 g dt) on every qubit."""
    c = circuit or Circuit(n)
    for _ in range(steps):
        for i in range(n - 1):
            c.rzz(i, i + 1, -2.0 * dt)
        for i in range(n):
            c.rx(i, -2.0 * g * dt)
    return c


def qft_trotter(n: int, steps: int = 20) -> Circuit:
    c = qft(n)
    return tfim_trotter(n, steps, circuit=c)


@dataclass
class ServedRequest:
    arrival: float          # seconds after the stream starts
    circuit: Circuit
    terms: list             # [(1.0, pauli string)]
    shots: int
    seed: int               # seed of the request's shot stream


SERVED_QUBITS = (10, 14, 18, 22, 26, 28)
SERVED_WEIGHTS = (0.30, 0.25, 0.20, 0.12, 0.08, 0.05)


def served_stream(count: int, qps: float, seed: int = 104, shots: int = 1024, nterms: int = 8,
                  qubits=SERVED_QUBITS, weights=SERVED_WEIGHTS, dmin: int = 10, dmax: int = 100):
    """Poisson arrivals at `qps` requests/s.  Per request, in stream order:
    the exponential inter-arrival gap, the qubit count (drawn with `weights`),
    the depth ~ U{dmin..dmax} of random U3 + CNOT-brickwork layers (circuit
    seed from the stream), `nterms` Pauli strings with every qubit ~ U{I,X,Y,Z},
    and the seed of its shot stream."""
    rng = SplitMix64(seed)
    cum = []
    acc = 0.0
    for w in weights:
        acc += w
        cum.append(acc)
    t = 0.0
    out = []
    for _ in range(count):
        t += -math.log(1.0 - rng.uniform()) / qps
        u = rng.uniform() * acc
        n = qubits[next(i for i, c in enumerate(cum) if u < c or i == len(cum) - 1)]
        depth = rng.randint(dmin, dmax)
        circ = random_u3_brickwork(n, depth, rng.next_u64())
        terms = [(1.0, "".join("IXYZ"[rng.randint(0, 3)] for _q in range(n))) for _k in range(nterms)]
        out.append(ServedRequest(t, circ, terms, shots, rng.next_u64()))
    return out
