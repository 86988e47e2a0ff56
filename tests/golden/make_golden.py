"""Generate golden vectors for the statevector path by dense brute force.

Independent of both the oracle and the product: gate matrices are written
here from GateType::matrix() (sdk/rust/src/core/gates/gates.rs:70-598) and
applied by tensor contraction of the full 2^n state (numpy.tensordot), and
Pauli expectations are <psi|P|psi> with P the dense Kronecker product.  The
reference publishes no golden files (its tests are closed-form assertions),
so these vectors are what the oracle and the GPU path are checked against.

Run:  python tests/golden/make_golden.py   (writes tests/golden/golden.npz)
"""
from __future__ import annotations

import os

import numpy as np

KINDS = ["h", "x", "y", "z", "s", "t", "rx", "ry", "rz", "u", "cnot", "cz", "swap",
         "toffoli", "crx", "cry", "crz", "cr", "sx", "phase", "iswap", "ccz", "rxx",
         "ryy", "rzz", "cswap", "cu", "custom1", "custom2"]
# (targets, controls, params) per kind
ARITY = {"h": (1, 0, 0), "x": (1, 0, 0), "y": (1, 0, 0), "z": (1, 0, 0), "s": (1, 0, 0), "t": (1, 0, 0),
         "rx": (1, 0, 1), "ry": (1, 0, 1), "rz": (1, 0, 1), "u": (1, 0, 3), "cnot": (1, 1, 0), "cz": (1, 1, 0),
         "swap": (2, 0, 0), "toffoli": (1, 2, 0), "crx": (1, 1, 1), "cry": (1, 1, 1), "crz": (1, 1, 1),
         "cr": (1, 1, 1), "sx": (1, 0, 0), "phase": (1, 0, 1), "iswap": (2, 0, 0), "ccz": (1, 2, 0),
         "rxx": (2, 0, 1), "ryy": (2, 0, 1), "rzz": (2, 0, 1), "cswap": (2, 1, 0), "cu": (1, 1, 4),
         "custom1": (1, 0, 0), "custom2": (2, 0, 0)}


def e(a):
    return np.exp(1j * a)


def matrix(kind, p, custom=None):
    """Dense matrix; index bits ordered as the qubit list from `order`."""
    c = lambda a: np.cos(a / 2)  # noqa: E731
    s = lambda a: np.sin(a / 2)  # noqa: E731
    I2 = np.eye(2, dtype=complex)
    if kind == "h":
        return np.array([[1, 1], [1, -1]], complex) / np.sqrt(2)
    if kind == "x":
        return np.array([[0, 1], [1, 0]], complex)
    if kind == "y":
        return np.array([[0, -1j], [1j, 0]], complex)
    if kind == "z":
        return np.diag([1, -1]).astype(complex)
    if kind == "s":
        return np.diag([1, 1j])
    if kind == "t":
        return np.diag([1, e(np.pi / 4)])
    if kind == "rx":
        return np.array([[c(p[0]), -1j * s(p[0])], [-1j * s(p[0]), c(p[0])]])
    if kind == "ry":
        return np.array([[c(p[0]), -s(p[0])], [s(p[0]), c(p[0])]], complex)
    if kind == "rz":
        return np.diag([e(-p[0] / 2), e(p[0] / 2)])
    if kind == "u":
        th, ph, la = p[:3]
        return np.array([[c(th), -e(la) * s(th)], [e(ph) * s(th), e(ph + la) * c(th)]])
    if kind == "sx":
        return 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]])
    if kind == "phase":
        return np.diag([1, e(p[0])])
    if kind == "custom1":
        return np.asarray(custom, complex).reshape(2, 2)
    if kind == "custom2":
        return np.asarray(custom, complex).reshape(4, 4)

    def ctrl(u):
        m = np.eye(4, dtype=complex)
        m[2:, 2:] = u
        return m
    if kind == "cnot":
        return ctrl(matrix("x", p))
    if kind == "cz":
        return ctrl(matrix("z", p))
    if kind == "crx":
        return ctrl(matrix("rx", p))
    if kind == "cry":
        return ctrl(matrix("ry", p))
    if kind == "crz":
        return ctrl(matrix("rz", p))
    if kind == "cr":
        return ctrl(np.diag([1, e(p[0])]))
    if kind == "cu":
        th, ph, la, ga = p[:4]
        return ctrl(e(ga) * matrix("u", (th, ph, la)))
    if kind == "swap":
        return np.eye(4, dtype=complex)[[0, 2, 1, 3]]
    if kind == "iswap":
        m = np.eye(4, dtype=complex)
        m[1, 1] = m[2, 2] = 0
        m[1, 2] = m[2, 1] = 1j
        return m
    X = matrix("x", p)
    Y = matrix("y", p)
    Z = matrix("z", p)
    if kind == "rxx":
        return c(p[0]) * np.eye(4) - 1j * s(p[0]) * np.kron(X, X)
    if kind == "ryy":
        return c(p[0]) * np.eye(4) - 1j * s(p[0]) * np.kron(Y, Y)
    if kind == "rzz":
        return c(p[0]) * np.eye(4) - 1j * s(p[0]) * np.kron(Z, Z)
    if kind == "toffoli":
        m = np.eye(8, dtype=complex)
        m[6:, 6:] = X
        return m
    if kind == "ccz":
        return np.diag([1, 1, 1, 1, 1, 1, 1, -1]).astype(complex)
    if kind == "cswap":
        m = np.eye(8, dtype=complex)
        m[4:, 4:] = matrix("swap", p)
        return m
    raise KeyError(kind)


def order(kind, t, cq):
    """Qubits of the matrix index, most significant first."""
    if kind in ("cnot", "cz", "crx", "cry", "crz", "cr", "cu"):
        return [cq[0], t[0]]
    if kind in ("toffoli", "ccz"):
        return [cq[0], cq[1], t[0]]
    if kind == "cswap":
        return [cq[0], t[0], t[1]]
    return list(t)


def apply(psi, n, m, qubits):
    k = len(qubits)
    t = psi.reshape([2] * n)
    axes = [n - 1 - q for q in qubits]
    mt = m.reshape([2] * (2 * k))
    t = np.tensordot(mt, t, axes=(list(range(k, 2 * k)), axes))
    t = np.moveaxis(t, list(range(k)), axes)
    return t.reshape(-1)


def pauli_dense(pstr):
    P = {"I": np.eye(2), "X": matrix("x", ()), "Y": matrix("y", ()), "Z": matrix("z", ())}
    m = np.array([[1.0 + 0j]])
    for ch in reversed(pstr):  # qubit n-1 is the most significant factor
        m = np.kron(m, P[ch])
    return m


def random_unitary(rng, d):
    z = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_gate(rng, n, kind=None):
    kind = kind or KINDS[rng.integers(len(KINDS))]
    nt, nc, np_ = ARITY[kind]
    qs = rng.permutation(n)[: nt + nc]
    t, cq = [int(q) for q in qs[:nt]], [int(q) for q in qs[nt:]]
    p = [float(v) for v in rng.uniform(-2 * np.pi, 2 * np.pi, size=np_)]
    cust = None
    if kind == "custom1":
        cust = random_unitary(rng, 2)
    elif kind == "custom2":
        cust = random_unitary(rng, 4)
    return kind, t, cq, p, cust


def encode(gates):
    G = len(gates)
    kind = np.zeros(G, np.int32)
    tg = np.zeros((G, 3), np.int32)
    ct = np.zeros((G, 3), np.int32)
    nt = np.zeros(G, np.int32)
    nc = np.zeros(G, np.int32)
    pr = np.zeros((G, 4))
    mx = np.zeros((G, 16), complex)
    for i, (k, t, cq, p, cust) in enumerate(gates):
        kind[i] = KINDS.index(k)
        nt[i], nc[i] = len(t), len(cq)
        tg[i, : len(t)] = t
        ct[i, : len(cq)] = cq
        pr[i, : len(p)] = p
        if cust is not None:
            mx[i, : cust.size] = cust.reshape(-1)
    return dict(kind=kind, targets=tg, controls=ct, nt=nt, nc=nc, params=pr, matrix=mx)


def run_dense(n, gates, psi0=None):
    psi = np.zeros(1 << n, complex)
    psi[0] = 1
    if psi0 is not None:
        psi = psi0.copy()
    for k, t, cq, p, cust in gates:
        psi = apply(psi, n, matrix(k, p, cust), order(k, t, cq))
    return psi


def main():
    rng = np.random.default_rng(20261203)
    out = {}
    # 1. every kind on a random 5-qubit state, three draws each
    n = 5
    single = []
    for kind in KINDS:
        for _ in range(3):
            single.append(random_gate(rng, n, kind))
    psi0 = rng.normal(size=(len(single), 1 << n)) + 1j * rng.normal(size=(len(single), 1 << n))
    psi0 /= np.linalg.norm(psi0, axis=1, keepdims=True)
    res = np.stack([run_dense(n, [g], psi0[i]) for i, g in enumerate(single)])
    for k, v in encode(single).items():
        out[f"single_{k}"] = v
    out["single_in"] = psi0
    out["single_out"] = res
    out["single_n"] = np.int32(n)
    # 2. random mixed circuits from |0..0>
    for idx, (nq, ng) in enumerate([(6, 80), (7, 120), (8, 160)]):
        gates = [random_gate(rng, nq) for _ in range(ng)]
        for k, v in encode(gates).items():
            out[f"circ{idx}_{k}"] = v
        out[f"circ{idx}_n"] = np.int32(nq)
        out[f"circ{idx}_out"] = run_dense(nq, gates)
    # 3. Pauli strings on the 8-qubit circuit state
    psi = out["circ2_out"]
    paulis = ["".join("IXYZ"[rng.integers(4)] for _ in range(8)) for _ in range(24)]
    paulis += ["YIIIIIII", "IYYIIIII", "ZZIIIIII", "XXXXXXXX", "IIIIIIII"]
    out["pauli_strings"] = np.array(paulis)
    out["pauli_values"] = np.array([np.vdot(psi, pauli_dense(p) @ psi).real for p in paulis])
    out["z_values"] = np.array([np.vdot(psi, pauli_dense("I" * q + "Z" + "I" * (7 - q)) @ psi).real
                                for q in range(8)])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
