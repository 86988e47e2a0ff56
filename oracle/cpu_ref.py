"""ctypes wrapper over the FP64 CPU oracle (oracle/cpu_ref.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by ``bench.py``'s ``cpu_baseline`` leg, always as the checker or the
timed CPU baseline, never as the product.  The product path (``nqpu_amd``)
never imports this module.

Gates are any objects with ``kind`` (a name from ``KINDS``), ``targets``,
``controls``, ``params`` and an optional ``matrix`` (complex, 2x2 or 4x4),
mirroring ``gates::Gate`` (sdk/rust/src/core/gates/gates.rs:632).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libcpu_ref.so")

# The order of GateType (gates.rs:8-64); values equal NqpuGateKind.
KINDS = ["h", "x", "y", "z", "s", "t", "rx", "ry", "rz", "u", "cnot", "cz", "swap",
         "toffoli", "crx", "cry", "crz", "cr", "sx", "phase", "iswap", "ccz", "rxx",
         "ryy", "rzz", "cswap", "cu", "custom1", "custom2"]
KIND_ID = {k: i for i, k in enumerate(KINDS)}


class OrcGate(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("nt", ct.c_int32), ("nc", ct.c_int32),
                ("t", ct.c_int32 * 3), ("c", ct.c_int32 * 3), ("pad", ct.c_int32),
                ("p", ct.c_double * 4), ("m", ct.c_double * 32)]


def build() -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        P = ct.c_void_p
        L.orc_init_zero.argtypes = [P, ct.c_int]
        L.orc_gate_matrix.argtypes = [ct.POINTER(OrcGate), P, ct.POINTER(ct.c_int)]
        L.orc_u1.argtypes = [P, ct.c_int, ct.c_int, P]
        L.orc_u2.argtypes = [P, ct.c_int, ct.c_int, ct.c_int, P]
        L.orc_apply_gate.argtypes = [P, ct.c_int, ct.POINTER(OrcGate), ct.c_int]
        L.orc_apply_circuit.argtypes = [P, ct.c_int, ct.POINTER(OrcGate), ct.c_int, ct.c_int]
        L.orc_execute_diagonal.argtypes = [P, ct.c_int, P, P, ct.c_int, P, P, ct.c_int]
        L.orc_norm.argtypes = [P, ct.c_int]
        L.orc_norm.restype = ct.c_double
        L.orc_probabilities.argtypes = [P, ct.c_int, P]
        for f in ("orc_expectation_z", "orc_expectation_x"):
            getattr(L, f).argtypes = [P, ct.c_int, ct.c_int]
            getattr(L, f).restype = ct.c_double
        L.orc_expectation_y.argtypes = [P, ct.c_int, ct.c_int, ct.c_int]
        L.orc_expectation_y.restype = ct.c_double
        L.orc_expectation_pauli.argtypes = [P, ct.c_int, ct.c_uint64, ct.c_uint64, ct.c_int]
        L.orc_expectation_pauli.restype = ct.c_double
        L.orc_fidelity.argtypes = [P, P, ct.c_int]
        L.orc_fidelity.restype = ct.c_double
        L.orc_rng_seed.argtypes = [ct.c_uint64]
        L.orc_rng_next.restype = ct.c_double
        L.orc_rng_state.restype = ct.c_uint64
        L.orc_reset_qubit.argtypes = [P, ct.c_int, ct.c_int]
        L.orc_initialize_from_amplitudes.argtypes = [P, ct.c_int, P, ct.c_uint64]
        L.orc_initialize_from_amplitudes.restype = ct.c_int
        L.orc_pauli_expectation_msb.argtypes = [P, ct.c_int, P, ct.c_char_p, ct.c_int, ct.c_double]
        L.orc_pauli_expectation_msb.restype = ct.c_double
        L.orc_sample_cabi.argtypes = [P, ct.c_int, ct.c_uint32, P]
        L.orc_sample_cdf.argtypes = [P, ct.c_int, P, ct.c_uint64, P]
        L.orc_measure_qubit.argtypes = [P, ct.c_int, ct.c_int, ct.POINTER(ct.c_double)]
        L.orc_measure_qubit.restype = ct.c_uint32
        L.orc_num_threads.restype = ct.c_int
        L.orc_set_num_threads.argtypes = [ct.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def to_struct(g) -> OrcGate:
    s = OrcGate()
    s.kind = KIND_ID[g.kind]
    s.nt = len(g.targets)
    s.nc = len(g.controls)
    for i, q in enumerate(g.targets):
        s.t[i] = int(q)
    for i, q in enumerate(g.controls):
        s.c[i] = int(q)
    for i, v in enumerate(g.params):
        s.p[i] = float(v)
    m = getattr(g, "matrix", None)
    if m is not None:
        flat = np.asarray(m, dtype=np.complex128).reshape(-1)
        for i, v in enumerate(flat):
            s.m[2 * i] = v.real
            s.m[2 * i + 1] = v.imag
    return s


def zero_state(n: int) -> np.ndarray:
    a = np.empty(1 << n, dtype=np.complex128)
    lib().orc_init_zero(_ptr(a), n)
    return a


def gate_matrix(g) -> np.ndarray:
    out = np.zeros(64, dtype=np.complex128)
    d = ct.c_int(0)
    s = to_struct(g)
    lib().orc_gate_matrix(ct.byref(s), _ptr(out), ct.byref(d))
    return out[: d.value * d.value].reshape(d.value, d.value)


def apply_gate(a: np.ndarray, n: int, g, quirk: bool = False) -> None:
    s = to_struct(g)
    lib().orc_apply_gate(_ptr(a), n, ct.byref(s), int(quirk))


def apply_circuit(a: np.ndarray, n: int, gates, quirk: bool = False) -> None:
    arr = (OrcGate * max(1, len(gates)))()
    for i, g in enumerate(gates):
        arr[i] = to_struct(g)
    lib().orc_apply_circuit(_ptr(a), n, arr, len(gates), int(quirk))


def run(n: int, gates, quirk: bool = False) -> np.ndarray:
    a = zero_state(n)
    apply_circuit(a, n, gates, quirk)
    return a


def u1(a, n, q, m):
    m = np.ascontiguousarray(m, dtype=np.complex128)
    lib().orc_u1(_ptr(a), n, q, _ptr(m))


def u2(a, n, qlo, qhi, m):
    m = np.ascontiguousarray(m, dtype=np.complex128)
    lib().orc_u2(_ptr(a), n, qlo, qhi, _ptr(m))


def execute_diagonal(a, n, single, pairs):
    """single: [(q, p0, p1)], pairs: [(lo, hi, p11)] (gate_fusion.rs:905)."""
    sq = np.array([s[0] for s in single] or [0], dtype=np.int32)
    sp = np.array([v for s in single for v in (s[1], s[2])] or [1, 1], dtype=np.complex128)
    pq = np.array([v for p in pairs for v in (p[0], p[1])] or [0, 0], dtype=np.int32)
    pp = np.array([p[2] for p in pairs] or [1], dtype=np.complex128)
    lib().orc_execute_diagonal(_ptr(a), n, _ptr(sq), _ptr(sp), len(single), _ptr(pq), _ptr(pp), len(pairs))


def norm(a, n):
    return lib().orc_norm(_ptr(a), n)


def probabilities(a, n):
    out = np.empty(1 << n, dtype=np.float64)
    lib().orc_probabilities(_ptr(a), n, _ptr(out))
    return out


def expectation_z(a, n, q):
    return lib().orc_expectation_z(_ptr(a), n, q)


def expectation_x(a, n, q):
    return lib().orc_expectation_x(_ptr(a), n, q)


def expectation_y(a, n, q, quirk=False):
    return lib().orc_expectation_y(_ptr(a), n, q, int(quirk))


def expectation_pauli(a, n, xmask, zmask, quirk=False):
    return lib().orc_expectation_pauli(_ptr(a), n, int(xmask), int(zmask), int(quirk))


def fidelity(a, b, n):
    return lib().orc_fidelity(_ptr(a), _ptr(b), n)


def rng_seed(seed: int):
    lib().orc_rng_seed(seed)


def rng_next() -> float:
    return lib().orc_rng_next()


def rng_state() -> int:
    """Raw 64-bit state of the generator."""
    return int(lib().orc_rng_state())


def reset_qubit(a, n, q):
    """reset_qubit (lib.rs:1705-1730), in place."""
    lib().orc_reset_qubit(_ptr(a), n, q)


def initialize_from_amplitudes(n, src):
    """initialize_from_amplitudes (lib.rs:1918-1935): the normalised copy, or
    None when the length is not 2^n or the norm is below 1e-10."""
    src = np.ascontiguousarray(src, dtype=np.complex128)
    dst = np.zeros(1 << n, dtype=np.complex128)
    ok = lib().orc_initialize_from_amplitudes(_ptr(dst), n, _ptr(src), len(src))
    return dst if ok else None


def pauli_expectation_msb(a, n, term, coeff=1.0):
    """primitives::pauli_expectation (primitives.rs:683-728): ``term`` is a
    list of (qubit, 'I'|'X'|'Y'|'Z') with qubit 0 the most significant bit."""
    qs = np.array([int(q) for q, _ in term] or [0], dtype=np.int32)
    ps = "".join(p.upper() for _, p in term).encode()
    return lib().orc_pauli_expectation_msb(_ptr(a), n, _ptr(qs), ps, len(term), float(coeff))


def sample_cabi(a, n, shots, seed=None):
    if seed is not None:
        rng_seed(seed)
    out = np.empty(shots, dtype=np.uint64)
    lib().orc_sample_cabi(_ptr(a), n, shots, _ptr(out))
    return out


def sample_cdf(a, n, r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.empty(len(r), dtype=np.uint64)
    lib().orc_sample_cdf(_ptr(a), n, _ptr(r), len(r), _ptr(out))
    return out


def measure_qubit(a, n, q):
    p = ct.c_double(0)
    o = lib().orc_measure_qubit(_ptr(a), n, q, ct.byref(p))
    return o, p.value


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(t: int) -> None:
    lib().orc_set_num_threads(t)
