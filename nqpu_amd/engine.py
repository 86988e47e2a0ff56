"""Python host API over the C ABI: gates, circuits and device-resident states.

Mirrors the reference's statevector surface — ``QuantumState`` /
``QuantumSimulator`` (sdk/rust/src/lib.rs:326-564, 1738-2070) and the C ABI
(sdk/rust/src/infra/bindings/c_ffi.rs) — with amplitudes resident in HBM.
Every call goes through libnqpu_hip.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import functools
from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

from . import _lib as L


@dataclass(frozen=True)
class Gate:
    """A gate as ``gates::Gate`` (sdk/rust/src/core/gates/gates.rs:632)."""

    kind: str
    targets: tuple = ()
    controls: tuple = ()
    params: tuple = ()
    matrix: object = None

    def qubits(self) -> tuple:
        return tuple(self.targets) + tuple(self.controls)


def _op(g: Gate) -> L.NqpuGateOp:
    op = L.NqpuGateOp()
    op.kind = L.GATE_ID[g.kind]
    op.num_targets = len(g.targets)
    op.num_controls = len(g.controls)
    for i, q in enumerate(g.targets):
        op.targets[i] = int(q)
    for i, q in enumerate(g.controls):
        op.controls[i] = int(q)
    for i, v in enumerate(g.params):
        op.params[i] = float(v)
    if g.matrix is not None:
        flat = np.asarray(g.matrix, dtype=np.complex128).reshape(-1)
        for i, v in enumerate(flat):
            op.matrix[i].re = float(v.real)
            op.matrix[i].im = float(v.imag)
    return op


@dataclass
class Circuit:
    """Ordered gate list on ``num_qubits`` qubits (LSB order: qubit q = bit q)."""

    num_qubits: int
    gates: list = field(default_factory=list)

    def add(self, kind, targets=(), controls=(), params=(), matrix=None) -> "Circuit":
        self.gates.append(Gate(kind, tuple(targets), tuple(controls), tuple(params), matrix))
        return self

    # constructors in the reference's argument order (gates.rs:669-760)
    def h(self, q): return self.add("h", (q,))
    def x(self, q): return self.add("x", (q,))
    def y(self, q): return self.add("y", (q,))
    def z(self, q): return self.add("z", (q,))
    def s(self, q): return self.add("s", (q,))
    def t(self, q): return self.add("t", (q,))
    def sx(self, q): return self.add("sx", (q,))
    def rx(self, q, a): return self.add("rx", (q,), (), (a,))
    def ry(self, q, a): return self.add("ry", (q,), (), (a,))
    def rz(self, q, a): return self.add("rz", (q,), (), (a,))
    def phase(self, q, a): return self.add("phase", (q,), (), (a,))
    def u(self, q, theta, phi, lam): return self.add("u", (q,), (), (theta, phi, lam))
    def cnot(self, c, t): return self.add("cnot", (t,), (c,))
    cx = cnot
    def cz(self, c, t): return self.add("cz", (t,), (c,))
    def cr(self, c, t, a): return self.add("cr", (t,), (c,), (a,))
    cp = cr
    def crx(self, c, t, a): return self.add("crx", (t,), (c,), (a,))
    def cry(self, c, t, a): return self.add("cry", (t,), (c,), (a,))
    def crz(self, c, t, a): return self.add("crz", (t,), (c,), (a,))
    def swap(self, a, b): return self.add("swap", (a, b))
    def iswap(self, a, b): return self.add("iswap", (a, b))
    def toffoli(self, c1, c2, t): return self.add("toffoli", (t,), (c1, c2))
    def ccz(self, c1, c2, t): return self.add("ccz", (t,), (c1, c2))
    def rxx(self, a, b, th): return self.add("rxx", (a, b), (), (th,))
    def ryy(self, a, b, th): return self.add("ryy", (a, b), (), (th,))
    def rzz(self, a, b, th): return self.add("rzz", (a, b), (), (th,))
    def cswap(self, c, a, b): return self.add("cswap", (a, b), (c,))
    def cu(self, c, t, theta, phi, lam, gamma): return self.add("cu", (t,), (c,), (theta, phi, lam, gamma))
    def custom1(self, q, m): return self.add("custom1", (q,), (), (), np.asarray(m, np.complex128))
    def custom2(self, q_hi, q_lo, m): return self.add("custom2", (q_hi, q_lo), (), (), np.asarray(m, np.complex128))

    def extend(self, gates: Iterable[Gate]) -> "Circuit":
        self.gates.extend(gates)
        return self

    def native(self) -> "NativeCircuit":
        return NativeCircuit(self)


class NativeCircuit:
    """A circuit uploaded into the library's circuit IR."""

    @classmethod
    def _adopt(cls, ptr, num_qubits: int) -> "NativeCircuit":
        nc = cls.__new__(cls)
        nc.ptr = ptr
        nc.num_qubits = num_qubits
        return nc

    def __init__(self, circuit: Circuit):
        lib = L.load()
        self.num_qubits = circuit.num_qubits
        self.ptr = lib.nqpu_circuit_create_ex(circuit.num_qubits)
        if not self.ptr:
            raise ValueError(f"invalid qubit count {circuit.num_qubits}")
        ops = (L.NqpuGateOp * max(1, len(circuit.gates)))()
        for i, g in enumerate(circuit.gates):
            ops[i] = _op(g)
        L.check(lib.nqpu_circuit_add_gates(self.ptr, ops, len(circuit.gates)), "nqpu_circuit_add_gates")

    def plan(self, precision="c64", fusion="tiled", semantics="standard", tile_qubits=0, low_qubits=0,
             lds_tiles=False, mfma=True, compile=False, qphase=True, fixed_low=False, untuned=False):
        lib = L.load()
        st = L.NqpuRunStats()
        o = run_options(fusion, semantics, tile_qubits, low_qubits, lds_tiles=lds_tiles, mfma=mfma,
                        compile=compile, qphase=qphase, fixed_low=fixed_low, untuned=untuned)
        L.check(lib.nqpu_circuit_plan(self.ptr, ct.byref(o), _prec(precision), ct.byref(st)), "plan")
        return _stats(st)

    def emulate(self, amps, precision="c64", semantics="standard", tile_qubits=0, low_qubits=0,
                lds_tiles=False, batch_layout=False, mfma=True, qphase=True, fixed_low=False, compile=False):
        """Host emulation of the tiled GPU execution (plan, per-pass programs
        and kernel semantics) in FP64; returns the new amplitudes.  With
        batch_layout, emulates the layout the batched kernel uses."""
        a = np.ascontiguousarray(amps, dtype=np.complex128).copy()
        o = run_options("tiled", semantics, tile_qubits, low_qubits, lds_tiles=lds_tiles, mfma=mfma,
                        qphase=qphase, fixed_low=fixed_low, compile=compile)
        if batch_layout:
            o.flags |= L.RUN_BATCH_LAYOUT
        L.check(L.load().nqpu_debug_emulate(self.ptr, ct.byref(o), _prec(precision), a.ctypes.data, len(a)),
                "emulate")
        return a

    def check_compile(self, semantics="standard", mfma=True, qphase=True, precision="c64", tile_qubits=0):
        """Generate and compile (hipRTC, no device) the straight-line pass
        kernels a run with compile=True would use; raises with the compiler
        log on failure."""
        o = run_options("tiled", semantics, tile_qubits=tile_qubits, mfma=mfma, compile=True, qphase=qphase,
                        untuned=True)
        L.check(L.load().nqpu_debug_compile_passes_ex(self.ptr, ct.byref(o), _prec(precision)), "compile passes")

    def fusion_counts(self):
        lib = L.load()
        fc = L.NqpuFusionCounts()
        L.check(lib.nqpu_circuit_fusion_counts(self.ptr, ct.byref(fc)), "fusion_counts")
        return {"original_count": fc.original_count, "fused_count": fc.fused_count,
                "gates_eliminated": fc.gates_eliminated, "items": fc.items}

    def close(self):
        if self.ptr:
            L.load().nqpu_circuit_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_FUSION = {"none": L.FUSION_NONE, "reference": L.FUSION_REFERENCE, "tiled": L.FUSION_TILED}
_SEMANTICS = {"standard": L.SEMANTICS_STANDARD, "reference": L.SEMANTICS_REFERENCE}


def run_options(fusion="tiled", semantics="standard", tile_qubits=0, low_qubits=0,
                time_passes=False, lds_tiles=False, mfma=True, compile=False, qphase=True,
                fixed_low=False, untuned=False, async_=False, from_zero=False) -> L.NqpuRunOptions:
    """mfma=False keeps every gate of complex64 register tiles on the vector
    ALUs (no fused blocks on the matrix cores); compile=True runs complex64
    register tile passes as straight-line kernels compiled at run time;
    qphase=False applies runs of diagonal gates one by one instead of as one
    quadratic phase per run; fixed_low=True keeps the state's lowest qubits
    in every complex64 register tile."""
    o = L.NqpuRunOptions()
    o.fusion = _FUSION[fusion]
    o.semantics = _SEMANTICS[semantics]
    o.tile_qubits = tile_qubits
    o.low_qubits = low_qubits
    o.flags = ((L.RUN_TIME_PASSES if time_passes else 0) | (L.RUN_LDS_TILES if lds_tiles else 0)
               | (0 if mfma else L.RUN_NO_MFMA) | (L.RUN_COMPILE if compile else 0)
               | (0 if qphase else L.RUN_NO_QPHASE) | (L.RUN_FIXED_LOW if fixed_low else 0)
               | (L.RUN_UNTUNED if untuned else 0) | (L.RUN_ASYNC if async_ else 0)
               | (L.RUN_FROM_ZERO if from_zero else 0))
    return o


def _stats(st: L.NqpuRunStats) -> dict:
    return {"gates": st.gates, "device_ops": st.device_ops, "passes": st.passes, "bytes": st.bytes_moved,
            "pass_ms": st.pass_ms, "timed_passes": st.timed_passes, "remaps": st.remaps,
            "mfma_blocks": st.mfma_blocks, "mfma_gates": st.mfma_gates,
            "exchanges": st.exchanges, "qphase_gates": st.qphase_gates}


def _prec(p) -> int:
    if p in ("c64", "complex64", L.PRECISION_C64):
        return L.PRECISION_C64
    if p in ("c128", "complex128", L.PRECISION_C128):
        return L.PRECISION_C128
    raise ValueError(f"unknown precision {p!r}")


def pauli_masks(pauli: str | Sequence[str]) -> tuple[int, int]:
    """Pauli string (character q acts on qubit q, as lib.rs:465) -> (xmask, zmask)."""
    if isinstance(pauli, str):
        return _pauli_masks_str(pauli)
    return _pauli_masks(pauli)


@functools.lru_cache(maxsize=1 << 16)
def _pauli_masks_str(pauli: str) -> tuple[int, int]:
    # strings repeat across the circuits of a batch (the same Hamiltonian terms)
    return _pauli_masks(pauli)


def _pauli_masks(pauli) -> tuple[int, int]:
    x = z = 0
    for q, c in enumerate(pauli):
        c = c.upper()
        if c == "X":
            x |= 1 << q
        elif c == "Y":
            x |= 1 << q
            z |= 1 << q
        elif c == "Z":
            z |= 1 << q
        elif c != "I":
            raise ValueError(f"unknown Pauli operator {c!r}")
    return x, z


def device_count() -> int:
    return int(L.load().nqpu_device_count())


def set_seed(seed: int) -> None:
    L.load().nqpu_set_seed(seed)


class StateVector:
    """A 2^n-amplitude state resident in GPU memory."""

    def __init__(self, num_qubits: int, precision="c128"):
        lib = L.load()
        self.num_qubits = num_qubits
        self.precision = _prec(precision)
        self.ptr = lib.nqpu_state_create_ex(num_qubits, self.precision)
        if not self.ptr:
            detail = (lib.nqpu_last_error_detail() or b"").decode()
            code = 9 if "no usable" in detail else (8 if "does not fit" in detail else 7)
            raise L.NqpuError(code, f"nqpu_state_create_ex({num_qubits}, {precision})")

    @property
    def dim(self) -> int:
        return 1 << self.num_qubits

    def close(self):
        if getattr(self, "ptr", None):
            L.load().nqpu_state_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- execution
    def run(self, circuit: Circuit | NativeCircuit, fusion="tiled", semantics="standard",
            tile_qubits=0, low_qubits=0, time_passes=False, lds_tiles=False, mfma=True, compile=False,
            qphase=True, fixed_low=False, untuned=False, async_=False, from_zero=False) -> dict:
        """async_=True returns once the passes are queued (call synchronize()
        before reading pass_ms); the state's later calls are ordered after it.
        from_zero=True runs the circuit from |0...0> (reset() and run() in
        one, without the separate fill pass)."""
        lib = L.load()
        nc = circuit if isinstance(circuit, NativeCircuit) else circuit.native()
        o = run_options(fusion, semantics, tile_qubits, low_qubits, time_passes, lds_tiles, mfma, compile, qphase,
                        fixed_low, untuned, async_, from_zero)
        L.check(lib.nqpu_run(nc.ptr, self.ptr, ct.byref(o)), "nqpu_run")
        return self.last_stats()

    def synchronize(self) -> dict:
        """Wait for runs queued with async_=True; returns the statistics
        (pass_ms covering every run since the last synchronisation)."""
        L.check(L.load().nqpu_state_synchronize(self.ptr), "synchronize")
        return self.last_stats()

    def last_stats(self) -> dict:
        st = L.NqpuRunStats()
        L.check(L.load().nqpu_last_run_stats(self.ptr, ct.byref(st)), "stats")
        return _stats(st)

    def reset(self):
        L.check(L.load().nqpu_state_reset(self.ptr), "reset")

    # --- data
    def amplitudes(self) -> np.ndarray:
        out = np.empty(self.dim, dtype=np.complex128)
        L.check(L.load().nqpu_state_get_amplitudes(self.ptr, out.ctypes.data, self.dim), "get_amplitudes")
        return out

    def set_amplitudes(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        L.check(L.load().nqpu_state_set_amplitudes(self.ptr, a.ctypes.data, len(a)), "set_amplitudes")

    def probabilities(self) -> np.ndarray:
        out = np.empty(self.dim, dtype=np.float64)
        L.check(L.load().nqpu_state_probabilities(self.ptr, out.ctypes.data, self.dim), "probabilities")
        return out

    def norm(self) -> float:
        return float(L.load().nqpu_state_norm(self.ptr))

    def normalize(self):
        L.check(L.load().nqpu_state_normalize(self.ptr), "normalize")

    def copy_from(self, other: "StateVector") -> None:
        """Device-side copy of another state of the same size and precision."""
        L.check(L.load().nqpu_state_copy(self.ptr, other.ptr), "copy")

    def overlap(self, other: "StateVector") -> complex:
        """<self|other> in FP64 on the device."""
        v = L.NqpuComplex()
        L.check(L.load().nqpu_state_overlap(self.ptr, other.ptr, ct.byref(v)), "overlap")
        return complex(v.re, v.im)

    def fidelity(self, other: "StateVector") -> float:
        """|<self|other>|^2 (QuantumState::fidelity, lib.rs:388-403)."""
        return abs(self.overlap(other)) ** 2

    def expectation_z(self, q: int) -> float:
        v = ct.c_double()
        L.check(L.load().nqpu_expectation_z(self.ptr, q, ct.byref(v)), "expectation_z")
        return v.value

    def expectation_z_all(self) -> np.ndarray:
        out = np.empty(self.num_qubits, dtype=np.float64)
        L.check(L.load().nqpu_expectation_z_all(self.ptr, out.ctypes.data, self.num_qubits), "z_all")
        return out

    def expectation_pauli(self, pauli=None, xmask=None, zmask=None) -> float:
        if pauli is not None:
            xmask, zmask = pauli_masks(pauli)
        v = ct.c_double()
        L.check(L.load().nqpu_expectation_pauli(self.ptr, int(xmask), int(zmask), ct.byref(v)), "pauli")
        return v.value

    def expectation_pauli_sum(self, terms) -> float:
        """terms: [(coeff, pauli string)] as expectation_hamiltonian (lib.rs:518)."""
        xs, zs, cs = [], [], []
        for c, p in terms:
            x, z = pauli_masks(p)
            xs.append(x); zs.append(z); cs.append(float(c))
        xa = np.array(xs or [0], dtype=np.uint64)
        za = np.array(zs or [0], dtype=np.uint64)
        ca = np.array(cs or [0.0], dtype=np.float64)
        v = ct.c_double()
        L.check(L.load().nqpu_expectation_pauli_sum(self.ptr, xa.ctypes.data, za.ctypes.data, ca.ctypes.data,
                                                    len(xs), ct.byref(v)), "pauli_sum")
        return v.value

    def expectation_pauli_values(self, paulis) -> np.ndarray:
        """<P> of every string (or (xmask, zmask) pair) in one grouped
        evaluation: strings with equal X/Y masks share a pass."""
        xs, zs = [], []
        for p in paulis:
            x, z = pauli_masks(p) if isinstance(p, str) else (int(p[0]), int(p[1]))
            xs.append(x)
            zs.append(z)
        xa = np.array(xs or [0], dtype=np.uint64)
        za = np.array(zs or [0], dtype=np.uint64)
        out = np.zeros(max(1, len(xs)), dtype=np.float64)
        L.check(L.load().nqpu_expectation_pauli_values(self.ptr, xa.ctypes.data, za.ctypes.data, len(xs),
                                                       out.ctypes.data), "pauli_values")
        return out[: len(xs)]

    def expectation_pauli_string(self, paulis: str, order="lsb") -> float:
        """<P> for a string of I/X/Y/Z.  order="lsb": character i acts on
        qubit i (lib.rs:465); order="msb": character i acts on the
        Estimator's qubit i, index bit n-1-i (primitives.rs:111-133, 683)."""
        o = {"lsb": L.ORDER_LSB, "msb": L.ORDER_MSB}[order]
        v = ct.c_double()
        L.check(L.load().nqpu_expectation_pauli_string(self.ptr, paulis.encode(), o, ct.byref(v)), "pauli_string")
        return v.value

    def estimator_expectation(self, terms) -> float:
        """Exact Estimator value of an observable (primitives.rs:1070-1083):
        terms = [(coeff, [(qubit, 'X'|'Y'|'Z'|'I'), ...]), ...] with qubit 0
        the most significant index bit (primitives.rs:683-728)."""
        qubits, paulis, offsets, coeffs = [], [], [0], []
        for c, factors in terms:
            for q, p in factors:
                qubits.append(int(q))
                paulis.append(str(p))
            offsets.append(len(qubits))
            coeffs.append(float(c))
        qa = np.array(qubits or [0], dtype=np.uint32)
        oa = np.array(offsets, dtype=np.uint64)
        ca = np.array(coeffs or [0.0], dtype=np.float64)
        v = ct.c_double()
        L.check(L.load().nqpu_estimator_expectation(self.ptr, qa.ctypes.data, "".join(paulis).encode(), oa.ctypes.data,
                                                    ca.ctypes.data, len(coeffs), ct.byref(v)), "estimator")
        return v.value

    def reset_qubit(self, q: int) -> None:
        """reset_qubit (lib.rs:1705-1730) on the device."""
        L.check(L.load().nqpu_state_reset_qubit(self.ptr, q), "reset_qubit")

    def initialize(self, amplitudes) -> bool:
        """initialize_from_amplitudes (lib.rs:1918-1935): False (state
        unchanged) when the length is not 2^n or the norm is below 1e-10."""
        a = np.ascontiguousarray(amplitudes, dtype=np.complex128)
        code = L.load().nqpu_state_initialize(self.ptr, a.ctypes.data, len(a))
        if code == L.NQPU_OK:
            return True
        if code == 2:  # INVALID_PARAMETER
            return False
        L.check(code, "initialize")
        return False

    def sample(self, shots: int, seed: int | None = None) -> np.ndarray:
        lib = L.load()
        if seed is not None:
            lib.nqpu_set_seed(seed)
        out = np.empty(max(shots, 1), dtype=np.uint64)
        L.check(lib.nqpu_sample(self.ptr, shots, out.ctypes.data), "sample")
        return out[:shots]

    def sample_counts(self, shots: int, seed: int | None = None) -> dict:
        """MSB-first bitstring counts (sample_bitstrings, lib.rs:553-564)."""
        idx = self.sample(shots, seed)
        vals, cnt = np.unique(idx, return_counts=True)
        return {format(int(v), f"0{self.num_qubits}b"): int(c) for v, c in zip(vals, cnt)}

    def measure(self, q: int) -> tuple[int, float]:
        m = L.NqpuMeasurement()
        L.check(L.load().nqpu_measure(self.ptr, q, ct.byref(m)), "measure")
        return int(m.outcome), float(m.probability)


class RunGraph:
    """A tiled run of `circuit` on `state` recorded once as a HIP graph
    (nqpu_graph_create) and replayed with launch(); update(circuit) swaps in
    a circuit of the same structure (new angles) by copying its register
    programs into the graph's device buffer, and records the graph again only
    when the new plan launches other kernels.  from_zero=True makes every
    replay start from |0...0>.  The state must outlive the graph."""

    def __init__(self, state: "StateVector", circuit: Circuit | NativeCircuit, from_zero=True,
                 semantics="standard", tile_qubits=0, low_qubits=0, lds_tiles=False, mfma=True, compile=False,
                 qphase=True, fixed_low=False, untuned=False):
        lib = L.load()
        self.state = state
        nc = circuit if isinstance(circuit, NativeCircuit) else circuit.native()
        o = run_options("tiled", semantics, tile_qubits, low_qubits, False, lds_tiles, mfma, compile, qphase,
                        fixed_low, untuned)
        p = ct.c_void_p()
        L.check(lib.nqpu_graph_create(nc.ptr, state.ptr, ct.byref(o), L.GRAPH_FROM_ZERO if from_zero else 0,
                                      ct.byref(p)), "nqpu_graph_create")
        self.ptr = p.value

    def update(self, circuit: Circuit | NativeCircuit) -> None:
        nc = circuit if isinstance(circuit, NativeCircuit) else circuit.native()
        L.check(L.load().nqpu_graph_update(self.ptr, nc.ptr), "nqpu_graph_update")

    def launch(self, async_=False) -> None:
        L.check(L.load().nqpu_graph_launch(self.ptr, L.RUN_ASYNC if async_ else 0), "nqpu_graph_launch")

    def info(self) -> dict:
        st = L.NqpuGraphInfo()
        L.check(L.load().nqpu_graph_info(self.ptr, ct.byref(st)), "nqpu_graph_info")
        return {k: getattr(st, k) for k, _ in L.NqpuGraphInfo._fields_}

    def close(self):
        if getattr(self, "ptr", None):
            L.load().nqpu_graph_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _terms(obs):
    """[(coeff, pauli string)] or [(coeff, xmask, zmask)] -> [(x, z, coeff)]."""
    out = []
    for t in obs:
        if len(t) == 2:
            x, z = pauli_masks(t[1])
            out.append((x, z, float(t[0])))
        else:
            out.append((int(t[1]), int(t[2]), float(t[0])))
    return out


_TERM_DTYPE = np.dtype([("xmask", np.uint64), ("zmask", np.uint64), ("coeff", np.float64)])


def _term_table(observables):
    """All circuits' terms as one NqpuPauliTerm array (numpy, same layout) and
    the per-circuit offsets."""
    terms, offsets = [], [0]
    for obs in observables:
        terms.extend(_terms(obs))
        offsets.append(len(terms))
    table = np.array(terms if terms else [(0, 0, 0.0)], dtype=_TERM_DTYPE)
    return table, np.array(offsets, dtype=np.uint64)


class Observables:
    """The Pauli sums of a batch in the library's term-table form
    (NqpuPauliTerm rows and per-circuit offsets), built once and reused by
    every Batch.expectation call that evaluates the same observables (the
    observables' counterpart of NativeCircuit)."""

    def __init__(self, observables):
        self.table, self.offsets = _term_table(observables)
        self.count = len(self.offsets) - 1


class Batch:
    """Many complex64 circuits in one persistent GPU launch, each returning
    the FP64 expectation of its own Pauli sum (the batched counterpart of
    the reference's serial vmap_simulate, jax_bridge.rs:447)."""

    def __init__(self):
        lib = L.load()
        self._pending = []
        self.ptr = lib.nqpu_batch_create()
        if not self.ptr:
            raise L.NqpuError(9, "nqpu_batch_create")

    def expectation(self, circuits, observables, semantics="standard", async_=False) -> np.ndarray:
        """async_=True returns once the batch is queued: the returned array is
        filled when synchronize() returns (the host plans the next batch
        meanwhile)."""
        lib = L.load()
        natives = [c if isinstance(c, NativeCircuit) else c.native() for c in circuits]
        arr = (ct.c_void_p * max(1, len(natives)))(*[nc.ptr for nc in natives])
        if not isinstance(observables, Observables):
            observables = Observables(observables)
        if observables.count != len(natives):
            raise ValueError(f"{observables.count} observables for {len(natives)} circuits")
        table, off = observables.table, observables.offsets
        out = np.empty(max(1, len(natives)), dtype=np.float64)
        o = run_options("tiled", semantics, async_=async_)
        L.check(lib.nqpu_batch_expectation(self.ptr, arr, len(natives), table.ctypes.data, off.ctypes.data,
                                           ct.byref(o), out.ctypes.data), "nqpu_batch_expectation")
        if async_:
            self._pending.append(out)  # written by the library until synchronize()
            self._pending = self._pending[-2:]
        return out[: len(natives)]

    def synchronize(self) -> None:
        L.check(L.load().nqpu_batch_synchronize(self.ptr), "nqpu_batch_synchronize")
        self._pending = []

    def last_stats(self) -> dict:
        st = L.NqpuBatchStats()
        L.check(L.load().nqpu_batch_last_stats(self.ptr, ct.byref(st)), "batch stats")
        return {k: getattr(st, k) for k, _ in L.NqpuBatchStats._fields_}

    @staticmethod
    def plan(circuits, observables, semantics="standard") -> dict:
        """The host half of expectation() only (no device): the batch's pass
        counts, planning time and the number of register-program ops."""
        lib = L.load()
        natives = [c if isinstance(c, NativeCircuit) else c.native() for c in circuits]
        arr = (ct.c_void_p * max(1, len(natives)))(*[nc.ptr for nc in natives])
        table, off = _term_table(observables)
        st = L.NqpuBatchStats()
        nops = ct.c_uint64()
        o = run_options("tiled", semantics)
        L.check(lib.nqpu_debug_batch_plan(arr, len(natives), table.ctypes.data, off.ctypes.data, ct.byref(o),
                                          ct.byref(st), ct.byref(nops)), "nqpu_debug_batch_plan")
        out = {k: getattr(st, k) for k, _ in L.NqpuBatchStats._fields_}
        out["program_ops"] = nops.value
        return out

    def close(self):
        if getattr(self, "ptr", None):
            L.load().nqpu_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request server on the GPU (nqpu_server_*): each
    request runs a circuit from |0...0>, returns coeff * <P> for its Pauli
    terms and `shots` sampled basis indices.  Requests are admitted in
    arrival order as slots and arena memory free up, and run concurrently;
    a request's results are the same whatever else is running (serving path
    replaced: sdk/rust/src/web/mod.rs:288-419)."""

    def __init__(self, slots=8, precision="c64", memory_bytes=0, fusion="tiled", semantics="standard",
                 tile_qubits=0, low_qubits=0):
        lib = L.load()
        cfg = L.NqpuServerConfig()
        cfg.slots = slots
        cfg.precision = _prec(precision)
        cfg.memory_bytes = int(memory_bytes)
        cfg.run = run_options(fusion, semantics, tile_qubits, low_qubits)
        self.ptr = lib.nqpu_server_create(ct.byref(cfg))
        if not self.ptr:
            detail = (lib.nqpu_last_error_detail() or b"").decode()
            raise L.NqpuError(9 if "no usable" in detail else 8, "nqpu_server_create")
        self._shape = {}

    def submit(self, circuit, terms=(), shots=0, seed=0) -> int:
        nc = circuit if isinstance(circuit, NativeCircuit) else circuit.native()
        ts = _terms(terms)
        tarr = (L.NqpuPauliTerm * max(1, len(ts)))()
        for i, (x, z, c) in enumerate(ts):
            tarr[i].xmask, tarr[i].zmask, tarr[i].coeff = x, z, c
        rid = ct.c_uint64()
        L.check(L.load().nqpu_server_submit(self.ptr, nc.ptr, tarr, len(ts), int(shots), int(seed),
                                            ct.byref(rid)), "nqpu_server_submit")
        self._shape[rid.value] = (len(ts), int(shots))
        return rid.value

    @staticmethod
    def _status(st: L.NqpuRequestStatus) -> dict:
        return {"state": L.REQUEST_STATES[st.state], "error": st.error, "queued_ms": st.queued_ms,
                "run_ms": st.run_ms, "latency_ms": st.latency_ms}

    def poll(self, rid: int) -> dict:
        st = L.NqpuRequestStatus()
        L.check(L.load().nqpu_server_poll(self.ptr, rid, ct.byref(st)), "nqpu_server_poll")
        return self._status(st)

    def wait(self, rid: int, timeout: float | None = None) -> dict:
        st = L.NqpuRequestStatus()
        ms = L.WAIT_FOREVER if timeout is None else max(0, int(timeout * 1000))
        L.check(L.load().nqpu_server_wait(self.ptr, rid, ms, ct.byref(st)), "nqpu_server_wait")
        return self._status(st)

    def take(self, rid: int):
        """(expectations, samples) of a finished request; forgets it."""
        nterms, shots = self._shape.pop(rid)
        e = np.zeros(max(1, nterms), dtype=np.float64)
        s = np.zeros(max(1, shots), dtype=np.uint64)
        L.check(L.load().nqpu_server_take(self.ptr, rid, e.ctypes.data, nterms, s.ctypes.data, shots),
                "nqpu_server_take")
        return e[:nterms], s[:shots]

    def result(self, rid: int, timeout: float | None = None) -> dict:
        st = self.wait(rid, timeout)
        e, s = self.take(rid)
        st.update(expectations=e, samples=s)
        return st

    def submit_json(self, text: str, seed: int = 0) -> int:
        """Queue a SampleRequest / SerializableCircuit request given as JSON
        (it needs "shots" or "observables")."""
        rid = ct.c_uint64()
        L.check(L.load().nqpu_server_submit_json(self.ptr, text.encode(), int(seed), ct.byref(rid)),
                "nqpu_server_submit_json")
        return rid.value

    def take_json(self, rid: int, timeout: float | None = None) -> dict:
        """Wait for a JSON request and return its SampleResponse-shaped result."""
        import json as _json
        lib = L.load()
        ms = L.WAIT_FOREVER if timeout is None else max(0, int(timeout * 1000))
        need = ct.c_uint64()
        size = 1 << 16
        while True:
            buf = ct.create_string_buffer(size)
            code = lib.nqpu_server_take_json(self.ptr, rid, ms, buf, size, ct.byref(need))
            if code == L.NQPU_OK:
                return _json.loads(buf.value.decode())
            if code == 2 and need.value > size:
                size = int(need.value)
                continue
            L.check(code, "nqpu_server_take_json")

    def stats(self) -> dict:
        st = L.NqpuServerStats()
        L.check(L.load().nqpu_server_stats(self.ptr, ct.byref(st)), "nqpu_server_stats")
        return {k: getattr(st, k) for k, _ in L.NqpuServerStats._fields_}

    def close(self):
        if getattr(self, "ptr", None):
            L.load().nqpu_server_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _tuned(t: L.NqpuTunedPlan) -> dict:
    return {"tile_qubits": t.tile_qubits, "low_qubits": t.low_qubits,
            "compile": bool(t.flags & L.RUN_COMPILE), "fixed_low": bool(t.flags & L.RUN_FIXED_LOW),
            "tiles_per_group": t.tiles_per_group, "ms": t.ms, "default_ms": t.default_ms}


def autotune(num_qubits: int, precision="c64", probe=None) -> dict:
    """Time the engine's plan choices for an n-qubit state on this device
    and make the fastest the default for (n, precision) in this process."""
    t = L.NqpuTunedPlan()
    nc = None if probe is None else (probe if isinstance(probe, NativeCircuit) else probe.native())
    L.check(L.load().nqpu_autotune(num_qubits, _prec(precision), nc.ptr if nc else None, ct.byref(t)),
            "nqpu_autotune")
    return _tuned(t)


def tuned_plan(num_qubits: int, precision="c64"):
    t = L.NqpuTunedPlan()
    code = L.load().nqpu_tuned_plan(num_qubits, _prec(precision), ct.byref(t))
    return _tuned(t) if code == L.NQPU_OK else None


def tune_load(path: str) -> None:
    L.check(L.load().nqpu_tune_load(path.encode()), "nqpu_tune_load")


def tune_reset() -> None:
    L.load().nqpu_tune_reset()


def circuit_from_json(text: str):
    """Parse a request in the web API's CircuitRequest/SampleRequest form
    (web/mod.rs:34-160) or circuit_serde's SerializableCircuit form
    (circuit_serde.rs:329-340): (NativeCircuit, shots)."""
    import json as _json
    lib = L.load()
    ptr = ct.c_void_p()
    shots = ct.c_uint32()
    L.check(lib.nqpu_circuit_from_json(text.encode(), ct.byref(ptr), ct.byref(shots)), "circuit_from_json")
    n = int(_json.loads(text)["num_qubits"])
    return NativeCircuit._adopt(ptr.value, n), int(shots.value)


def execute_json(text: str, precision="c128") -> dict:
    """ExecuteResponse for a CircuitRequest (web/mod.rs:288-375): the run's
    probabilities, time, backend, gate count and fidelity against the
    one-pass-per-gate run."""
    import json as _json
    lib = L.load()
    need = ct.c_uint64()
    code = lib.nqpu_execute_json(text.encode(), _prec(precision), None, 0, ct.byref(need))
    if code not in (L.NQPU_OK, 2):
        L.check(code, "execute_json")
    buf = ct.create_string_buffer(int(need.value))
    L.check(lib.nqpu_execute_json(text.encode(), _prec(precision), buf, len(buf), ct.byref(need)), "execute_json")
    return _json.loads(buf.value.decode())


def simulate(circuit: Circuit, precision="c128", **run_kw) -> StateVector:
    sv = StateVector(circuit.num_qubits, precision)
    sv.run(circuit, **run_kw)
    return sv
