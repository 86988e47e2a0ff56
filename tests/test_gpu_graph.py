"""Run graphs (nqpu_graph_*): a tiled run recorded once as a HIP graph and
replayed.  A replay launches the same kernels in the same order as nqpu_run
on the same plan, so its result must equal the eager run's bit for bit; the
complex128 replays are also checked against the FP64 oracle at 1e-12.  Each
test replays a graph several times with the program's own eager work and
synchronisations in between and checks every replay."""
import ctypes as ct
import math

import numpy as np
import pytest

from nqpu_amd import _lib as L
from nqpu_amd import engine as E
from nqpu_amd.engine import Circuit, RunGraph, StateVector
from nqpu_amd.workloads import random_u3_brickwork

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if E.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X")


def ansatz(n, layers, seed):
    """RY.RZ on every qubit per layer, then a CZ ladder (same structure for
    every seed)."""
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for _ in range(layers):
        for q in range(n):
            c.ry(q, float(rng.uniform(-math.pi, math.pi)))
            c.rz(q, float(rng.uniform(-math.pi, math.pi)))
        for q in range(n - 1):
            c.cz(q, q + 1)
    return c


# (qubits, precision, run options): LDS tile passes (8 qubits), interpreted
# register passes (complex64 11-qubit tiles, complex128), compiled passes
CASES = [
    (8, "c128", {}),
    (8, "c64", {}),
    (14, "c64", {}),
    (14, "c64", {"compile": True}),
    (16, "c128", {}),
    (20, "c64", {"compile": True}),
    (22, "c128", {}),
]


@pytest.mark.parametrize("n,prec,kw", CASES, ids=[f"{n}q-{p}-{'compiled' if k else 'default'}" for n, p, k in CASES])
def test_replays_equal_the_eager_run(oracle, n, prec, kw):
    c = random_u3_brickwork(n, 12, 500 + n)
    c.rzz(0, n - 1, 0.37).cr(1, n - 2, 0.9).swap(2, n - 3)
    other = random_u3_brickwork(n, 3, 77)
    with StateVector(n, prec) as ref:
        ref.run(c, **kw)
        want = ref.amplitudes()
    if prec == "c128":
        exp = oracle.run(n, c.gates)
        assert np.max(np.abs(want - exp)) < 1e-12
    with StateVector(n, prec) as sv, RunGraph(sv, c, from_zero=True, **kw) as g:
        info = g.info()
        # from |0...0>: the first register pass generates the start state;
        # LDS tile passes (tiles under 11 qubits) get a fill kernel first
        fill = 1 if min(n, 11 if prec == "c64" else (10 if n <= 20 else 11)) < 11 else 0
        assert info["passes"] >= 1 and info["launches"] == info["passes"] + fill
        for k in range(3):
            # eager work on the same state between replays
            sv.run(other)
            assert abs(sv.norm() - 1.0) < 1e-5
            g.launch()
            got = sv.amplitudes()
            assert np.array_equal(got, want), f"replay {k}"
        assert g.info()["replays"] == 3


@pytest.mark.parametrize("prec,kw", [("c64", {"compile": True}), ("c64", {}), ("c128", {})])
def test_update_swaps_angles_in_place(oracle, prec, kw):
    n = 18
    c1, c2, c3 = ansatz(n, 4, 1), ansatz(n, 4, 2), ansatz(n, 4, 3)
    wants = []
    for c in (c1, c2, c3):
        with StateVector(n, prec) as ref:
            ref.run(c, **kw)
            wants.append(ref.amplitudes())
    with StateVector(n, prec) as sv, RunGraph(sv, c1, **kw) as g:
        g.launch()
        assert np.array_equal(sv.amplitudes(), wants[0])
        for i, c in enumerate((c2, c3, c1)):
            g.update(c)
            g.launch()
            assert np.array_equal(sv.amplitudes(), wants[(i + 1) % 3]), f"update {i}"
        info = g.info()
        assert info["updates"] == 3 and info["recaptures"] == 0
        if prec == "c128":
            exp = oracle.run(n, c1.gates)
            assert np.max(np.abs(sv.amplitudes() - exp)) < 1e-12


def test_update_with_another_structure_records_again(oracle):
    n = 16
    a = ansatz(n, 3, 5)
    b = random_u3_brickwork(n, 9, 6)
    with StateVector(n, "c128") as sv, RunGraph(sv, a) as g:
        g.launch()
        assert np.max(np.abs(sv.amplitudes() - oracle.run(n, a.gates))) < 1e-12
        g.update(b)
        g.launch()
        assert g.info()["recaptures"] == 1
        assert np.max(np.abs(sv.amplitudes() - oracle.run(n, b.gates))) < 1e-12
        g.update(a)
        g.launch()
        assert g.info()["recaptures"] == 2
        assert np.max(np.abs(sv.amplitudes() - oracle.run(n, a.gates))) < 1e-12


def test_queued_replays_apply_a_step_repeatedly(oracle):
    """Without the |0...0> fill a replay applies the circuit to the current
    state: five queued replays of one Trotter-like step, with an update queued
    between them, equal the eager runs in the same order."""
    n = 20
    step = Circuit(n)
    for q in range(n - 1):
        step.rzz(q, q + 1, 0.1)
    for q in range(n):
        step.rx(q, 0.2)
    step2 = Circuit(n)
    for q in range(n - 1):
        step2.rzz(q, q + 1, 0.15)
    for q in range(n):
        step2.rx(q, 0.25)
    for prec in ("c64", "c128"):
        with StateVector(n, prec) as ref:
            h = Circuit(n)
            for q in range(n):
                h.h(q)
            ref.run(h)
            for c in (step, step, step2, step2, step2):
                ref.run(c)
            want = ref.amplitudes()
        with StateVector(n, prec) as sv:
            sv.run(h)
            with RunGraph(sv, step, from_zero=False) as g:
                g.launch(async_=True)
                g.launch(async_=True)
                g.update(step2)
                for _ in range(3):
                    g.launch(async_=True)
                sv.synchronize()
                assert g.info()["recaptures"] == 0
                assert np.array_equal(sv.amplitudes(), want), prec
        if prec == "c128":
            exp = oracle.run(n, h.gates + 2 * step.gates + 3 * step2.gates)
            assert np.max(np.abs(want - exp)) < 1e-12


def test_graph_argument_errors():
    n = 10
    c = random_u3_brickwork(n, 2, 1)
    with StateVector(n, "c64") as sv:
        with pytest.raises(L.NqpuError) as e:
            RunGraph(sv, random_u3_brickwork(n + 1, 2, 1))
        assert e.value.code == 2
        lib = L.load()
        o = E.run_options("none")
        p = ct.c_void_p()
        nc = c.native()
        assert lib.nqpu_graph_create(nc.ptr, sv.ptr, ct.byref(o), 0, ct.byref(p)) == 2
        with RunGraph(sv, c) as g:
            assert lib.nqpu_graph_launch(g.ptr, 1 << 20) == 2
            with pytest.raises(L.NqpuError):
                g.update(random_u3_brickwork(n + 2, 2, 1))
            g.launch()  # still usable after a rejected update
