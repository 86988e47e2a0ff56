"""The plan autotuner on the GPU: every candidate plan it may choose gives
the oracle's state (complex128 within 1e-12, complex64 within 2e-5), a tuned
plan becomes the default and reproduces the oracle, and the choice is
appended to the NQPU_TUNE_CACHE file."""
import os

import numpy as np
import pytest

from nqpu_amd import engine as E
from nqpu_amd.engine import StateVector
from nqpu_amd.workloads import random_u3_brickwork

pytestmark = pytest.mark.gpu

# (tile, low, flags, tiles_per_group): the candidates of the tuner (tune.hip)
C64 = [(12, 5, 0, 1), (11, 5, 0, 1), (12, 5, 16, 1), (12, 5, 16 | 64, 1), (12, 5, 16, 2), (12, 6, 16, 1),
       (13, 5, 0, 1), (10, 5, 0, 1)]
C128 = [(11, 4, 0, 1), (10, 4, 0, 1), (11, 4, 16, 1), (12, 4, 16, 1), (11, 4, 0, 2), (11, 5, 0, 1),
        (12, 4, 0, 1)]


@pytest.fixture(autouse=True)
def clean():
    if E.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X")
    E.tune_reset()
    yield
    E.tune_reset()


@pytest.fixture(scope="module")
def case(oracle):
    n = 17
    c = random_u3_brickwork(n, 10, 61)
    return n, c, oracle.run(n, c.gates)


@pytest.mark.parametrize("prec,cands,tol", [("c64", C64, 2e-5), ("c128", C128, 1e-12)])
def test_every_candidate_plan_matches_oracle(tmp_path, case, prec, cands, tol):
    n, c, exp = case
    code = 1 if prec == "c64" else 2
    for tile, low, flags, tpg in cands:
        f = tmp_path / f"{prec}_{tile}_{low}_{flags}_{tpg}.txt"
        f.write_text(f"{n} {code} {tile} {low} {flags} {tpg} 1 1\n")
        E.tune_reset()
        E.tune_load(str(f))
        with StateVector(n, prec) as sv:
            st = sv.run(c)
            out = sv.amplitudes()
        assert np.max(np.abs(out - exp)) < tol, (tile, low, flags, tpg)


@pytest.mark.parametrize("prec,tol", [("c64", 2e-5), ("c128", 1e-12)])
def test_autotune_then_default_runs_match_oracle(tmp_path, case, prec, tol, monkeypatch):
    n, c, exp = case
    cache = tmp_path / "cache.txt"
    monkeypatch.setenv("NQPU_TUNE_CACHE", str(cache))
    t = E.autotune(n, prec, probe=c)
    assert t["ms"] > 0 and t["ms"] <= t["default_ms"] * 1.0001
    assert E.tuned_plan(n, prec) == t
    with StateVector(n, prec) as sv:
        sv.run(c)
        out = sv.amplitudes()
    assert np.max(np.abs(out - exp)) < tol
    line = cache.read_text().strip().split("\n")[-1].split()
    assert int(line[0]) == n and int(line[2]) == t["tile_qubits"] and int(line[5]) == t["tiles_per_group"]
