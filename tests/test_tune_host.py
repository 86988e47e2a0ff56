"""The plan autotuner's host side (no device): tuned plans read from a cache
file become the default plan of runs that leave the plan settings open,
NQPU_RUN_UNTUNED keeps the fixed default, and nqpu_tune_reset forgets them.
Measuring needs a device (NQPU_UNAVAILABLE here)."""
import pytest

from nqpu_amd import _lib as L
from nqpu_amd import engine as E
from nqpu_amd.workloads import random_u3_brickwork


@pytest.fixture(autouse=True)
def clean():
    E.tune_reset()
    yield
    E.tune_reset()


def test_loaded_plan_is_the_default(tmp_path):
    nc = random_u3_brickwork(20, 10, 3).native()
    fixed = nc.plan("c64")["passes"]
    t10 = nc.plan("c64", tile_qubits=10)["passes"]
    assert t10 != fixed
    f = tmp_path / "tune.txt"
    f.write_text("20 1 10 5 0 1 1.5 2.5\n19 2 12 4 0 2 3.0 3.5\nbogus line\n")
    E.tune_load(str(f))
    assert E.tuned_plan(20, "c64") == {"tile_qubits": 10, "low_qubits": 5, "compile": False, "fixed_low": False,
                                       "tiles_per_group": 1, "ms": 1.5, "default_ms": 2.5}
    assert E.tuned_plan(19, "c128")["tiles_per_group"] == 2
    assert E.tuned_plan(21, "c64") is None
    assert nc.plan("c64")["passes"] == t10                 # the tuned plan
    assert nc.plan("c64", untuned=True)["passes"] == fixed  # the fixed default
    assert nc.plan("c64", tile_qubits=12)["passes"] == nc.plan("c64", tile_qubits=12, untuned=True)["passes"]
    E.tune_reset()
    assert E.tuned_plan(20, "c64") is None and nc.plan("c64")["passes"] == fixed


def test_later_lines_win_and_bad_files(tmp_path):
    f = tmp_path / "tune.txt"
    f.write_text("20 1 10 5 0 1 1 2\n20 1 12 5 16 1 0.5 2\n20 1 99 5 0 1 0.1 2\n")
    E.tune_load(str(f))
    t = E.tuned_plan(20, "c64")
    assert t["tile_qubits"] == 12 and t["compile"]
    with pytest.raises(L.NqpuError):
        E.tune_load(str(tmp_path / "missing.txt"))


def test_measuring_needs_a_device():
    if E.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(L.NqpuError) as ei:
        E.autotune(16, "c64")
    assert ei.value.code == 9
