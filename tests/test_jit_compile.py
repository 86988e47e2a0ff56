"""The straight-line pass kernels of compile=True runs, on the CPU: the
generated HIP source of every register tile pass must compile with hipRTC for
gfx950 (no device needed; nothing is loaded).  Their results on the GPU are
checked against the oracle in test_gpu_parity.py."""
import pytest

from nqpu_amd.workloads import qft, tfim_trotter
from tests.test_planner_emulation import CIRCUITS


@pytest.mark.parametrize("tile", [12, 13])
@pytest.mark.parametrize("name", ["mixed16", "qft_trotter15"])
def test_generated_passes_compile(name, tile):
    # mixed16: every op kind, controls and diagonal factors outside the tile;
    # qft_trotter15: diagonal runs and matrix-core blocks with selector qubits
    CIRCUITS[name]().native().check_compile(tile_qubits=tile)


def test_reference_semantics_and_vector_only_passes_compile():
    c = tfim_trotter(13, 2, circuit=qft(13)).native()
    c.check_compile(semantics="reference")
    c.check_compile(mfma=False)


@pytest.mark.parametrize("tile", [11, 12])
def test_complex128_passes_compile(tile):
    """complex128 register tile passes (LDS remaps and vector-ALU gates) as
    straight-line kernels: every op kind, controls and diagonals outside the
    tile."""
    CIRCUITS["mixed16"]().native().check_compile(precision="c128", tile_qubits=tile)
