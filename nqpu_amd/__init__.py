"""nqpu_amd — MI355X (gfx950) statevector engine for the nQPU simulator interface.

The compute path is libnqpu_hip.so (hand-written HIP kernels behind a C ABI,
include/nqpu_hip.h); this package is its Python host side.
"""
from .engine import (Circuit, Gate, NativeCircuit, StateVector, device_count, pauli_masks,  # noqa: F401
                     set_seed, simulate)

__all__ = ["Circuit", "Gate", "NativeCircuit", "StateVector", "device_count", "pauli_masks", "set_seed",
           "simulate"]
