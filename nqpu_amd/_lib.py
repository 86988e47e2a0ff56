"""ctypes binding of libnqpu_hip.so (include/nqpu_hip.h).

The library must exist (``python -m nqpu_amd.build`` or
``__graft_entry__.build()``); there is no CPU fallback.  Loading fails loudly
if it is missing.
"""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnqpu_hip.so")

NQPU_OK = 0
ERROR_NAMES = {0: "OK", 1: "INVALID_QUBIT", 2: "INVALID_PARAMETER", 3: "ALLOCATION_FAILED",
               4: "SIMULATION_FAILED", 5: "NULL_POINTER", 6: "INVALID_HANDLE", 7: "DEVICE_ERROR",
               8: "OUT_OF_DEVICE_MEMORY", 9: "UNAVAILABLE"}

PRECISION_C64 = 1
PRECISION_C128 = 2
FUSION_NONE, FUSION_REFERENCE, FUSION_TILED = 0, 1, 2
SEMANTICS_STANDARD, SEMANTICS_REFERENCE = 0, 1

GATE_KINDS = ["h", "x", "y", "z", "s", "t", "rx", "ry", "rz", "u", "cnot", "cz", "swap",
              "toffoli", "crx", "cry", "crz", "cr", "sx", "phase", "iswap", "ccz", "rxx",
              "ryy", "rzz", "cswap", "cu", "custom1", "custom2"]
GATE_ID = {k: i for i, k in enumerate(GATE_KINDS)}


class NqpuComplex(ct.Structure):
    _fields_ = [("re", ct.c_double), ("im", ct.c_double)]


class NqpuMeasurement(ct.Structure):
    _fields_ = [("outcome", ct.c_uint32), ("probability", ct.c_double)]


class NqpuVersion(ct.Structure):
    _fields_ = [("major", ct.c_uint32), ("minor", ct.c_uint32), ("patch", ct.c_uint32)]


class NqpuGateOp(ct.Structure):
    _fields_ = [("kind", ct.c_uint32), ("num_targets", ct.c_uint32), ("num_controls", ct.c_uint32),
                ("targets", ct.c_uint32 * 3), ("controls", ct.c_uint32 * 3), ("reserved", ct.c_uint32),
                ("params", ct.c_double * 4), ("matrix", NqpuComplex * 16)]


class NqpuRunOptions(ct.Structure):
    _fields_ = [("fusion", ct.c_uint32), ("semantics", ct.c_uint32), ("tile_qubits", ct.c_uint32),
                ("low_qubits", ct.c_uint32), ("flags", ct.c_uint32)]


RUN_TIME_PASSES = 1
RUN_LDS_TILES = 2


class NqpuRunStats(ct.Structure):
    _fields_ = [("gates", ct.c_uint64), ("device_ops", ct.c_uint64), ("passes", ct.c_uint64),
                ("bytes_moved", ct.c_double), ("pass_ms", ct.c_double), ("timed_passes", ct.c_uint64),
                ("remaps", ct.c_uint64), ("mfma_blocks", ct.c_uint64), ("mfma_gates", ct.c_uint64),
                ("exchanges", ct.c_uint64), ("qphase_gates", ct.c_uint64)]


class NqpuPauliTerm(ct.Structure):
    _fields_ = [("xmask", ct.c_uint64), ("zmask", ct.c_uint64), ("coeff", ct.c_double)]


class NqpuBatchStats(ct.Structure):
    _fields_ = [("circuits", ct.c_uint64), ("resident", ct.c_uint64), ("gate_passes", ct.c_uint64),
                ("observable_passes", ct.c_uint64), ("plan_ms", ct.c_double), ("kernel_ms", ct.c_double),
                ("state_bytes", ct.c_double), ("hbm_bytes", ct.c_double)]


RUN_BATCH_LAYOUT = 4
RUN_NO_MFMA = 8
RUN_COMPILE = 16
RUN_NO_QPHASE = 32
RUN_FIXED_LOW = 64
RUN_UNTUNED = 128
RUN_ASYNC = 256
RUN_FROM_ZERO = 512


class NqpuFusionCounts(ct.Structure):
    _fields_ = [("original_count", ct.c_uint64), ("fused_count", ct.c_uint64),
                ("gates_eliminated", ct.c_uint64), ("items", ct.c_uint64)]


class NqpuServerConfig(ct.Structure):
    _fields_ = [("slots", ct.c_uint32), ("precision", ct.c_uint32), ("memory_bytes", ct.c_uint64),
                ("run", NqpuRunOptions)]


class NqpuRequestStatus(ct.Structure):
    _fields_ = [("state", ct.c_uint32), ("error", ct.c_int32), ("queued_ms", ct.c_double),
                ("run_ms", ct.c_double), ("latency_ms", ct.c_double)]


class NqpuServerStats(ct.Structure):
    _fields_ = [("submitted", ct.c_uint64), ("completed", ct.c_uint64), ("failed", ct.c_uint64),
                ("peak_running", ct.c_uint64), ("peak_bytes", ct.c_uint64), ("memory_bytes", ct.c_uint64)]


ORDER_LSB = 0
ORDER_MSB = 1

REQUEST_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}
WAIT_FOREVER = 0xFFFFFFFF

class NqpuTunedPlan(ct.Structure):
    _fields_ = [("tile_qubits", ct.c_uint32), ("low_qubits", ct.c_uint32), ("flags", ct.c_uint32),
                ("tiles_per_group", ct.c_uint32), ("ms", ct.c_double), ("default_ms", ct.c_double)]


GRAPH_FROM_ZERO = 1


class NqpuGraphInfo(ct.Structure):
    _fields_ = [("launches", ct.c_uint64), ("passes", ct.c_uint64), ("replays", ct.c_uint64),
                ("updates", ct.c_uint64), ("recaptures", ct.c_uint64), ("image_bytes", ct.c_uint64),
                ("plan_ms", ct.c_double)]


# name -> (restype, argtypes)
_P = ct.c_void_p
_E = ct.c_int
SIGNATURES = {
    "nqpu_state_create": (_P, [ct.c_uint32]),
    "nqpu_state_destroy": (None, [_P]),
    "nqpu_state_clone": (_P, [_P]),
    "nqpu_state_num_qubits": (ct.c_uint32, [_P]),
    "nqpu_state_get_amplitude": (_E, [_P, ct.c_uint64, ct.POINTER(NqpuComplex)]),
    "nqpu_state_set_amplitude": (_E, [_P, ct.c_uint64, NqpuComplex]),
    "nqpu_state_probabilities": (_E, [_P, _P, ct.c_uint64]),
    "nqpu_state_norm": (ct.c_double, [_P]),
    "nqpu_state_reset": (_E, [_P]),
    "nqpu_gate_h": (_E, [_P, ct.c_uint32]),
    "nqpu_gate_x": (_E, [_P, ct.c_uint32]),
    "nqpu_gate_y": (_E, [_P, ct.c_uint32]),
    "nqpu_gate_z": (_E, [_P, ct.c_uint32]),
    "nqpu_gate_s": (_E, [_P, ct.c_uint32]),
    "nqpu_gate_t": (_E, [_P, ct.c_uint32]),
    "nqpu_gate_rx": (_E, [_P, ct.c_uint32, ct.c_double]),
    "nqpu_gate_ry": (_E, [_P, ct.c_uint32, ct.c_double]),
    "nqpu_gate_rz": (_E, [_P, ct.c_uint32, ct.c_double]),
    "nqpu_gate_cx": (_E, [_P, ct.c_uint32, ct.c_uint32]),
    "nqpu_gate_cz": (_E, [_P, ct.c_uint32, ct.c_uint32]),
    "nqpu_gate_swap": (_E, [_P, ct.c_uint32, ct.c_uint32]),
    "nqpu_gate_custom_1q": (_E, [_P, ct.c_uint32, _P]),
    "nqpu_gate_custom_2q": (_E, [_P, ct.c_uint32, ct.c_uint32, _P]),
    "nqpu_measure": (_E, [_P, ct.c_uint32, ct.POINTER(NqpuMeasurement)]),
    "nqpu_measure_all": (_E, [_P, _P, ct.c_uint32]),
    "nqpu_expectation_z": (_E, [_P, ct.c_uint32, ct.POINTER(ct.c_double)]),
    "nqpu_sample": (_E, [_P, ct.c_uint32, _P]),
    "nqpu_circuit_create": (_P, [ct.c_uint32]),
    "nqpu_circuit_destroy": (None, [_P]),
    "nqpu_circuit_add_h": (_E, [_P, ct.c_uint32]),
    "nqpu_circuit_add_cx": (_E, [_P, ct.c_uint32, ct.c_uint32]),
    "nqpu_circuit_add_rx": (_E, [_P, ct.c_uint32, ct.c_double]),
    "nqpu_circuit_execute": (_E, [_P, _P]),
    "nqpu_circuit_gate_count": (ct.c_uint32, [_P]),
    "nqpu_circuit_depth": (ct.c_uint32, [_P]),
    "nqpu_version": (NqpuVersion, []),
    "nqpu_error_message": (ct.c_char_p, [_E]),
    "nqpu_set_seed": (None, [ct.c_uint64]),
    "nqpu_set_num_threads": (None, [ct.c_uint32]),
    "nqpu_state_create_ex": (_P, [ct.c_uint32, ct.c_uint32]),
    "nqpu_state_precision": (ct.c_uint32, [_P]),
    "nqpu_circuit_add_gate": (_E, [_P, ct.POINTER(NqpuGateOp)]),
    "nqpu_circuit_add_gates": (_E, [_P, _P, ct.c_uint64]),
    "nqpu_circuit_create_ex": (_P, [ct.c_uint32]),
    "nqpu_run": (_E, [_P, _P, ct.POINTER(NqpuRunOptions)]),
    "nqpu_last_run_stats": (_E, [_P, ct.POINTER(NqpuRunStats)]),
    "nqpu_state_synchronize": (_E, [_P]),
    "nqpu_graph_create": (_E, [_P, _P, ct.POINTER(NqpuRunOptions), ct.c_uint32, ct.POINTER(_P)]),
    "nqpu_graph_update": (_E, [_P, _P]),
    "nqpu_graph_launch": (_E, [_P, ct.c_uint32]),
    "nqpu_graph_info": (_E, [_P, ct.POINTER(NqpuGraphInfo)]),
    "nqpu_graph_destroy": (None, [_P]),
    "nqpu_expectation_z_all": (_E, [_P, _P, ct.c_uint32]),
    "nqpu_expectation_pauli": (_E, [_P, ct.c_uint64, ct.c_uint64, ct.POINTER(ct.c_double)]),
    "nqpu_expectation_pauli_sum": (_E, [_P, _P, _P, _P, ct.c_uint64, ct.POINTER(ct.c_double)]),
    "nqpu_expectation_pauli_sum_noalloc": (_E, [_P, _P, _P, _P, ct.c_uint64, ct.POINTER(ct.c_double)]),
    "nqpu_expectation_pauli_values": (_E, [_P, _P, _P, ct.c_uint64, _P]),
    "nqpu_expectation_pauli_string": (_E, [_P, ct.c_char_p, ct.c_uint32, ct.POINTER(ct.c_double)]),
    "nqpu_estimator_expectation": (_E, [_P, _P, ct.c_char_p, _P, _P, ct.c_uint64, ct.POINTER(ct.c_double)]),
    "nqpu_state_reset_qubit": (_E, [_P, ct.c_uint32]),
    "nqpu_state_initialize": (_E, [_P, _P, ct.c_uint64]),
    "nqpu_state_normalize": (_E, [_P]),
    "nqpu_state_overlap": (_E, [_P, _P, ct.POINTER(NqpuComplex)]),
    "nqpu_state_copy": (_E, [_P, _P]),
    "nqpu_state_get_amplitudes": (_E, [_P, _P, ct.c_uint64]),
    "nqpu_state_set_amplitudes": (_E, [_P, _P, ct.c_uint64]),
    "nqpu_last_error_detail": (ct.c_char_p, []),
    "nqpu_device_count": (ct.c_uint32, []),
    "nqpu_circuit_plan": (_E, [_P, ct.POINTER(NqpuRunOptions), ct.c_uint32, ct.POINTER(NqpuRunStats)]),
    "nqpu_circuit_fusion_counts": (_E, [_P, ct.POINTER(NqpuFusionCounts)]),
    "nqpu_debug_emulate": (_E, [_P, ct.POINTER(NqpuRunOptions), ct.c_uint32, _P, ct.c_uint64]),
    "nqpu_debug_compile_passes": (_E, [_P, ct.POINTER(NqpuRunOptions)]),
    "nqpu_debug_compile_passes_ex": (_E, [_P, ct.POINTER(NqpuRunOptions), ct.c_uint32]),
    "nqpu_debug_state_layout": (_E, [_P, _P, ct.c_uint32]),
    "nqpu_batch_create": (_P, []),
    "nqpu_batch_destroy": (None, [_P]),
    "nqpu_batch_expectation": (_E, [_P, _P, ct.c_uint64, _P, _P, ct.POINTER(NqpuRunOptions), _P]),
    "nqpu_batch_last_stats": (_E, [_P, ct.POINTER(NqpuBatchStats)]),
    "nqpu_batch_synchronize": (_E, [_P]),
    "nqpu_debug_batch_plan": (_E, [_P, ct.c_uint64, _P, _P, ct.POINTER(NqpuRunOptions),
                                   ct.POINTER(NqpuBatchStats), ct.POINTER(ct.c_uint64)]),
    "nqpu_server_create": (_P, [ct.POINTER(NqpuServerConfig)]),
    "nqpu_server_destroy": (None, [_P]),
    "nqpu_server_submit": (_E, [_P, _P, _P, ct.c_uint32, ct.c_uint32, ct.c_uint64, ct.POINTER(ct.c_uint64)]),
    "nqpu_server_poll": (_E, [_P, ct.c_uint64, ct.POINTER(NqpuRequestStatus)]),
    "nqpu_server_wait": (_E, [_P, ct.c_uint64, ct.c_uint32, ct.POINTER(NqpuRequestStatus)]),
    "nqpu_server_take": (_E, [_P, ct.c_uint64, _P, ct.c_uint32, _P, ct.c_uint32]),
    "nqpu_server_stats": (_E, [_P, ct.POINTER(NqpuServerStats)]),
    "nqpu_autotune": (_E, [ct.c_uint32, ct.c_uint32, _P, ct.POINTER(NqpuTunedPlan)]),
    "nqpu_tuned_plan": (_E, [ct.c_uint32, ct.c_uint32, ct.POINTER(NqpuTunedPlan)]),
    "nqpu_tune_load": (_E, [ct.c_char_p]),
    "nqpu_tune_reset": (None, []),
    "nqpu_circuit_from_json": (_E, [ct.c_char_p, ct.POINTER(_P), ct.POINTER(ct.c_uint32)]),
    "nqpu_execute_json": (_E, [ct.c_char_p, ct.c_uint32, _P, ct.c_uint64, ct.POINTER(ct.c_uint64)]),
    "nqpu_server_submit_json": (_E, [_P, ct.c_char_p, ct.c_uint64, ct.POINTER(ct.c_uint64)]),
    "nqpu_server_take_json": (_E, [_P, ct.c_uint64, ct.c_uint32, _P, ct.c_uint64, ct.POINTER(ct.c_uint64)]),
}

_lib = None


class NqpuLibraryMissing(RuntimeError):
    pass


def load():
    """Load libnqpu_hip.so and declare every prototype; raises if missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NqpuLibraryMissing(
            f"{LIB_PATH} is not built; run `python -m nqpu_amd.build` (hipcc, gfx950)")
    lib = ct.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class NqpuError(RuntimeError):
    def __init__(self, code: int, where: str = ""):
        lib = load()
        detail = (lib.nqpu_last_error_detail() or b"").decode(errors="replace")
        self.code = code
        super().__init__(f"{where}: {ERROR_NAMES.get(code, code)} ({detail})")


def check(code: int, where: str = "") -> None:
    if code != NQPU_OK:
        raise NqpuError(code, where)
