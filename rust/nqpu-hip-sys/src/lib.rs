//! Raw bindings to `libnqpu_hip.so` (declared in `include/nqpu_hip.h` of
//! this repository) plus owning handles for states and circuits.
//!
//! The first group of functions is the reference's C ABI
//! (`sdk/rust/src/infra/bindings/c_ffi.rs:788-1467`) with the same names and
//! meaning; the rest are the GPU engine's extensions.  Every function returns
//! an `NqpuError` code (0 = OK) unless noted; `nqpu_last_error_detail` gives
//! the calling thread's detail string.
//!
//! Threading: a state handle is bound to one HIP stream and is not
//! thread-safe (the owning wrappers below are `Send`, not `Sync`); distinct
//! handles may be used from distinct threads.  Host buffers are owned by the
//! caller; device memory by the handle.
#![allow(non_camel_case_types)]

use std::ffi::CStr;
use std::os::raw::{c_char, c_int};

// ------------------------------------------------------------------ types

#[repr(C)]
pub struct NqpuState {
    _private: [u8; 0],
}
#[repr(C)]
pub struct NqpuCircuit {
    _private: [u8; 0],
}
#[repr(C)]
pub struct NqpuBatch {
    _private: [u8; 0],
}
#[repr(C)]
pub struct NqpuRunGraph {
    _private: [u8; 0],
}
#[repr(C)]
pub struct NqpuServer {
    _private: [u8; 0],
}

pub type NqpuError = c_int;
pub const NQPU_OK: NqpuError = 0;
pub const NQPU_INVALID_QUBIT: NqpuError = 1;
pub const NQPU_INVALID_PARAMETER: NqpuError = 2;
pub const NQPU_ALLOCATION_FAILED: NqpuError = 3;
pub const NQPU_SIMULATION_FAILED: NqpuError = 4;
pub const NQPU_NULL_POINTER: NqpuError = 5;
pub const NQPU_INVALID_HANDLE: NqpuError = 6;
pub const NQPU_DEVICE_ERROR: NqpuError = 7;
pub const NQPU_OUT_OF_DEVICE_MEMORY: NqpuError = 8;
pub const NQPU_UNAVAILABLE: NqpuError = 9;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct NqpuComplex {
    pub re: f64,
    pub im: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuMeasurement {
    pub outcome: u32,
    pub probability: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuVersion {
    pub major: u32,
    pub minor: u32,
    pub patch: u32,
}

pub const NQPU_PRECISION_C64: u32 = 1;
pub const NQPU_PRECISION_C128: u32 = 2;

// NqpuGateKind: the order of GateType (sdk/rust/src/core/gates/gates.rs:8-64)
pub const NQPU_GATE_H: u32 = 0;
pub const NQPU_GATE_X: u32 = 1;
pub const NQPU_GATE_Y: u32 = 2;
pub const NQPU_GATE_Z: u32 = 3;
pub const NQPU_GATE_S: u32 = 4;
pub const NQPU_GATE_T: u32 = 5;
pub const NQPU_GATE_RX: u32 = 6;
pub const NQPU_GATE_RY: u32 = 7;
pub const NQPU_GATE_RZ: u32 = 8;
pub const NQPU_GATE_U: u32 = 9;
pub const NQPU_GATE_CNOT: u32 = 10;
pub const NQPU_GATE_CZ: u32 = 11;
pub const NQPU_GATE_SWAP: u32 = 12;
pub const NQPU_GATE_TOFFOLI: u32 = 13;
pub const NQPU_GATE_CRX: u32 = 14;
pub const NQPU_GATE_CRY: u32 = 15;
pub const NQPU_GATE_CRZ: u32 = 16;
pub const NQPU_GATE_CR: u32 = 17;
pub const NQPU_GATE_SX: u32 = 18;
pub const NQPU_GATE_PHASE: u32 = 19;
pub const NQPU_GATE_ISWAP: u32 = 20;
pub const NQPU_GATE_CCZ: u32 = 21;
pub const NQPU_GATE_RXX: u32 = 22;
pub const NQPU_GATE_RYY: u32 = 23;
pub const NQPU_GATE_RZZ: u32 = 24;
pub const NQPU_GATE_CSWAP: u32 = 25;
pub const NQPU_GATE_CU: u32 = 26;
pub const NQPU_GATE_CUSTOM1: u32 = 27;
pub const NQPU_GATE_CUSTOM2: u32 = 28;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct NqpuGateOp {
    pub kind: u32,
    pub num_targets: u32,
    pub num_controls: u32,
    pub targets: [u32; 3],
    pub controls: [u32; 3],
    pub reserved: u32,
    pub params: [f64; 4],
    pub matrix: [NqpuComplex; 16],
}

impl Default for NqpuGateOp {
    fn default() -> Self {
        NqpuGateOp {
            kind: 0,
            num_targets: 0,
            num_controls: 0,
            targets: [0; 3],
            controls: [0; 3],
            reserved: 0,
            params: [0.0; 4],
            matrix: [NqpuComplex::default(); 16],
        }
    }
}

pub const NQPU_SEMANTICS_STANDARD: u32 = 0;
pub const NQPU_SEMANTICS_REFERENCE: u32 = 1;
pub const NQPU_FUSION_NONE: u32 = 0;
pub const NQPU_FUSION_REFERENCE: u32 = 1;
pub const NQPU_FUSION_TILED: u32 = 2;

pub const NQPU_RUN_TIME_PASSES: u32 = 1;
pub const NQPU_RUN_LDS_TILES: u32 = 2;
pub const NQPU_RUN_BATCH_LAYOUT: u32 = 4;
pub const NQPU_RUN_NO_MFMA: u32 = 8;
pub const NQPU_RUN_COMPILE: u32 = 16;
pub const NQPU_RUN_NO_QPHASE: u32 = 32;
pub const NQPU_RUN_FIXED_LOW: u32 = 64;
pub const NQPU_RUN_UNTUNED: u32 = 128;
pub const NQPU_RUN_ASYNC: u32 = 256;
pub const NQPU_RUN_FROM_ZERO: u32 = 512;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuRunOptions {
    pub fusion: u32,
    pub semantics: u32,
    pub tile_qubits: u32,
    pub low_qubits: u32,
    pub flags: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuRunStats {
    pub gates: u64,
    pub device_ops: u64,
    pub passes: u64,
    pub bytes_moved: f64,
    pub pass_ms: f64,
    pub timed_passes: u64,
    pub remaps: u64,
    pub mfma_blocks: u64,
    pub mfma_gates: u64,
    pub exchanges: u64,
    pub qphase_gates: u64,
}

/// Counters of a run graph (nqpu_graph_info).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuGraphInfo {
    pub launches: u64,
    pub passes: u64,
    pub replays: u64,
    pub updates: u64,
    pub recaptures: u64,
    pub image_bytes: u64,
    pub plan_ms: f64,
}

pub const NQPU_GRAPH_FROM_ZERO: u32 = 1;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuTunedPlan {
    pub tile_qubits: u32,
    pub low_qubits: u32,
    pub flags: u32,
    pub tiles_per_group: u32,
    pub ms: f64,
    pub default_ms: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuFusionCounts {
    pub original_count: u64,
    pub fused_count: u64,
    pub gates_eliminated: u64,
    pub items: u64,
}

pub const NQPU_ORDER_LSB: u32 = 0;
pub const NQPU_ORDER_MSB: u32 = 1;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuPauliTerm {
    pub xmask: u64,
    pub zmask: u64,
    pub coeff: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuBatchStats {
    pub circuits: u64,
    pub resident: u64,
    pub gate_passes: u64,
    pub observable_passes: u64,
    pub plan_ms: f64,
    pub kernel_ms: f64,
    pub state_bytes: f64,
    pub hbm_bytes: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuServerConfig {
    pub slots: u32,
    pub precision: u32,
    pub memory_bytes: u64,
    pub run: NqpuRunOptions,
}

pub const NQPU_REQUEST_QUEUED: u32 = 0;
pub const NQPU_REQUEST_RUNNING: u32 = 1;
pub const NQPU_REQUEST_DONE: u32 = 2;
pub const NQPU_REQUEST_FAILED: u32 = 3;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuRequestStatus {
    pub state: u32,
    pub error: i32,
    pub queued_ms: f64,
    pub run_ms: f64,
    pub latency_ms: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NqpuServerStats {
    pub submitted: u64,
    pub completed: u64,
    pub failed: u64,
    pub peak_running: u64,
    pub peak_bytes: u64,
    pub memory_bytes: u64,
}

// -------------------------------------------------------------- functions

extern "C" {
    // reference C ABI (c_ffi.rs:788-1467)
    pub fn nqpu_state_create(num_qubits: u32) -> *mut NqpuState;
    pub fn nqpu_state_destroy(state: *mut NqpuState);
    pub fn nqpu_state_clone(state: *const NqpuState) -> *mut NqpuState;
    pub fn nqpu_state_num_qubits(state: *const NqpuState) -> u32;
    pub fn nqpu_state_get_amplitude(state: *const NqpuState, index: u64, out: *mut NqpuComplex) -> NqpuError;
    pub fn nqpu_state_set_amplitude(state: *mut NqpuState, index: u64, amp: NqpuComplex) -> NqpuError;
    pub fn nqpu_state_probabilities(state: *const NqpuState, out: *mut f64, len: u64) -> NqpuError;
    pub fn nqpu_state_norm(state: *const NqpuState) -> f64;
    pub fn nqpu_state_reset(state: *mut NqpuState) -> NqpuError;
    pub fn nqpu_gate_h(state: *mut NqpuState, qubit: u32) -> NqpuError;
    pub fn nqpu_gate_x(state: *mut NqpuState, qubit: u32) -> NqpuError;
    pub fn nqpu_gate_y(state: *mut NqpuState, qubit: u32) -> NqpuError;
    pub fn nqpu_gate_z(state: *mut NqpuState, qubit: u32) -> NqpuError;
    pub fn nqpu_gate_s(state: *mut NqpuState, qubit: u32) -> NqpuError;
    pub fn nqpu_gate_t(state: *mut NqpuState, qubit: u32) -> NqpuError;
    pub fn nqpu_gate_rx(state: *mut NqpuState, qubit: u32, angle: f64) -> NqpuError;
    pub fn nqpu_gate_ry(state: *mut NqpuState, qubit: u32, angle: f64) -> NqpuError;
    pub fn nqpu_gate_rz(state: *mut NqpuState, qubit: u32, angle: f64) -> NqpuError;
    pub fn nqpu_gate_cx(state: *mut NqpuState, control: u32, target: u32) -> NqpuError;
    pub fn nqpu_gate_cz(state: *mut NqpuState, control: u32, target: u32) -> NqpuError;
    pub fn nqpu_gate_swap(state: *mut NqpuState, q1: u32, q2: u32) -> NqpuError;
    pub fn nqpu_gate_custom_1q(state: *mut NqpuState, qubit: u32, matrix: *const NqpuComplex) -> NqpuError;
    pub fn nqpu_gate_custom_2q(state: *mut NqpuState, q1: u32, q2: u32, matrix: *const NqpuComplex) -> NqpuError;
    pub fn nqpu_measure(state: *mut NqpuState, qubit: u32, out: *mut NqpuMeasurement) -> NqpuError;
    pub fn nqpu_measure_all(state: *mut NqpuState, out: *mut u32, num_qubits: u32) -> NqpuError;
    pub fn nqpu_expectation_z(state: *const NqpuState, qubit: u32, out: *mut f64) -> NqpuError;
    pub fn nqpu_sample(state: *const NqpuState, num_shots: u32, out: *mut u64) -> NqpuError;
    pub fn nqpu_circuit_create(num_qubits: u32) -> *mut NqpuCircuit;
    pub fn nqpu_circuit_destroy(circuit: *mut NqpuCircuit);
    pub fn nqpu_circuit_add_h(circuit: *mut NqpuCircuit, qubit: u32) -> NqpuError;
    pub fn nqpu_circuit_add_cx(circuit: *mut NqpuCircuit, control: u32, target: u32) -> NqpuError;
    pub fn nqpu_circuit_add_rx(circuit: *mut NqpuCircuit, qubit: u32, angle: f64) -> NqpuError;
    pub fn nqpu_circuit_execute(circuit: *const NqpuCircuit, state: *mut NqpuState) -> NqpuError;
    pub fn nqpu_circuit_gate_count(circuit: *const NqpuCircuit) -> u32;
    pub fn nqpu_circuit_depth(circuit: *const NqpuCircuit) -> u32;
    pub fn nqpu_version() -> NqpuVersion;
    pub fn nqpu_error_message(error: NqpuError) -> *const c_char;
    pub fn nqpu_set_seed(seed: u64);
    pub fn nqpu_set_num_threads(threads: u32);

    // GPU extensions
    pub fn nqpu_state_create_ex(num_qubits: u32, precision: u32) -> *mut NqpuState;
    pub fn nqpu_state_precision(state: *const NqpuState) -> u32;
    pub fn nqpu_circuit_add_gate(circuit: *mut NqpuCircuit, op: *const NqpuGateOp) -> NqpuError;
    pub fn nqpu_circuit_add_gates(circuit: *mut NqpuCircuit, ops: *const NqpuGateOp, count: u64) -> NqpuError;
    pub fn nqpu_circuit_create_ex(num_qubits: u32) -> *mut NqpuCircuit;
    pub fn nqpu_run(circuit: *const NqpuCircuit, state: *mut NqpuState, opts: *const NqpuRunOptions) -> NqpuError;
    pub fn nqpu_last_run_stats(state: *const NqpuState, out: *mut NqpuRunStats) -> NqpuError;
    pub fn nqpu_state_synchronize(state: *mut NqpuState) -> NqpuError;
    pub fn nqpu_graph_create(
        circuit: *const NqpuCircuit,
        state: *mut NqpuState,
        opts: *const NqpuRunOptions,
        graph_flags: u32,
        out: *mut *mut NqpuRunGraph,
    ) -> NqpuError;
    pub fn nqpu_graph_update(graph: *mut NqpuRunGraph, circuit: *const NqpuCircuit) -> NqpuError;
    pub fn nqpu_graph_launch(graph: *mut NqpuRunGraph, flags: u32) -> NqpuError;
    pub fn nqpu_graph_info(graph: *const NqpuRunGraph, out: *mut NqpuGraphInfo) -> NqpuError;
    pub fn nqpu_graph_destroy(graph: *mut NqpuRunGraph);
    pub fn nqpu_expectation_z_all(state: *const NqpuState, out: *mut f64, len: u32) -> NqpuError;
    pub fn nqpu_expectation_pauli(state: *const NqpuState, xmask: u64, zmask: u64, out: *mut f64) -> NqpuError;
    pub fn nqpu_expectation_pauli_sum(
        state: *const NqpuState,
        xmasks: *const u64,
        zmasks: *const u64,
        coeffs: *const f64,
        count: u64,
        out: *mut f64,
    ) -> NqpuError;
    /// nqpu_expectation_pauli_sum without device allocation (a fixed 1 MiB
    /// array per device; at most 1024 terms; overlapping calls on one device race)
    pub fn nqpu_expectation_pauli_sum_noalloc(
        state: *const NqpuState,
        xmasks: *const u64,
        zmasks: *const u64,
        coeffs: *const f64,
        count: u64,
        out: *mut f64,
    ) -> NqpuError;
    pub fn nqpu_expectation_pauli_values(
        state: *const NqpuState,
        xmasks: *const u64,
        zmasks: *const u64,
        count: u64,
        out: *mut f64,
    ) -> NqpuError;
    pub fn nqpu_expectation_pauli_string(
        state: *const NqpuState,
        paulis: *const c_char,
        order: u32,
        out: *mut f64,
    ) -> NqpuError;
    pub fn nqpu_estimator_expectation(
        state: *const NqpuState,
        qubits: *const u32,
        paulis: *const c_char,
        offsets: *const u64,
        coeffs: *const f64,
        num_terms: u64,
        out: *mut f64,
    ) -> NqpuError;
    pub fn nqpu_state_reset_qubit(state: *mut NqpuState, qubit: u32) -> NqpuError;
    pub fn nqpu_state_initialize(state: *mut NqpuState, amplitudes: *const NqpuComplex, len: u64) -> NqpuError;
    pub fn nqpu_state_normalize(state: *mut NqpuState) -> NqpuError;
    pub fn nqpu_state_copy(dst: *mut NqpuState, src: *const NqpuState) -> NqpuError;
    pub fn nqpu_state_overlap(a: *const NqpuState, b: *const NqpuState, out: *mut NqpuComplex) -> NqpuError;
    pub fn nqpu_state_get_amplitudes(state: *const NqpuState, out: *mut NqpuComplex, len: u64) -> NqpuError;
    pub fn nqpu_state_set_amplitudes(state: *mut NqpuState, input: *const NqpuComplex, len: u64) -> NqpuError;
    pub fn nqpu_circuit_plan(
        circuit: *const NqpuCircuit,
        opts: *const NqpuRunOptions,
        precision: u32,
        out: *mut NqpuRunStats,
    ) -> NqpuError;
    pub fn nqpu_circuit_fusion_counts(circuit: *const NqpuCircuit, out: *mut NqpuFusionCounts) -> NqpuError;
    pub fn nqpu_batch_create() -> *mut NqpuBatch;
    pub fn nqpu_batch_destroy(batch: *mut NqpuBatch);
    pub fn nqpu_batch_expectation(
        batch: *mut NqpuBatch,
        circuits: *const *const NqpuCircuit,
        count: u64,
        terms: *const NqpuPauliTerm,
        offsets: *const u64,
        opts: *const NqpuRunOptions,
        out: *mut f64,
    ) -> NqpuError;
    pub fn nqpu_batch_synchronize(batch: *mut NqpuBatch) -> NqpuError;
    pub fn nqpu_batch_last_stats(batch: *const NqpuBatch, out: *mut NqpuBatchStats) -> NqpuError;
    pub fn nqpu_debug_batch_plan(
        circuits: *const *const NqpuCircuit,
        count: u64,
        terms: *const NqpuPauliTerm,
        offsets: *const u64,
        opts: *const NqpuRunOptions,
        out: *mut NqpuBatchStats,
        program_ops: *mut u64,
    ) -> NqpuError;
    pub fn nqpu_debug_emulate(
        circuit: *const NqpuCircuit,
        opts: *const NqpuRunOptions,
        precision: u32,
        amps: *mut NqpuComplex,
        len: u64,
    ) -> NqpuError;
    pub fn nqpu_debug_compile_passes(circuit: *const NqpuCircuit, opts: *const NqpuRunOptions) -> NqpuError;
    pub fn nqpu_debug_compile_passes_ex(
        circuit: *const NqpuCircuit,
        opts: *const NqpuRunOptions,
        precision: u32,
    ) -> NqpuError;
    pub fn nqpu_debug_state_layout(state: *const NqpuState, positions: *mut i32, len: u32) -> NqpuError;
    pub fn nqpu_last_error_detail() -> *const c_char;
    pub fn nqpu_device_count() -> u32;
    pub fn nqpu_server_create(config: *const NqpuServerConfig) -> *mut NqpuServer;
    pub fn nqpu_server_destroy(server: *mut NqpuServer);
    pub fn nqpu_server_submit(
        server: *mut NqpuServer,
        circuit: *const NqpuCircuit,
        terms: *const NqpuPauliTerm,
        num_terms: u32,
        shots: u32,
        seed: u64,
        request_id: *mut u64,
    ) -> NqpuError;
    pub fn nqpu_server_poll(server: *mut NqpuServer, request_id: u64, out: *mut NqpuRequestStatus) -> NqpuError;
    pub fn nqpu_server_wait(
        server: *mut NqpuServer,
        request_id: u64,
        timeout_ms: u32,
        out: *mut NqpuRequestStatus,
    ) -> NqpuError;
    pub fn nqpu_server_take(
        server: *mut NqpuServer,
        request_id: u64,
        expectations: *mut f64,
        num_terms: u32,
        samples: *mut u64,
        shots: u32,
    ) -> NqpuError;
    pub fn nqpu_server_stats(server: *const NqpuServer, out: *mut NqpuServerStats) -> NqpuError;

    // plan autotuning
    pub fn nqpu_autotune(num_qubits: u32, precision: u32, probe: *const NqpuCircuit, out: *mut NqpuTunedPlan) -> NqpuError;
    pub fn nqpu_tuned_plan(num_qubits: u32, precision: u32, out: *mut NqpuTunedPlan) -> NqpuError;
    pub fn nqpu_tune_load(path: *const c_char) -> NqpuError;
    pub fn nqpu_tune_reset();

    // request formats (web/mod.rs:34-160, circuit_serde.rs:329-340)
    pub fn nqpu_circuit_from_json(json: *const c_char, circuit: *mut *mut NqpuCircuit, shots: *mut u32) -> NqpuError;
    pub fn nqpu_execute_json(
        request: *const c_char,
        precision: u32,
        response: *mut c_char,
        len: u64,
        needed: *mut u64,
    ) -> NqpuError;
    pub fn nqpu_server_submit_json(
        server: *mut NqpuServer,
        request: *const c_char,
        seed: u64,
        request_id: *mut u64,
    ) -> NqpuError;
    pub fn nqpu_server_take_json(
        server: *mut NqpuServer,
        request_id: u64,
        timeout_ms: u32,
        response: *mut c_char,
        len: u64,
        needed: *mut u64,
    ) -> NqpuError;
}

// ------------------------------------------------------- owning wrappers

/// An error code of the library with the calling thread's detail string.
#[derive(Clone, Debug)]
pub struct HipError {
    pub code: NqpuError,
    pub detail: String,
}

impl std::fmt::Display for HipError {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        write!(f, "nqpu_hip error {}: {}", self.code, self.detail)
    }
}

impl std::error::Error for HipError {}

/// The detail string of the last error on this thread.
pub fn last_error_detail() -> String {
    unsafe {
        let p = nqpu_last_error_detail();
        if p.is_null() {
            String::new()
        } else {
            CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    }
}

/// Ok(()) for NQPU_OK, else the code with its detail string.
pub fn check(code: NqpuError) -> Result<(), HipError> {
    if code == NQPU_OK {
        Ok(())
    } else {
        Err(HipError { code, detail: last_error_detail() })
    }
}

/// A device-resident state (complex64 or complex128), freed on drop.
pub struct HipState {
    ptr: *mut NqpuState,
    num_qubits: usize,
}

// The handle is used by one thread at a time (no interior sharing).
unsafe impl Send for HipState {}

impl HipState {
    pub fn new(num_qubits: usize, precision: u32) -> Result<Self, HipError> {
        let ptr = unsafe { nqpu_state_create_ex(num_qubits as u32, precision) };
        if ptr.is_null() {
            return Err(HipError { code: NQPU_DEVICE_ERROR, detail: last_error_detail() });
        }
        Ok(HipState { ptr, num_qubits })
    }
    pub fn num_qubits(&self) -> usize {
        self.num_qubits
    }
    pub fn as_ptr(&self) -> *const NqpuState {
        self.ptr
    }
    pub fn as_mut_ptr(&mut self) -> *mut NqpuState {
        self.ptr
    }
}

impl Drop for HipState {
    fn drop(&mut self) {
        unsafe { nqpu_state_destroy(self.ptr) }
    }
}

/// A circuit in the library's IR, freed on drop.
pub struct HipCircuit {
    ptr: *mut NqpuCircuit,
}

unsafe impl Send for HipCircuit {}

impl HipCircuit {
    pub fn new(num_qubits: usize) -> Result<Self, HipError> {
        let ptr = unsafe { nqpu_circuit_create_ex(num_qubits as u32) };
        if ptr.is_null() {
            return Err(HipError { code: NQPU_INVALID_PARAMETER, detail: last_error_detail() });
        }
        Ok(HipCircuit { ptr })
    }
    pub fn add_gates(&mut self, ops: &[NqpuGateOp]) -> Result<(), HipError> {
        check(unsafe { nqpu_circuit_add_gates(self.ptr, ops.as_ptr(), ops.len() as u64) })
    }
    pub fn run(&self, state: &mut HipState, opts: Option<&NqpuRunOptions>) -> Result<(), HipError> {
        let o = opts.map_or(std::ptr::null(), |o| o as *const NqpuRunOptions);
        check(unsafe { nqpu_run(self.ptr, state.as_mut_ptr(), o) })
    }
    pub fn as_ptr(&self) -> *const NqpuCircuit {
        self.ptr
    }
}

impl Drop for HipCircuit {
    fn drop(&mut self) {
        unsafe { nqpu_circuit_destroy(self.ptr) }
    }
}

#[cfg(test)]
mod tests {
    use super::*;
    use std::mem::{align_of, size_of};

    // layouts of include/nqpu_hip.h (x86-64 / aarch64 C ABI)
    #[test]
    fn struct_layouts_match_the_header() {
        assert_eq!(size_of::<NqpuComplex>(), 16);
        assert_eq!(size_of::<NqpuMeasurement>(), 16);
        assert_eq!(size_of::<NqpuGateOp>(), 4 * 3 + 4 * 3 + 4 * 3 + 4 + 8 * 4 + 16 * 16);
        assert_eq!(align_of::<NqpuGateOp>(), 8);
        assert_eq!(size_of::<NqpuRunOptions>(), 20);
        assert_eq!(size_of::<NqpuRunStats>(), 11 * 8);
        assert_eq!(size_of::<NqpuBatchStats>(), 8 * 8);
        assert_eq!(size_of::<NqpuServerConfig>(), 4 + 4 + 8 + 20 + 4);
        assert_eq!(size_of::<NqpuRequestStatus>(), 32);
        assert_eq!(size_of::<NqpuPauliTerm>(), 24);
    }

    // needs the library and a device: cargo test -- --ignored
    #[test]
    #[ignore]
    fn bell_state_on_the_device() {
        if unsafe { nqpu_device_count() } == 0 {
            return;
        }
        let mut st = HipState::new(2, NQPU_PRECISION_C128).unwrap();
        let mut c = HipCircuit::new(2).unwrap();
        let mut h = NqpuGateOp::default();
        h.kind = NQPU_GATE_H;
        h.num_targets = 1;
        let mut cx = NqpuGateOp::default();
        cx.kind = NQPU_GATE_CNOT;
        cx.num_targets = 1;
        cx.num_controls = 1;
        cx.targets[0] = 1;
        cx.controls[0] = 0;
        c.add_gates(&[h, cx]).unwrap();
        c.run(&mut st, None).unwrap();
        let mut p = [0.0f64; 4];
        check(unsafe { nqpu_state_probabilities(st.as_ptr(), p.as_mut_ptr(), 4) }).unwrap();
        assert!((p[0] - 0.5).abs() < 1e-12 && (p[3] - 0.5).abs() < 1e-12);
    }
}
