//! `QuantumBackend` over the MI355X statevector engine (libnqpu_hip).
//!
//! Goes into the reference crate as `sdk/rust/src/backends/gpu/hip_backend.rs`
//! behind its `rocm` feature (see `rocm_backend.patch` beside this file, which
//! also adds the `SimBackend::RocmGPU` / `RocmGPUOnly` arms that call
//! `run_probabilities`).  It implements `trait QuantumBackend`
//! (`sdk/rust/src/infra/execution/traits.rs:54-110`) with the error type of
//! `traits.rs:17-41`.
//!
//! Gates are queued on the host and run as one fused `nqpu_run` (tiled
//! multi-gate passes) when a result is read.  The trait's reads take `&self`,
//! so the queue and the device handle sit in a `RefCell`: the backend is
//! `Send` (the trait's bound) and not `Sync`, matching the reference's
//! single-owner model.  State sizes: complex128 up to 30 qubits (the
//! reference's FP64 semantics), complex64 from 31 to 34.
//!
//! This file is source for a maintainer: it is written against the
//! reference's types but has not been compiled here (no Rust toolchain in
//! this environment).
use std::cell::RefCell;
use std::collections::HashMap;

use nqpu_hip_sys as sys;

use crate::gates::{Gate, GateType};
use crate::traits::{BackendError, BackendResult, QuantumBackend};

/// True when libnqpu_hip sees a usable gfx950 device.
pub fn rocm_device_available() -> bool {
    unsafe { sys::nqpu_device_count() > 0 }
}

fn internal(code: sys::NqpuError) -> BackendError {
    BackendError::Internal(format!("nqpu_hip error {}: {}", code, sys::last_error_detail()))
}

fn qubit(q: usize, n: usize) -> BackendResult<u32> {
    if q >= n {
        return Err(BackendError::QubitOutOfRange { qubit: q, num_qubits: n });
    }
    Ok(q as u32)
}

/// One reference gate as the engine's packed op: kinds in `GateType` order,
/// qubit roles as the reference's constructors (gates.rs:669-760).
pub fn to_gate_op(g: &Gate, n: usize) -> BackendResult<sys::NqpuGateOp> {
    let mut op = sys::NqpuGateOp::default();
    let mut params: Vec<f64> = Vec::new();
    let (kind, nt, nc) = match &g.gate_type {
        GateType::H => (sys::NQPU_GATE_H, 1, 0),
        GateType::X => (sys::NQPU_GATE_X, 1, 0),
        GateType::Y => (sys::NQPU_GATE_Y, 1, 0),
        GateType::Z => (sys::NQPU_GATE_Z, 1, 0),
        GateType::S => (sys::NQPU_GATE_S, 1, 0),
        GateType::T => (sys::NQPU_GATE_T, 1, 0),
        GateType::SX => (sys::NQPU_GATE_SX, 1, 0),
        GateType::Rx(a) => {
            params.push(*a);
            (sys::NQPU_GATE_RX, 1, 0)
        }
        GateType::Ry(a) => {
            params.push(*a);
            (sys::NQPU_GATE_RY, 1, 0)
        }
        GateType::Rz(a) => {
            params.push(*a);
            (sys::NQPU_GATE_RZ, 1, 0)
        }
        GateType::Phase(a) => {
            params.push(*a);
            (sys::NQPU_GATE_PHASE, 1, 0)
        }
        GateType::U { theta, phi, lambda } => {
            params.extend_from_slice(&[*theta, *phi, *lambda]);
            (sys::NQPU_GATE_U, 1, 0)
        }
        GateType::CNOT => (sys::NQPU_GATE_CNOT, 1, 1),
        GateType::CZ => (sys::NQPU_GATE_CZ, 1, 1),
        GateType::SWAP => (sys::NQPU_GATE_SWAP, 2, 0),
        GateType::ISWAP => (sys::NQPU_GATE_ISWAP, 2, 0),
        GateType::Toffoli => (sys::NQPU_GATE_TOFFOLI, 1, 2),
        GateType::CCZ => (sys::NQPU_GATE_CCZ, 1, 2),
        GateType::CSWAP => (sys::NQPU_GATE_CSWAP, 2, 1),
        GateType::CRx(a) => {
            params.push(*a);
            (sys::NQPU_GATE_CRX, 1, 1)
        }
        GateType::CRy(a) => {
            params.push(*a);
            (sys::NQPU_GATE_CRY, 1, 1)
        }
        GateType::CRz(a) => {
            params.push(*a);
            (sys::NQPU_GATE_CRZ, 1, 1)
        }
        GateType::CR(a) => {
            params.push(*a);
            (sys::NQPU_GATE_CR, 1, 1)
        }
        GateType::Rxx(a) => {
            params.push(*a);
            (sys::NQPU_GATE_RXX, 2, 0)
        }
        GateType::Ryy(a) => {
            params.push(*a);
            (sys::NQPU_GATE_RYY, 2, 0)
        }
        GateType::Rzz(a) => {
            params.push(*a);
            (sys::NQPU_GATE_RZZ, 2, 0)
        }
        GateType::CU { theta, phi, lambda, gamma } => {
            params.extend_from_slice(&[*theta, *phi, *lambda, *gamma]);
            (sys::NQPU_GATE_CU, 1, 1)
        }
        GateType::Custom(m) => {
            // 2x2 on targets[0]; 4x4 with targets[0] as the high bit of the index
            let dim = m.len();
            let nt = match dim {
                2 => 1,
                4 => 2,
                _ => return Err(BackendError::UnsupportedGate(format!("custom {}x{} matrix", dim, dim))),
            };
            for (r, row) in m.iter().enumerate() {
                if row.len() != dim {
                    return Err(BackendError::UnsupportedGate("non-square custom matrix".into()));
                }
                for (c, v) in row.iter().enumerate() {
                    op.matrix[r * dim + c] = sys::NqpuComplex { re: v.re, im: v.im };
                }
            }
            (if nt == 1 { sys::NQPU_GATE_CUSTOM1 } else { sys::NQPU_GATE_CUSTOM2 }, nt, 0)
        }
    };
    if g.targets.len() < nt || g.controls.len() < nc {
        return Err(BackendError::Internal(format!("{:?}: expected {} targets and {} controls", g.gate_type, nt, nc)));
    }
    op.kind = kind;
    op.num_targets = nt as u32;
    op.num_controls = nc as u32;
    for i in 0..nt {
        op.targets[i] = qubit(g.targets[i], n)?;
    }
    for i in 0..nc {
        op.controls[i] = qubit(g.controls[i], n)?;
    }
    for (i, p) in params.iter().enumerate() {
        op.params[i] = *p;
    }
    Ok(op)
}

struct Inner {
    state: sys::HipState,
    pending: Vec<sys::NqpuGateOp>,
}

impl Inner {
    /// Run the queued gates as one circuit.
    fn flush(&mut self) -> BackendResult<()> {
        if self.pending.is_empty() {
            return Ok(());
        }
        let n = self.state.num_qubits();
        let mut c = sys::HipCircuit::new(n).map_err(|e| internal(e.code))?;
        let ops = std::mem::take(&mut self.pending);
        c.add_gates(&ops).map_err(|e| internal(e.code))?;
        c.run(&mut self.state, None).map_err(|e| internal(e.code))?;
        Ok(())
    }
}

/// Dense statevector on one MI355X.
pub struct HipStateVectorBackend {
    n: usize,
    inner: RefCell<Inner>,
}

impl HipStateVectorBackend {
    pub fn new(num_qubits: usize) -> BackendResult<Self> {
        let prec = if num_qubits <= 30 { sys::NQPU_PRECISION_C128 } else { sys::NQPU_PRECISION_C64 };
        Self::with_precision(num_qubits, prec)
    }

    /// `precision`: `sys::NQPU_PRECISION_C64` or `sys::NQPU_PRECISION_C128`.
    pub fn with_precision(num_qubits: usize, precision: u32) -> BackendResult<Self> {
        if !rocm_device_available() {
            return Err(BackendError::Unavailable("no ROCm (gfx950) device".into()));
        }
        let state = sys::HipState::new(num_qubits, precision).map_err(|e| match e.code {
            sys::NQPU_UNAVAILABLE => BackendError::Unavailable(e.detail),
            _ => BackendError::Internal(e.detail),
        })?;
        Ok(HipStateVectorBackend { n: num_qubits, inner: RefCell::new(Inner { state, pending: Vec::new() }) })
    }

    /// <P> of a Pauli string (qubit q = bit q) on the device, in FP64.
    pub fn expectation_pauli(&self, xmask: u64, zmask: u64) -> BackendResult<f64> {
        let mut inner = self.inner.borrow_mut();
        inner.flush()?;
        let mut v = 0.0f64;
        let rc = unsafe { sys::nqpu_expectation_pauli(inner.state.as_ptr(), xmask, zmask, &mut v) };
        sys::check(rc).map_err(|e| internal(e.code))?;
        Ok(v)
    }
}

impl QuantumBackend for HipStateVectorBackend {
    fn num_qubits(&self) -> usize {
        self.n
    }

    fn apply_gate(&mut self, gate: &Gate) -> BackendResult<()> {
        let op = to_gate_op(gate, self.n)?;
        self.inner.get_mut().pending.push(op);
        Ok(())
    }

    fn probabilities(&self) -> BackendResult<Vec<f64>> {
        let mut inner = self.inner.borrow_mut();
        inner.flush()?;
        let dim = 1usize << self.n;
        let mut out = vec![0.0f64; dim];
        let rc = unsafe { sys::nqpu_state_probabilities(inner.state.as_ptr(), out.as_mut_ptr(), dim as u64) };
        sys::check(rc).map_err(|e| internal(e.code))?;
        Ok(out)
    }

    fn sample(&self, n_shots: usize) -> BackendResult<HashMap<usize, usize>> {
        let mut inner = self.inner.borrow_mut();
        inner.flush()?;
        let mut counts = HashMap::new();
        let mut left = n_shots;
        let mut buf: Vec<u64> = Vec::new();
        while left > 0 {
            let k = left.min(u32::MAX as usize);
            buf.resize(k, 0);
            let rc = unsafe { sys::nqpu_sample(inner.state.as_ptr(), k as u32, buf.as_mut_ptr()) };
            sys::check(rc).map_err(|e| internal(e.code))?;
            for &i in &buf {
                *counts.entry(i as usize).or_insert(0) += 1;
            }
            left -= k;
        }
        Ok(counts)
    }

    fn measure_qubit(&mut self, q: usize) -> BackendResult<u8> {
        let qq = qubit(q, self.n)?;
        let inner = self.inner.get_mut();
        inner.flush()?;
        let mut m = sys::NqpuMeasurement::default();
        let rc = unsafe { sys::nqpu_measure(inner.state.as_mut_ptr(), qq, &mut m) };
        sys::check(rc).map_err(|e| internal(e.code))?;
        Ok(m.outcome as u8)
    }

    fn reset(&mut self) {
        let inner = self.inner.get_mut();
        inner.pending.clear();
        unsafe {
            let _ = sys::nqpu_state_reset(inner.state.as_mut_ptr());
        }
    }

    fn name(&self) -> &str {
        "StateVector (ROCm MI355X)"
    }

    fn apply_circuit(&mut self, gates: &[Gate]) -> BackendResult<()> {
        let n = self.n;
        let inner = self.inner.get_mut();
        inner.pending.reserve(gates.len());
        for g in gates {
            inner.pending.push(to_gate_op(g, n)?);
        }
        Ok(())
    }

    fn expectation_z(&self, q: usize) -> BackendResult<f64> {
        let qq = qubit(q, self.n)?;
        let mut inner = self.inner.borrow_mut();
        inner.flush()?;
        let mut v = 0.0f64;
        let rc = unsafe { sys::nqpu_expectation_z(inner.state.as_ptr(), qq, &mut v) };
        sys::check(rc).map_err(|e| internal(e.code))?;
        Ok(v)
    }
}

/// What `SimBackend::RocmGPU` / `RocmGPUOnly` run: the circuit from |0...0>
/// in one fused execution, then the probability vector (as every arm of
/// `AutoSimulator::execute_inner` returns).
pub fn run_probabilities(num_qubits: usize, gates: &[Gate]) -> BackendResult<Vec<f64>> {
    let mut b = HipStateVectorBackend::new(num_qubits)?;
    b.apply_circuit(gates)?;
    b.probabilities()
}
