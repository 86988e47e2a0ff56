/*
 * nqpu_hip.h — C ABI of the MI355X (gfx950) statevector engine.
 *
 * The first block of declarations is the reference's C interface, with the
 * same names, argument meaning and error behaviour, so that a C caller of
 * the reference (and the reference's own C tests) can link against
 * libnqpu_hip.so instead.  The reference's header is generated at
 * sdk/rust/src/infra/bindings/c_ffi.rs:1480-1620; each function cites the
 * export it replaces.  Amplitudes live in GPU HBM; host buffers are always
 * caller-owned.  A handle is bound to one HIP stream and is not thread-safe;
 * distinct handles may be used from distinct threads.
 *
 * The second block holds the extensions the GPU engine adds (precision
 * selection, lifted qubit cap, packed circuit upload, fused execution, FP64
 * Pauli reductions, bulk amplitude transfer, batched execution).
 */
#ifndef NQPU_HIP_H
#define NQPU_HIP_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */

/* Opaque handles (c_ffi.rs:77-92). */
typedef struct NqpuState NqpuState;
typedef struct NqpuCircuit NqpuCircuit;

/* Error codes (c_ffi.rs:95-113); 7..9 are GPU-specific additions. */
typedef enum {
  NQPU_OK = 0,
  NQPU_INVALID_QUBIT = 1,
  NQPU_INVALID_PARAMETER = 2,
  NQPU_ALLOCATION_FAILED = 3,
  NQPU_SIMULATION_FAILED = 4,
  NQPU_NULL_POINTER = 5,
  NQPU_INVALID_HANDLE = 6,
  NQPU_DEVICE_ERROR = 7,          /* a HIP call or kernel launch failed */
  NQPU_OUT_OF_DEVICE_MEMORY = 8,  /* the state does not fit in HBM */
  NQPU_UNAVAILABLE = 9            /* no usable gfx950 device */
} NqpuError;

/* c_ffi.rs:118-123 */
typedef struct {
  double re;
  double im;
} NqpuComplex;

/* c_ffi.rs:161-168 */
typedef struct {
  uint32_t outcome;
  double probability;
} NqpuMeasurement;

/* c_ffi.rs:171-178 (reserved, as in the reference) */
typedef struct {
  uint32_t num_qubits;
  uint64_t seed;
  bool use_gpu;
  uint32_t num_threads;
} NqpuConfig;

/* c_ffi.rs:181-187 */
typedef struct {
  uint32_t major;
  uint32_t minor;
  uint32_t patch;
} NqpuVersion;

/* ---------------------------------------------- reference C ABI (drop-in) */

/* c_ffi.rs:789  — complex128 state, 1 <= n <= 30, NULL on failure */
NqpuState* nqpu_state_create(uint32_t num_qubits);
/* c_ffi.rs:805  — NULL-safe */
void nqpu_state_destroy(NqpuState* state);
/* c_ffi.rs:817  — device-to-device copy */
NqpuState* nqpu_state_clone(const NqpuState* state);
/* c_ffi.rs:836  — 0 on NULL */
uint32_t nqpu_state_num_qubits(const NqpuState* state);
/* c_ffi.rs:850  — index >= 2^n -> INVALID_PARAMETER */
NqpuError nqpu_state_get_amplitude(const NqpuState* state, uint64_t index, NqpuComplex* out);
/* c_ffi.rs:877  — does not renormalise */
NqpuError nqpu_state_set_amplitude(NqpuState* state, uint64_t index, NqpuComplex amp);
/* c_ffi.rs:900  — len < 2^n -> INVALID_PARAMETER */
NqpuError nqpu_state_probabilities(const NqpuState* state, double* out, uint64_t len);
/* c_ffi.rs:929  — sum |a|^2 in FP64; -1.0 on NULL */
double nqpu_state_norm(const NqpuState* state);
/* c_ffi.rs:941 */
NqpuError nqpu_state_reset(NqpuState* state);

/* c_ffi.rs:987-1004 (macro-generated gate exports) */
NqpuError nqpu_gate_h(NqpuState* state, uint32_t qubit);
NqpuError nqpu_gate_x(NqpuState* state, uint32_t qubit);
NqpuError nqpu_gate_y(NqpuState* state, uint32_t qubit);
NqpuError nqpu_gate_z(NqpuState* state, uint32_t qubit);
NqpuError nqpu_gate_s(NqpuState* state, uint32_t qubit);
NqpuError nqpu_gate_t(NqpuState* state, uint32_t qubit);
/* non-finite angle -> INVALID_PARAMETER (c_ffi.rs:987) */
NqpuError nqpu_gate_rx(NqpuState* state, uint32_t qubit, double angle);
NqpuError nqpu_gate_ry(NqpuState* state, uint32_t qubit, double angle);
NqpuError nqpu_gate_rz(NqpuState* state, uint32_t qubit, double angle);
/* c_ffi.rs:1010, 1032, 1054 — equal qubits -> INVALID_PARAMETER */
NqpuError nqpu_gate_cx(NqpuState* state, uint32_t control, uint32_t target);
NqpuError nqpu_gate_cz(NqpuState* state, uint32_t control, uint32_t target);
NqpuError nqpu_gate_swap(NqpuState* state, uint32_t q1, uint32_t q2);
/* c_ffi.rs:1079 — row-major [m00, m01, m10, m11] */
NqpuError nqpu_gate_custom_1q(NqpuState* state, uint32_t qubit, const NqpuComplex* matrix);
/* c_ffi.rs:1111 — row-major 4x4, basis |q1 q2> with q1 the high bit of the
 * matrix index (the ordering documented at c_ffi.rs:300-303; every 4-amplitude
 * group is transformed exactly once) */
NqpuError nqpu_gate_custom_2q(NqpuState* state, uint32_t q1, uint32_t q2, const NqpuComplex* matrix);

/* c_ffi.rs:1151 — collapses; uses the seeded xorshift64* stream */
NqpuError nqpu_measure(NqpuState* state, uint32_t qubit, NqpuMeasurement* out);
/* c_ffi.rs:1177 — measures qubit 0..n-1 in order */
NqpuError nqpu_measure_all(NqpuState* state, uint32_t* out, uint32_t num_qubits);
/* c_ffi.rs:1205 */
NqpuError nqpu_expectation_z(const NqpuState* state, uint32_t qubit, double* out);
/* c_ffi.rs:1231 — basis indices; outcome = first i with r < cumsum(|a|^2) */
NqpuError nqpu_sample(const NqpuState* state, uint32_t num_shots, uint64_t* out);

/* c_ffi.rs:1259-1398 */
NqpuCircuit* nqpu_circuit_create(uint32_t num_qubits);
void nqpu_circuit_destroy(NqpuCircuit* circuit);
NqpuError nqpu_circuit_add_h(NqpuCircuit* circuit, uint32_t qubit);
NqpuError nqpu_circuit_add_cx(NqpuCircuit* circuit, uint32_t control, uint32_t target);
NqpuError nqpu_circuit_add_rx(NqpuCircuit* circuit, uint32_t qubit, double angle);
NqpuError nqpu_circuit_execute(const NqpuCircuit* circuit, NqpuState* state);
uint32_t nqpu_circuit_gate_count(const NqpuCircuit* circuit);
uint32_t nqpu_circuit_depth(const NqpuCircuit* circuit);

/* c_ffi.rs:1403-1467 */
NqpuVersion nqpu_version(void);
const char* nqpu_error_message(NqpuError error);
void nqpu_set_seed(uint64_t seed);
/* host-side worker threads (circuit planning); the GPU path ignores it */
void nqpu_set_num_threads(uint32_t threads);

/* ------------------------------------------------------- GPU extensions */

/* Amplitude precision of a state. */
typedef enum {
  NQPU_PRECISION_C64 = 1,   /* complex64: 8 bytes per amplitude */
  NQPU_PRECISION_C128 = 2   /* complex128: 16 bytes per amplitude */
} NqpuPrecision;

/* Gate kinds: the order of GateType (sdk/rust/src/core/gates/gates.rs:8-64). */
typedef enum {
  NQPU_GATE_H = 0, NQPU_GATE_X, NQPU_GATE_Y, NQPU_GATE_Z, NQPU_GATE_S, NQPU_GATE_T,
  NQPU_GATE_RX, NQPU_GATE_RY, NQPU_GATE_RZ, NQPU_GATE_U, NQPU_GATE_CNOT, NQPU_GATE_CZ,
  NQPU_GATE_SWAP, NQPU_GATE_TOFFOLI, NQPU_GATE_CRX, NQPU_GATE_CRY, NQPU_GATE_CRZ,
  NQPU_GATE_CR, NQPU_GATE_SX, NQPU_GATE_PHASE, NQPU_GATE_ISWAP, NQPU_GATE_CCZ,
  NQPU_GATE_RXX, NQPU_GATE_RYY, NQPU_GATE_RZZ, NQPU_GATE_CSWAP, NQPU_GATE_CU,
  NQPU_GATE_CUSTOM1, NQPU_GATE_CUSTOM2, NQPU_GATE_NUM_KINDS
} NqpuGateKind;

/* One gate, as gates::Gate (gates.rs:632): targets, controls and parameters.
 * Qubit roles follow the reference constructors (gates.rs:669-760):
 * CNOT/CZ/CR/CRx/CRy/CRz/CU: controls[0], targets[0]; SWAP/ISWAP/Rxx/Ryy/Rzz:
 * targets[0..1]; Toffoli/CCZ: controls[0..1], targets[0]; CSWAP: controls[0],
 * targets[0..1].  params: Rx, Ry, Rz, Phase, CRx, CRy, CRz, CR, Rxx, Ryy, Rzz: params[0]; U: theta, phi,
 * lambda; CU: theta, phi, lambda, gamma.  matrix: CUSTOM1 2x2 / CUSTOM2 4x4,
 * row-major, with targets[0] as the high bit of a CUSTOM2 index. */
typedef struct {
  uint32_t kind;
  uint32_t num_targets;
  uint32_t num_controls;
  uint32_t targets[3];
  uint32_t controls[3];
  uint32_t reserved;
  double params[4];
  NqpuComplex matrix[16];
} NqpuGateOp;

/* Gate semantics of a run. */
typedef enum {
  /* the matrices of gates.rs for every kind */
  NQPU_SEMANTICS_STANDARD = 0,
  /* what the reference's dispatcher executes (ascii_viz.rs:124-179): CRx acts
   * as lib.rs:1292, CRz as lib.rs:1439-1473, Rxx/Ryy/Rzz/CSWAP/CU/Custom are
   * no-ops */
  NQPU_SEMANTICS_REFERENCE = 1
} NqpuSemantics;

/* How a circuit is turned into device passes. */
typedef enum {
  /* one pass per gate, as GateOperations (lib.rs:666-1700) */
  NQPU_FUSION_NONE = 0,
  /* the reference's fuse_gates + fuse_diagonal_pass (gate_fusion.rs:537-851),
   * one pass per fused item, as execute_fused_circuit (gate_fusion.rs:1158) */
  NQPU_FUSION_REFERENCE = 1,
  /* LDS-tiled multi-gate passes: every gate whose qubits fit one tile of up
   * to `tile_qubits` qubits is applied in a single HBM sweep */
  NQPU_FUSION_TILED = 2
} NqpuFusion;

typedef struct {
  uint32_t fusion;       /* NqpuFusion; default NQPU_FUSION_TILED */
  uint32_t semantics;    /* NqpuSemantics; default NQPU_SEMANTICS_STANDARD */
  uint32_t tile_qubits;  /* 0 = default: the plan nqpu_autotune measured for this qubit count and
                            precision, else 12 (complex64) / 11 (complex128), one less for small states */
  uint32_t low_qubits;   /* contiguous low qubits always in a tile; 0 = default */
  uint32_t flags;        /* NQPU_RUN_TIME_PASSES: time every pass with HIP events */
} NqpuRunOptions;

#define NQPU_RUN_TIME_PASSES 1u
/* run tiled passes with the LDS-resident tile kernel (one LDS sweep per gate)
 * instead of the register-resident one */
#define NQPU_RUN_LDS_TILES 2u
/* nqpu_debug_emulate only: emulate the layout the batched kernel uses */
#define NQPU_RUN_BATCH_LAYOUT 4u
/* complex64 register tiles: apply every gate on the vector ALUs instead of
 * fusing runs of gates on four qubits into dense blocks for the matrix cores */
#define NQPU_RUN_NO_MFMA 8u
/* register tiles (complex64: 12-qubit tiles; complex128: any width): run each
 * pass as its own straight-line kernel,
 * generated from the pass's register program and compiled with hipRTC on
 * first use (complex64: about a second per eight passes, complex128 several
 * times that, on parallel threads); the kernels depend only on the program's
 * structure, not on gate angles, and stay cached for the life of the process.
 * Compiled complex128 passes move bits with register/lane exchanges instead
 * of LDS remaps. */
#define NQPU_RUN_COMPILE 16u
/* complex64 register tiles: apply runs of diagonal gates one by one instead
 * of merging each run into one quadratic phase evaluated per amplitude */
#define NQPU_RUN_NO_QPHASE 32u
/* complex64 register tiles: keep the lowest qubits of the state in every
 * tile instead of handing the tile's low positions to the qubits the next
 * pass works on (which needs passes that only permute at the end of a run) */
#define NQPU_RUN_FIXED_LOW 64u
/* ignore plans measured by nqpu_autotune; use the fixed defaults */
#define NQPU_RUN_UNTUNED 128u
/* return once every pass of the run is queued on the state's stream: the
 * host can plan the next run while this one executes.  Every later call on
 * the state is ordered after it; nqpu_state_synchronize waits for it.  With
 * NQPU_RUN_TIME_PASSES, pass_ms then covers all runs queued since the last
 * synchronisation (read it after nqpu_state_synchronize). */
#define NQPU_RUN_ASYNC 256u
/* run the circuit from |0...0> instead of from the state's current contents
 * (the same as nqpu_state_reset first, without its separate fill pass and
 * synchronisation): the first tile pass generates the start state instead of
 * reading it.  A tiled complex64 run may leave the state's qubits in the
 * layout its last pass stored them in; every call that reads or writes the
 * state by index restores the canonical layout first, so the layout is never
 * visible through this interface. */
#define NQPU_RUN_FROM_ZERO 512u

/* Statistics of the last run on a state. */
typedef struct {
  uint64_t gates;           /* logical gates in the circuit */
  uint64_t device_ops;      /* gate operations after lowering and 1q merging */
  uint64_t passes;          /* full-state HBM sweeps (kernel launches) */
  double bytes_moved;       /* algorithmic HBM bytes of those sweeps */
  double pass_ms;           /* with NQPU_RUN_TIME_PASSES: summed kernel time of the passes */
  uint64_t timed_passes;    /* passes that pass_ms covers */
  uint64_t remaps;          /* LDS remaps inside register-resident tile passes */
  uint64_t mfma_blocks;     /* fused 4-qubit blocks applied on the matrix cores */
  uint64_t mfma_gates;      /* device ops those blocks absorbed */
  uint64_t exchanges;       /* register <-> lane-bit exchanges inside register tile passes (no LDS) */
  uint64_t qphase_gates;    /* diagonal device ops merged into quadratic phase ops */
} NqpuRunStats;

/* 1 <= n <= 34 (c64) or 33 (c128), bounded by free device memory. */
NqpuState* nqpu_state_create_ex(uint32_t num_qubits, uint32_t precision);
uint32_t nqpu_state_precision(const NqpuState* state);

/* Append any gate to a circuit; validates qubits and angles. */
NqpuError nqpu_circuit_add_gate(NqpuCircuit* circuit, const NqpuGateOp* op);
NqpuError nqpu_circuit_add_gates(NqpuCircuit* circuit, const NqpuGateOp* ops, uint64_t count);
/* nqpu_circuit_create without the 30-qubit cap (n <= 34). */
NqpuCircuit* nqpu_circuit_create_ex(uint32_t num_qubits);

/* Execute with explicit fusion/semantics options (NULL = defaults). */
NqpuError nqpu_run(const NqpuCircuit* circuit, NqpuState* state, const NqpuRunOptions* opts);
NqpuError nqpu_last_run_stats(const NqpuState* state, NqpuRunStats* out);
/* Wait for the runs queued on the state with NQPU_RUN_ASYNC. */
NqpuError nqpu_state_synchronize(NqpuState* state);

/* Run graphs: a tiled run recorded once as a HIP graph and replayed.
 * No counterpart in the reference, whose simulators dispatch every gate from
 * the host (QuantumSimulator, sdk/rust/src/lib.rs:1738-2070); this serves
 * callers that run the same circuit structure many times: variational loops,
 * repeated Trotter steps, a service replaying a circuit template.
 * nqpu_graph_create plans `circuit` (NQPU_FUSION_TILED) for `state` and
 * records its passes (after a |0...0> fill with NQPU_GRAPH_FROM_ZERO) into one
 * HIP graph on the state's stream.  The graph reads its register programs and
 * operand blobs from a device buffer it owns.  nqpu_graph_update plans another
 * circuit on the same qubits (new angles, for one) and copies its numbers into
 * that buffer in stream order, after the replays queued before it; the graph
 * is recorded again only when the new plan launches other kernels or the same
 * kernels with other arguments.  nqpu_graph_launch replays the graph on the
 * state's stream (NQPU_RUN_ASYNC: return once queued).  A graph belongs to
 * its state, which must outlive it; like the state, it is not thread-safe. */
typedef struct NqpuRunGraph NqpuRunGraph;
#define NQPU_GRAPH_FROM_ZERO 1u /* the graph first sets the state to |0...0> */
typedef struct {
  uint64_t launches;    /* kernel launches one replay makes (passes, plus the |0...0> fill) */
  uint64_t passes;      /* full-state gate passes among them */
  uint64_t replays;     /* nqpu_graph_launch calls */
  uint64_t updates;     /* nqpu_graph_update calls */
  uint64_t recaptures;  /* updates that recorded the graph again */
  uint64_t image_bytes; /* device buffer of register programs and operand blobs */
  double plan_ms;       /* host planning time of the last create or update */
} NqpuGraphInfo;
NqpuError nqpu_graph_create(const NqpuCircuit* circuit, NqpuState* state, const NqpuRunOptions* opts,
                            uint32_t graph_flags, NqpuRunGraph** out);
NqpuError nqpu_graph_update(NqpuRunGraph* graph, const NqpuCircuit* circuit);
NqpuError nqpu_graph_launch(NqpuRunGraph* graph, uint32_t flags);
NqpuError nqpu_graph_info(const NqpuRunGraph* graph, NqpuGraphInfo* out);
void nqpu_graph_destroy(NqpuRunGraph* graph);

/* FP64 reductions on the device. */
NqpuError nqpu_expectation_z_all(const NqpuState* state, double* out, uint32_t len);
/* <P> for the Pauli string with X on xmask&~zmask, Y on xmask&zmask, Z on
 * zmask&~xmask (qubit q = bit q). */
NqpuError nqpu_expectation_pauli(const NqpuState* state, uint64_t xmask, uint64_t zmask, double* out);
/* sum_k coeffs[k] * <P_k> */
NqpuError nqpu_expectation_pauli_sum(const NqpuState* state, const uint64_t* xmasks,
                                     const uint64_t* zmasks, const double* coeffs,
                                     uint64_t count, double* out);
/* out[t] = <P_t> for count Pauli strings (masks as nqpu_expectation_pauli).
 * Terms whose X/Y masks are equal share one pass over the state (all
 * diagonal strings of a Hamiltonian in one pass; up to 32 terms per pass);
 * nqpu_expectation_pauli_sum evaluates its terms the same way.  The
 * reference evaluates each string in its own loop
 * (expectation_hamiltonian, sdk/rust/src/lib.rs:518). */
NqpuError nqpu_expectation_pauli_values(const NqpuState* state, const uint64_t* xmasks, const uint64_t* zmasks,
                                        uint64_t count, double* out);

/* nqpu_expectation_pauli_sum with every device buffer in a fixed array of
 * the library (no counterpart in the reference).  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): it calls no hipMalloc/hipFree and no
 * allocating routine.  At most 1024 terms (more -> NQPU_INVALID_PARAMETER).
 * Every call on one device shares that array: calls that overlap on one
 * device race and must be serialised by the caller. */
NqpuError nqpu_expectation_pauli_sum_noalloc(const NqpuState* state, const uint64_t* xmasks,
                                             const uint64_t* zmasks, const double* coeffs, uint64_t count,
                                             double* out);

/* Qubit order of a Pauli string. */
typedef enum {
  /* character i acts on qubit i = index bit i (expectation_pauli_string,
   * sdk/rust/src/lib.rs:465-514) */
  NQPU_ORDER_LSB = 0,
  /* character i acts on qubit i of the Estimator's convention, index bit
   * n-1-i (Observable::from_string + pauli_expectation,
   * sdk/rust/src/core/foundations/primitives.rs:111-133, 683-728) */
  NQPU_ORDER_MSB = 1
} NqpuQubitOrder;

/* <P> for a string of 'I', 'X', 'Y', 'Z' (either case) of length <= n.
 * Longer strings -> NQPU_INVALID_QUBIT, other characters ->
 * NQPU_INVALID_PARAMETER (the reference panics).  The value is the physical
 * <psi|P|psi>; lib.rs:465-514 returns its negative for an odd number of Y
 * factors, which NQPU_SEMANTICS_REFERENCE reproduces nowhere here. */
NqpuError nqpu_expectation_pauli_string(const NqpuState* state, const char* paulis, uint32_t order, double* out);

/* The Estimator's exact value of an observable (Estimator::run with
 * shots = None, estimate_observable, primitives.rs:1021-1083): sum over terms
 * t of coeffs[t] * pauli_expectation(term t) (primitives.rs:683-728).  Term t
 * is the factors (qubits[k], paulis[k]) for k in [offsets[t], offsets[t+1]),
 * qubit 0 being the most significant index bit; a state built by a circuit
 * whose qubit q is this engine's qubit n-1-q.  Factors are applied as the
 * reference does (each against the original index bits), so repeated
 * qubits compose as there. */
NqpuError nqpu_estimator_expectation(const NqpuState* state, const uint32_t* qubits, const char* paulis,
                                     const uint64_t* offsets, const double* coeffs, uint64_t num_terms,
                                     double* out);

/* reset_qubit (sdk/rust/src/lib.rs:1705-1730): zero every amplitude whose
 * bit `qubit` is 1, then divide by the square root of the remaining mass
 * when it is positive (FP64, on the device). */
NqpuError nqpu_state_reset_qubit(NqpuState* state, uint32_t qubit);
/* initialize_from_amplitudes (sdk/rust/src/lib.rs:1918-1935): len must be
 * 2^n and the FP64 norm at least 1e-10, else NQPU_INVALID_PARAMETER and the
 * state is unchanged; the state becomes amplitudes / ||amplitudes||. */
NqpuError nqpu_state_initialize(NqpuState* state, const NqpuComplex* amplitudes, uint64_t len);

/* Scale the state to unit norm (FP64 norm). */
NqpuError nqpu_state_normalize(NqpuState* state);
/* Copy src's amplitudes into dst (same qubit count and precision), one
 * device-side pass. */
NqpuError nqpu_state_copy(NqpuState* dst, const NqpuState* src);
/* <a|b> = sum conj(a_i) b_i in FP64 over two states of equal qubit count
 * (any precisions); |<a|b>|^2 is QuantumState::fidelity (lib.rs:388-403). */
NqpuError nqpu_state_overlap(const NqpuState* a, const NqpuState* b, NqpuComplex* out);

/* Bulk transfer of all 2^n amplitudes as complex128 (len >= 2^n). */
NqpuError nqpu_state_get_amplitudes(const NqpuState* state, NqpuComplex* out, uint64_t len);
NqpuError nqpu_state_set_amplitudes(NqpuState* state, const NqpuComplex* in, uint64_t len);

/* The passes nqpu_run would launch for `circuit` on a state of the given
 * precision, computed on the host (no device needed). */
NqpuError nqpu_circuit_plan(const NqpuCircuit* circuit, const NqpuRunOptions* opts,
                            uint32_t precision, NqpuRunStats* out);

/* Counters of the reference fusion pass (FusionResult, gate_fusion.rs:96-101). */
typedef struct {
  uint64_t original_count;
  uint64_t fused_count;
  uint64_t gates_eliminated;
  uint64_t items;
} NqpuFusionCounts;
NqpuError nqpu_circuit_fusion_counts(const NqpuCircuit* circuit, NqpuFusionCounts* out);

/* ------------------------------------------------------------ batching */

/* One weighted Pauli string: X on xmask&~zmask, Y on xmask&zmask, Z on
 * zmask&~xmask. */
typedef struct {
  uint64_t xmask;
  uint64_t zmask;
  double coeff;
} NqpuPauliTerm;

typedef struct NqpuBatch NqpuBatch;

typedef struct {
  uint64_t circuits;
  uint64_t resident;        /* circuits held entirely in registers/LDS */
  uint64_t gate_passes;     /* tile passes over HBM-resident states */
  uint64_t observable_passes;
  double plan_ms;           /* host planning and queueing (the upload is asynchronous) */
  double kernel_ms;         /* the persistent launch and the launches of split circuits */
  double state_bytes;       /* HBM for the states of circuits above 12 qubits: one slot of the
                               largest such circuit below 18 qubits per workgroup of the persistent
                               launch (nqpu_debug_batch_plan: one slot), plus the states of the
                               circuits of 18 qubits or more, which run their passes as grids of
                               work items beside the persistent launch, in waves of at most 8 GiB
                               of states */
  double hbm_bytes;         /* algorithmic state traffic of the launch: sum over circuits above 12
                               qubits of (2 x gate passes - 1 + observable passes) x 2^n x 8 bytes
                               (the first gate pass only writes) */
} NqpuBatchStats;

/* A reusable batch context (stream and device buffers).  Not thread-safe. */
NqpuBatch* nqpu_batch_create(void);
void nqpu_batch_destroy(NqpuBatch* batch);
/* Run each circuit k from |0...0> as a complex64 state and write
 * sum_t coeff_t <P_t> (FP64) for its terms terms[offsets[k] .. offsets[k+1])
 * to out[k].  All circuits run in one persistent launch; every result is
 * independent of which other circuits share the batch.  Circuits need at
 * least 4 qubits; the jax_bridge vmap_simulate map (serial) is what this
 * batches (sdk/rust/src/quantum_ml/bridges/jax_bridge.rs:447).
 * With NQPU_RUN_ASYNC in opts->flags the call returns once the batch is
 * queued: the host can plan the next batch while this one runs (two batches
 * are kept in flight; a third call first completes the oldest).  `out` is
 * written when nqpu_batch_synchronize returns (or a later call completes the
 * batch) and must stay valid until then; circuits and terms may be released
 * as soon as the call returns. */
NqpuError nqpu_batch_expectation(NqpuBatch* batch, const NqpuCircuit* const* circuits, uint64_t count,
                                 const NqpuPauliTerm* terms, const uint64_t* offsets,
                                 const NqpuRunOptions* opts, double* out);
/* Wait for the batches queued with NQPU_RUN_ASYNC and write their results. */
NqpuError nqpu_batch_synchronize(NqpuBatch* batch);
NqpuError nqpu_batch_last_stats(const NqpuBatch* batch, NqpuBatchStats* out);
/* Host half of nqpu_batch_expectation only (planning, no device): fills the
 * pass counts of `out` and the number of register-program operations. */
NqpuError nqpu_debug_batch_plan(const NqpuCircuit* const* circuits, uint64_t count, const NqpuPauliTerm* terms,
                                const uint64_t* offsets, const NqpuRunOptions* opts, NqpuBatchStats* out,
                                uint64_t* program_ops);

/* Host-only emulation of nqpu_run's tiled execution for a state of the given
 * precision: the same plan and per-pass programs, interpreted on the CPU in
 * FP64 over `amps` (len >= 2^n, updated in place).  For testing the planner
 * without a device; slow. */
NqpuError nqpu_debug_emulate(const NqpuCircuit* circuit, const NqpuRunOptions* opts, uint32_t precision,
                             NqpuComplex* amps, uint64_t len);

/* Host-only check of NQPU_RUN_COMPILE for a complex64 state: generate the
 * straight-line kernels of every register tile pass of `circuit` and compile
 * them with hipRTC (nothing is loaded or cached; no device needed).
 * NQPU_ERROR_DEVICE_ERROR with the compiler log in nqpu_last_error_detail()
 * when compilation fails. */
NqpuError nqpu_debug_compile_passes(const NqpuCircuit* circuit, const NqpuRunOptions* opts);
/* The same for a state of the given precision (complex128: register tiles
 * of 11-13 qubits). */
NqpuError nqpu_debug_compile_passes_ex(const NqpuCircuit* circuit, const NqpuRunOptions* opts, uint32_t precision);

/* The state's current qubit layout: positions[q] = the bit of an
 * amplitude's position in device memory that holds qubit q (q itself in the
 * canonical layout; see NQPU_RUN_FROM_ZERO).  len >= the state's qubits. */
NqpuError nqpu_debug_state_layout(const NqpuState* state, int32_t* positions, uint32_t len);

/* Detail string of the last error on the calling thread. */
const char* nqpu_last_error_detail(void);

/* Number of usable devices (0 when none or when NQPU_HIP_FORCE_NO_DEVICE=1). */
uint32_t nqpu_device_count(void);

/* ---- Request server (continuous batching) ----
 * Replaces the one-request-at-a-time serving path of ExecuteResponse::execute
 * and SampleResponse::sample (sdk/rust/src/web/mod.rs:288-419), which
 * re-simulates the circuit for every shot.  Each request is a circuit run from
 * |0...0>, a list of weighted Pauli strings and a number of shots.
 * Requests are admitted in arrival order as soon as a slot and room in the
 * server's HBM arena are free.  Admitted requests run concurrently, one per
 * slot, each on its own stream.  A request's results do not depend on which
 * other requests share the device: its shots come from its own seeded
 * xorshift64* stream (seed 0 picks the C ABI's default state) and its
 * reductions are summed in a fixed order.
 * Every call may be made from any thread. */
typedef struct NqpuServer NqpuServer;

typedef struct {
  uint32_t slots;          /* requests on the device at once (0 = 8, at most 64) */
  uint32_t precision;      /* NqpuPrecision of every request's state (0 = complex64) */
  uint64_t memory_bytes;   /* HBM reserved for request states (0 = half the free memory) */
  NqpuRunOptions run;      /* how each circuit runs (all zero = defaults: tiled, standard) */
} NqpuServerConfig;

typedef enum {
  NQPU_REQUEST_QUEUED = 0,
  NQPU_REQUEST_RUNNING = 1,
  NQPU_REQUEST_DONE = 2,
  NQPU_REQUEST_FAILED = 3
} NqpuRequestState;

typedef struct {
  uint32_t state;      /* NqpuRequestState */
  int32_t error;       /* NqpuError of a failed request */
  double queued_ms;    /* submission to admission (so far, while queued) */
  double run_ms;       /* admission to completion (so far, while running) */
  double latency_ms;   /* submission to completion (so far, while unfinished) */
} NqpuRequestStatus;

typedef struct {
  uint64_t submitted, completed, failed;
  uint64_t peak_running;   /* most requests admitted at once */
  uint64_t peak_bytes;     /* most arena bytes held by admitted requests at once */
  uint64_t memory_bytes;   /* size of the arena */
} NqpuServerStats;

/* NULL on failure (nqpu_last_error_detail says why). */
NqpuServer* nqpu_server_create(const NqpuServerConfig* config);
/* Finishes the running requests, fails the queued ones, frees the arena. */
void nqpu_server_destroy(NqpuServer* server);
/* Queue a request; the circuit and terms are copied.  A state larger than the
 * whole arena is refused with NQPU_ERROR_OUT_OF_DEVICE_MEMORY. */
NqpuError nqpu_server_submit(NqpuServer* server, const NqpuCircuit* circuit, const NqpuPauliTerm* terms,
                             uint32_t num_terms, uint32_t shots, uint64_t seed, uint64_t* request_id);
NqpuError nqpu_server_poll(NqpuServer* server, uint64_t request_id, NqpuRequestStatus* out);
/* Block until the request finishes (timeout_ms = UINT32_MAX: no limit);
 * NQPU_ERROR_SIMULATION_FAILED when the timeout passes first. */
NqpuError nqpu_server_wait(NqpuServer* server, uint64_t request_id, uint32_t timeout_ms, NqpuRequestStatus* out);
/* Copy a finished request's results (coeff * <P_k> per term; sampled basis
 * indices) and forget the request.  A failed request returns its error. */
NqpuError nqpu_server_take(NqpuServer* server, uint64_t request_id, double* expectations, uint32_t num_terms,
                           uint64_t* samples, uint32_t shots);
NqpuError nqpu_server_stats(const NqpuServer* server, NqpuServerStats* out);

/* ---- Plan autotuning ----
 * The reference's tuner targets Apple GPUs (backends/runtime/auto_tuning.rs);
 * this one times the engine's plan choices on this device: tile qubits, the
 * contiguous low qubits of every tile, compiled or interpreted pass
 * kernels, fixed or moving low positions, and tiles per workgroup (the
 * grid).  The fastest becomes the default plan for states of that qubit
 * count and precision in this process (runs whose options leave tile_qubits,
 * low_qubits and those flags at zero, without NQPU_RUN_UNTUNED), and is
 * appended to the file named by NQPU_TUNE_CACHE when set (read back by
 * later processes).  Results of any plan agree to rounding. */
typedef struct {
  uint32_t tile_qubits;
  uint32_t low_qubits;
  uint32_t flags;            /* NQPU_RUN_COMPILE / NQPU_RUN_FIXED_LOW of the chosen plan */
  uint32_t tiles_per_group;  /* register tile passes: tiles per workgroup */
  double ms;                 /* the chosen plan's time for the probe circuit */
  double default_ms;         /* the fixed default plan's time for it */
} NqpuTunedPlan;

/* Time the candidates for an n-qubit state with `probe` (NULL: a random
 * U3 + CNOT brickwork of depth 8; else its qubit count must be n); n >= 11. */
NqpuError nqpu_autotune(uint32_t num_qubits, uint32_t precision, const NqpuCircuit* probe, NqpuTunedPlan* out);
/* The tuned plan for (num_qubits, precision), NQPU_INVALID_PARAMETER if none. */
NqpuError nqpu_tuned_plan(uint32_t num_qubits, uint32_t precision, NqpuTunedPlan* out);
/* Read plans from a file in the NQPU_TUNE_CACHE format (lines "n precision
 * tile low flags tiles_per_group ms default_ms"). */
NqpuError nqpu_tune_load(const char* path);
/* Forget every tuned plan of this process. */
void nqpu_tune_reset(void);

/* ---- Request formats (JSON) ----
 * A request is either the web API's CircuitRequest / SampleRequest
 * (sdk/rust/src/web/mod.rs:34-131, 149-160): {"num_qubits": n, "gates":
 * [{"type": "h", "target": 0}, {"type": "cnot", "control": 0, "target": 1},
 * ...], "backend": "auto", "shots": 1024}; or circuit_serde's
 * SerializableCircuit (sdk/rust/src/circuits/formats/circuit_serde.rs:
 * 329-340, 870-930): {"version": 1, "num_qubits": n, "gates": [{"gate_type":
 * "CNOT", "targets": [1], "controls": [0], "params": []}]}.  Either may add
 * "shots" and "observables": [{"pauli": "ZZI", "coeff": -1.0}] (character q
 * acts on qubit q).  Malformed requests -> NQPU_INVALID_PARAMETER (or
 * NQPU_INVALID_QUBIT for qubit numbers) with the reason in
 * nqpu_last_error_detail(). */

/* Parse a request into a new circuit (destroy with nqpu_circuit_destroy);
 * *shots = its "shots" (0 when absent; shots may be NULL). */
NqpuError nqpu_circuit_from_json(const char* json, NqpuCircuit** circuit, uint32_t* shots);
/* ExecuteResponse::execute (web/mod.rs:288-375): run the request from
 * |0...0> (precision 0 = complex128) and write {"probabilities": [...],
 * "time_ms", "backend_used", "num_gates", "fidelity"} (fidelity of the fused
 * run against the one-pass-per-gate run) as a NUL-terminated string.
 * *needed = the bytes this call requires (an upper bound, known before the
 * run); a smaller buffer -> NQPU_INVALID_PARAMETER without running. */
NqpuError nqpu_execute_json(const char* request, uint32_t precision, char* response, uint64_t len,
                            uint64_t* needed);
/* SampleResponse::sample (web/mod.rs:378-419) on the request server: the
 * request needs "shots" or "observables"; its shots come from its own
 * xorshift64* stream seeded by `seed`. */
NqpuError nqpu_server_submit_json(NqpuServer* server, const char* request, uint64_t seed, uint64_t* request_id);
/* Wait (up to timeout_ms; UINT32_MAX = no limit) and write {"outcomes":
 * {"<MSB-first bitstring>": count, ...}, "shots", "time_ms", "latency_ms",
 * "backend_used", "expectations": [...] (when observables were given)}.
 * *needed = the exact size; with a smaller buffer the response is kept for
 * the next call (NQPU_INVALID_PARAMETER); otherwise the request is
 * forgotten. */
NqpuError nqpu_server_take_json(NqpuServer* server, uint64_t request_id, uint32_t timeout_ms, char* response,
                                uint64_t len, uint64_t* needed);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* NQPU_HIP_H */
