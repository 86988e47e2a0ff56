/*
 * cpu_ref — FP64 CPU oracle for the nQPU dense statevector path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (nqpu_amd, libnqpu_hip.so) never links, loads or calls it.
 *
 * It restates, in plain C99 + OpenMP, the complex128 engine of the reference:
 *   sdk/rust/src/lib.rs:326-1732          QuantumState, GateOperations, sample
 *   sdk/rust/src/core/gates/gates.rs:70-598  GateType::matrix()
 *   sdk/rust/src/circuits/visualization/ascii_viz.rs:124-179  apply_gate_to_state
 *   sdk/rust/src/infra/bindings/c_ffi.rs:48-67, 211-560  C ABI RNG, measure, sample
 *   sdk/rust/src/circuits/synthesis/gate_fusion.rs:905-1033  execute_diagonal
 *
 * Amplitudes are interleaved (re, im) doubles, LSB qubit order (qubit q is bit
 * q of the basis index), exactly as QuantumState (lib.rs:326-346).
 *
 * Two semantics are selectable per call:
 *   quirk = 0  "standard": every gate is the matrix of gates.rs, including the
 *              gates the reference dispatcher drops (Rxx/Ryy/Rzz/CSWAP/CU/Custom).
 *   quirk = 1  "reference": what the reference code actually does — CRx applies
 *              a controlled-Rz (lib.rs:1292-1358), CRz only rotates |c=1,t=0> by
 *              e^{+i theta/2} (parallel branch, lib.rs:1439-1473), the dropped gates
 *              are no-ops (ascii_viz.rs:166-177), and <Y>-bearing Pauli strings
 *              carry the sign of lib.rs:444-514.
 */
#ifndef NQPU_ORACLE_CPU_REF_H
#define NQPU_ORACLE_CPU_REF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } orc_c;

/* Gate kinds: same values as NqpuGateKind in include/nqpu_hip.h (the order of
 * GateType in gates.rs:8-64). */
enum {
  ORC_H = 0, ORC_X, ORC_Y, ORC_Z, ORC_S, ORC_T, ORC_RX, ORC_RY, ORC_RZ, ORC_U,
  ORC_CNOT, ORC_CZ, ORC_SWAP, ORC_TOFFOLI, ORC_CRX, ORC_CRY, ORC_CRZ, ORC_CR,
  ORC_SX, ORC_PHASE, ORC_ISWAP, ORC_CCZ, ORC_RXX, ORC_RYY, ORC_RZZ, ORC_CSWAP,
  ORC_CU, ORC_CUSTOM1, ORC_CUSTOM2, ORC_NUM_KINDS
};

/* Mirrors gates::Gate (gates.rs:632): targets, controls, parameters, and the
 * matrix of a Custom gate (row-major, 2x2 or 4x4, interleaved re/im). */
typedef struct {
  int32_t kind;
  int32_t nt, nc;
  int32_t t[3];
  int32_t c[3];
  int32_t pad;
  double p[4];
  double m[32];
} orc_gate;

/* state */
void orc_init_zero(orc_c* a, int n);
/* GateType::matrix() (gates.rs:70-598); *dim receives 2, 4 or 8. */
void orc_gate_matrix(const orc_gate* g, orc_c* out, int* dim);
/* GateOperations::u (lib.rs:1493) and u2 (lib.rs:1536; sub-index (b_hi<<1)|b_lo). */
void orc_u1(orc_c* a, int n, int q, const orc_c* m);
void orc_u2(orc_c* a, int n, int qlo, int qhi, const orc_c* m);
/* apply_gate_to_state (ascii_viz.rs:124-179) in the chosen semantics. */
void orc_apply_gate(orc_c* a, int n, const orc_gate* g, int quirk);
void orc_apply_circuit(orc_c* a, int n, const orc_gate* g, int count, int quirk);
/* execute_diagonal (gate_fusion.rs:905): single phases (q,p0,p1) and pair phases (lo,hi,p11). */
void orc_execute_diagonal(orc_c* a, int n, const int32_t* sq, const orc_c* sp, int ns,
                          const int32_t* pq, const orc_c* pp, int np);

/* reductions (lib.rs:354-524, c_ffi.rs:243-249) */
double orc_norm(const orc_c* a, int n);
void orc_probabilities(const orc_c* a, int n, double* out);
double orc_expectation_z(const orc_c* a, int n, int q);
double orc_expectation_x(const orc_c* a, int n, int q);
double orc_expectation_y(const orc_c* a, int n, int q, int quirk);
/* Pauli string as masks: qubit q carries X if x&~z, Y if x&z, Z if z&~x. */
double orc_expectation_pauli(const orc_c* a, int n, uint64_t xmask, uint64_t zmask, int quirk);
double orc_fidelity(const orc_c* a, const orc_c* b, int n);

/* C-ABI RNG and sampling (c_ffi.rs:48-67, 1444-1452, 506-525, 473-500) */
void orc_rng_seed(uint64_t seed);
double orc_rng_next(void);
uint64_t orc_rng_state(void);
void orc_sample_cabi(const orc_c* a, int n, uint32_t shots, uint64_t* out);
/* lib.rs:527-550 with caller-supplied uniforms r[k] (first index with cdf >= r). */
void orc_sample_cdf(const orc_c* a, int n, const double* r, uint64_t shots, uint64_t* out);
/* c_ffi.rs:473-500: measure one qubit with the C-ABI RNG, collapse, renormalise. */
uint32_t orc_measure_qubit(orc_c* a, int n, int q, double* prob);

/* lib.rs:1705-1730, lib.rs:1918-1935, primitives.rs:683-728 */
void orc_reset_qubit(orc_c* a, int n, int q);
int orc_initialize_from_amplitudes(orc_c* dst, int n, const orc_c* src, uint64_t len);
double orc_pauli_expectation_msb(const orc_c* a, int n, const int32_t* qubits, const char* paulis, int count,
                                 double coeff);

int orc_num_threads(void);
void orc_set_num_threads(int t);

#ifdef __cplusplus
}
#endif
#endif
