// capi.hip — the extern "C" boundary (include/nqpu_hip.h).
//
// Argument checks and error codes follow the reference's C ABI
// (sdk/rust/src/infra/bindings/c_ffi.rs:788-1467) in the same order, so a
// caller sees the same code for the same mistake.
#include "runtime.h"

#include <cmath>
#include <algorithm>
#include <cstring>
#include <new>

using namespace nqpu;

namespace {

State* S(NqpuState* p) { return reinterpret_cast<State*>(p); }
const State* S(const NqpuState* p) { return reinterpret_cast<const State*>(p); }
Circuit* Cc(NqpuCircuit* p) { return reinterpret_cast<Circuit*>(p); }
const Circuit* Cc(const NqpuCircuit* p) { return reinterpret_cast<const Circuit*>(p); }

Gate make_gate(int kind, std::initializer_list<int> targets, std::initializer_list<int> controls,
               std::initializer_list<double> params) {
  Gate g;
  g.kind = kind;
  for (int q : targets) g.t[g.nt++] = q;
  for (int q : controls) g.c[g.nc++] = q;
  int i = 0;
  for (double v : params) g.p[i++] = v;
  return g;
}

NqpuError gate1(NqpuState* st, uint32_t q, int kind) {
  if (!st) return NQPU_NULL_POINTER;
  State* s = S(st);
  if ((int)q >= s->n) return NQPU_INVALID_QUBIT;
  return apply_gate_now(s, make_gate(kind, {(int)q}, {}, {}));
}

NqpuError gate1p(NqpuState* st, uint32_t q, double angle, int kind) {
  if (!st) return NQPU_NULL_POINTER;
  State* s = S(st);
  if ((int)q >= s->n) return NQPU_INVALID_QUBIT;
  if (!std::isfinite(angle)) return NQPU_INVALID_PARAMETER;
  return apply_gate_now(s, make_gate(kind, {(int)q}, {}, {angle}));
}

NqpuError gate2(NqpuState* st, uint32_t a, uint32_t b, int kind) {
  if (!st) return NQPU_NULL_POINTER;
  State* s = S(st);
  if ((int)a >= s->n) return NQPU_INVALID_QUBIT;
  if ((int)b >= s->n) return NQPU_INVALID_QUBIT;
  if (a == b) return NQPU_INVALID_PARAMETER;
  if (kind == NQPU_GATE_SWAP) return apply_gate_now(s, make_gate(kind, {(int)a, (int)b}, {}, {}));
  return apply_gate_now(s, make_gate(kind, {(int)b}, {(int)a}, {}));
}

// Number of qubits, parameters and matrices a gate kind needs, and its
// validation (qubit range, distinct qubits, finite parameters).
NqpuError validate_op(const NqpuGateOp* op, uint32_t n, Gate* out) {
  if (op->kind >= NQPU_GATE_NUM_KINDS) return NQPU_INVALID_PARAMETER;
  static const int need_t[NQPU_GATE_NUM_KINDS] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 1, 1,
                                                  1, 1, 1, 1, 1, 2, 1, 2, 2, 2, 2, 1, 1, 2};
  static const int need_c[NQPU_GATE_NUM_KINDS] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 2, 1,
                                                  1, 1, 1, 0, 0, 0, 2, 0, 0, 0, 1, 1, 0, 0};
  static const int need_p[NQPU_GATE_NUM_KINDS] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 3, 0, 0, 0, 0, 1,
                                                  1, 1, 1, 0, 1, 0, 0, 1, 1, 1, 0, 4, 0, 0};
  const int k = (int)op->kind;
  if ((int)op->num_targets != need_t[k] || (int)op->num_controls != need_c[k]) return NQPU_INVALID_PARAMETER;
  Gate g;
  g.kind = k;
  g.nt = need_t[k];
  g.nc = need_c[k];
  uint64_t seen = 0;
  for (int i = 0; i < g.nt; ++i) {
    if (op->targets[i] >= n) return NQPU_INVALID_QUBIT;
    g.t[i] = (int)op->targets[i];
  }
  for (int i = 0; i < g.nc; ++i) {
    if (op->controls[i] >= n) return NQPU_INVALID_QUBIT;
    g.c[i] = (int)op->controls[i];
  }
  for (int i = 0; i < g.nt; ++i) {
    if (seen >> g.t[i] & 1) return NQPU_INVALID_PARAMETER;
    seen |= 1ull << g.t[i];
  }
  for (int i = 0; i < g.nc; ++i) {
    if (seen >> g.c[i] & 1) return NQPU_INVALID_PARAMETER;
    seen |= 1ull << g.c[i];
  }
  for (int i = 0; i < need_p[k]; ++i) {
    if (!std::isfinite(op->params[i])) return NQPU_INVALID_PARAMETER;
    g.p[i] = op->params[i];
  }
  int nm = k == NQPU_GATE_CUSTOM1 ? 4 : (k == NQPU_GATE_CUSTOM2 ? 16 : 0);
  for (int i = 0; i < nm; ++i) {
    if (!std::isfinite(op->matrix[i].re) || !std::isfinite(op->matrix[i].im)) return NQPU_INVALID_PARAMETER;
    g.m[i] = cd(op->matrix[i].re, op->matrix[i].im);
  }
  *out = g;
  return NQPU_OK;
}

NqpuState* create(uint32_t n, uint32_t prec) {
  State* s = nullptr;
  try {
    if (create_state((int)n, (int)prec, &s) != NQPU_OK) return nullptr;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return nullptr;
  }
  return reinterpret_cast<NqpuState*>(s);
}

}  // namespace

extern "C" {

NqpuState* nqpu_state_create(uint32_t num_qubits) {
  if (num_qubits == 0 || num_qubits > 30) {
    set_error("num_qubits must be in [1, 30]");
    return nullptr;
  }
  return create(num_qubits, NQPU_PRECISION_C128);
}

NqpuState* nqpu_state_create_ex(uint32_t num_qubits, uint32_t precision) {
  if (precision != NQPU_PRECISION_C64 && precision != NQPU_PRECISION_C128) {
    set_error("precision must be NQPU_PRECISION_C64 or NQPU_PRECISION_C128");
    return nullptr;
  }
  uint32_t cap = precision == NQPU_PRECISION_C64 ? 34 : 33;
  if (num_qubits == 0 || num_qubits > cap) {
    set_error("num_qubits out of range for this precision");
    return nullptr;
  }
  return create(num_qubits, precision);
}

void nqpu_state_destroy(NqpuState* state) { delete S(state); }

NqpuState* nqpu_state_clone(const NqpuState* state) {
  if (!state) return nullptr;
  State* c = nullptr;
  if (clone_state(S(state), &c) != NQPU_OK) return nullptr;
  return reinterpret_cast<NqpuState*>(c);
}

uint32_t nqpu_state_num_qubits(const NqpuState* state) { return state ? (uint32_t)S(state)->n : 0; }

uint32_t nqpu_state_precision(const NqpuState* state) { return state ? (uint32_t)S(state)->prec : 0; }

NqpuError nqpu_state_get_amplitude(const NqpuState* state, uint64_t index, NqpuComplex* out) {
  if (!state) return NQPU_NULL_POINTER;
  if (!out) return NQPU_NULL_POINTER;
  if (index >= S(state)->dim()) return NQPU_INVALID_PARAMETER;
  return get_amplitudes(S(state), index, 1, out);
}

NqpuError nqpu_state_set_amplitude(NqpuState* state, uint64_t index, NqpuComplex amp) {
  if (!state) return NQPU_NULL_POINTER;
  if (index >= S(state)->dim()) return NQPU_INVALID_PARAMETER;
  return set_amplitudes(S(state), index, 1, &amp);
}

NqpuError nqpu_state_get_amplitudes(const NqpuState* state, NqpuComplex* out, uint64_t len) {
  if (!state || !out) return NQPU_NULL_POINTER;
  if (len < S(state)->dim()) return NQPU_INVALID_PARAMETER;
  return get_amplitudes(S(state), 0, S(state)->dim(), out);
}

NqpuError nqpu_state_set_amplitudes(NqpuState* state, const NqpuComplex* in, uint64_t len) {
  if (!state || !in) return NQPU_NULL_POINTER;
  if (len < S(state)->dim()) return NQPU_INVALID_PARAMETER;
  return set_amplitudes(S(state), 0, S(state)->dim(), in);
}

NqpuError nqpu_state_probabilities(const NqpuState* state, double* out, uint64_t len) {
  if (!state) return NQPU_NULL_POINTER;
  if (!out) return NQPU_NULL_POINTER;
  if (len < S(state)->dim()) return NQPU_INVALID_PARAMETER;
  return probabilities(S(state), out);
}

double nqpu_state_norm(const NqpuState* state) {
  if (!state) return -1.0;
  double v = 0;
  if (norm(S(state), &v) != NQPU_OK) return -1.0;
  return v;
}

NqpuError nqpu_state_reset(NqpuState* state) {
  if (!state) return NQPU_NULL_POINTER;
  return reset_state(S(state));
}

NqpuError nqpu_state_normalize(NqpuState* state) {
  if (!state) return NQPU_NULL_POINTER;
  double v = 0;
  NqpuError e = norm(S(state), &v);
  if (e != NQPU_OK) return e;
  if (!(v > 0)) return NQPU_SIMULATION_FAILED;
  return scale(S(state), 1.0 / std::sqrt(v));
}

NqpuError nqpu_state_copy(NqpuState* dst, const NqpuState* src) {
  if (!dst || !src) return NQPU_NULL_POINTER;
  State* d = S(dst);
  const State* s = S(src);
  if (d->n != s->n || d->prec != s->prec) return NQPU_INVALID_PARAMETER;
  NqpuError e = hip_check(hipStreamSynchronize(s->stream), "sync");
  if (e != NQPU_OK) return e;
  launch_copy(d->amps, s->amps, d->dim(), d->prec == NQPU_PRECISION_C64, d->stream);
  e = hip_check(hipGetLastError(), "copy launch");
  if (e != NQPU_OK) return e;
  d->perm = s->perm;  // the copy keeps the layout
  d->perm_k = s->perm_k;
  return hip_check(hipStreamSynchronize(d->stream), "copy");
}

NqpuError nqpu_state_overlap(const NqpuState* a, const NqpuState* b, NqpuComplex* out) {
  if (!a || !b || !out) return NQPU_NULL_POINTER;
  if (S(a)->n != S(b)->n) return NQPU_INVALID_PARAMETER;
  return overlap(S(a), S(b), out);
}

NqpuError nqpu_gate_h(NqpuState* s, uint32_t q) { return gate1(s, q, NQPU_GATE_H); }
NqpuError nqpu_gate_x(NqpuState* s, uint32_t q) { return gate1(s, q, NQPU_GATE_X); }
NqpuError nqpu_gate_y(NqpuState* s, uint32_t q) { return gate1(s, q, NQPU_GATE_Y); }
NqpuError nqpu_gate_z(NqpuState* s, uint32_t q) { return gate1(s, q, NQPU_GATE_Z); }
NqpuError nqpu_gate_s(NqpuState* s, uint32_t q) { return gate1(s, q, NQPU_GATE_S); }
NqpuError nqpu_gate_t(NqpuState* s, uint32_t q) { return gate1(s, q, NQPU_GATE_T); }
NqpuError nqpu_gate_rx(NqpuState* s, uint32_t q, double a) { return gate1p(s, q, a, NQPU_GATE_RX); }
NqpuError nqpu_gate_ry(NqpuState* s, uint32_t q, double a) { return gate1p(s, q, a, NQPU_GATE_RY); }
NqpuError nqpu_gate_rz(NqpuState* s, uint32_t q, double a) { return gate1p(s, q, a, NQPU_GATE_RZ); }
NqpuError nqpu_gate_cx(NqpuState* s, uint32_t c, uint32_t t) { return gate2(s, c, t, NQPU_GATE_CNOT); }
NqpuError nqpu_gate_cz(NqpuState* s, uint32_t c, uint32_t t) { return gate2(s, c, t, NQPU_GATE_CZ); }
NqpuError nqpu_gate_swap(NqpuState* s, uint32_t a, uint32_t b) { return gate2(s, a, b, NQPU_GATE_SWAP); }

NqpuError nqpu_gate_custom_1q(NqpuState* st, uint32_t q, const NqpuComplex* m) {
  if (!st) return NQPU_NULL_POINTER;
  if (!m) return NQPU_NULL_POINTER;
  State* s = S(st);
  if ((int)q >= s->n) return NQPU_INVALID_QUBIT;
  Gate g = make_gate(NQPU_GATE_CUSTOM1, {(int)q}, {}, {});
  for (int i = 0; i < 4; ++i) g.m[i] = cd(m[i].re, m[i].im);
  return apply_gate_now(s, g);
}

NqpuError nqpu_gate_custom_2q(NqpuState* st, uint32_t q1, uint32_t q2, const NqpuComplex* m) {
  if (!st) return NQPU_NULL_POINTER;
  if (!m) return NQPU_NULL_POINTER;
  State* s = S(st);
  if ((int)q1 >= s->n) return NQPU_INVALID_QUBIT;
  if ((int)q2 >= s->n) return NQPU_INVALID_QUBIT;
  if (q1 == q2) return NQPU_INVALID_PARAMETER;
  Gate g = make_gate(NQPU_GATE_CUSTOM2, {(int)q1, (int)q2}, {}, {});
  for (int i = 0; i < 16; ++i) g.m[i] = cd(m[i].re, m[i].im);
  return apply_gate_now(s, g);
}

NqpuError nqpu_measure(NqpuState* st, uint32_t q, NqpuMeasurement* out) {
  if (!st) return NQPU_NULL_POINTER;
  if (!out) return NQPU_NULL_POINTER;
  State* s = S(st);
  if ((int)q >= s->n) return NQPU_INVALID_QUBIT;
  uint32_t o;
  double p;
  NqpuError e = measure_qubit(s, (int)q, rng_next(), &o, &p);
  if (e != NQPU_OK) return e;
  out->outcome = o;
  out->probability = p;
  return NQPU_OK;
}

NqpuError nqpu_measure_all(NqpuState* st, uint32_t* out, uint32_t num_qubits) {
  if (!st) return NQPU_NULL_POINTER;
  if (!out) return NQPU_NULL_POINTER;
  State* s = S(st);
  if ((int)num_qubits != s->n) return NQPU_INVALID_PARAMETER;
  for (int q = 0; q < s->n; ++q) {
    uint32_t o;
    double p;
    NqpuError e = measure_qubit(s, q, rng_next(), &o, &p);
    if (e != NQPU_OK) return e;
    out[q] = o;
  }
  return NQPU_OK;
}

NqpuError nqpu_expectation_z(const NqpuState* st, uint32_t q, double* out) {
  if (!st) return NQPU_NULL_POINTER;
  if (!out) return NQPU_NULL_POINTER;
  if ((int)q >= S(st)->n) return NQPU_INVALID_QUBIT;
  double tot;
  std::vector<double> s1;
  NqpuError e = zsum_all(S(st), &tot, &s1);
  if (e != NQPU_OK) return e;
  *out = tot - 2.0 * s1[q];
  return NQPU_OK;
}

NqpuError nqpu_expectation_z_all(const NqpuState* st, double* out, uint32_t len) {
  if (!st) return NQPU_NULL_POINTER;
  if (!out) return NQPU_NULL_POINTER;
  if ((int)len < S(st)->n) return NQPU_INVALID_PARAMETER;
  double tot;
  std::vector<double> s1;
  NqpuError e = zsum_all(S(st), &tot, &s1);
  if (e != NQPU_OK) return e;
  for (int q = 0; q < S(st)->n; ++q) out[q] = tot - 2.0 * s1[q];
  return NQPU_OK;
}

NqpuError nqpu_expectation_pauli(const NqpuState* st, uint64_t xmask, uint64_t zmask, double* out) {
  if (!st) return NQPU_NULL_POINTER;
  if (!out) return NQPU_NULL_POINTER;
  int n = S(st)->n;
  uint64_t lim = n >= 64 ? ~0ull : ((1ull << n) - 1);
  if ((xmask | zmask) & ~lim) return NQPU_INVALID_QUBIT;
  return pauli(S(st), xmask, zmask, out);
}

namespace {
NqpuError check_masks(const NqpuState* st, const uint64_t* xm, const uint64_t* zm, uint64_t count) {
  const int n = S(st)->n;
  const uint64_t lim = n >= 64 ? ~0ull : ((1ull << n) - 1);
  for (uint64_t k = 0; k < count; ++k)
    if ((xm[k] | zm[k]) & ~lim) return NQPU_INVALID_QUBIT;
  return NQPU_OK;
}
}  // namespace

NqpuError nqpu_expectation_pauli_sum(const NqpuState* st, const uint64_t* xm, const uint64_t* zm,
                                     const double* coeffs, uint64_t count, double* out) {
  if (!st) return NQPU_NULL_POINTER;
  if (!out || (count && (!xm || !zm || !coeffs))) return NQPU_NULL_POINTER;
  NqpuError e = check_masks(st, xm, zm, count);
  if (e != NQPU_OK) return e;
  if (count == 0) {
    *out = 0.0;
    return NQPU_OK;
  }
  return pauli_sum(S(st), xm, zm, coeffs, count, out);
}

NqpuError nqpu_expectation_pauli_sum_noalloc(const NqpuState* st, const uint64_t* xm, const uint64_t* zm,
                                             const double* coeffs, uint64_t count, double* out) {
  if (!st) return NQPU_NULL_POINTER;
  if (!out || (count && (!xm || !zm || !coeffs))) return NQPU_NULL_POINTER;
  if (count > kPauliSumNoAllocMaxTerms) {
    set_error("nqpu_expectation_pauli_sum_noalloc: at most 1024 terms");
    return NQPU_INVALID_PARAMETER;
  }
  NqpuError e = check_masks(st, xm, zm, count);
  if (e != NQPU_OK) return e;
  return pauli_sum_noalloc(S(st), xm, zm, coeffs, count, out);
}

NqpuError nqpu_expectation_pauli_values(const NqpuState* st, const uint64_t* xm, const uint64_t* zm, uint64_t count,
                                        double* out) {
  if (!st) return NQPU_NULL_POINTER;
  if (count && (!out || !xm || !zm)) return NQPU_NULL_POINTER;
  NqpuError e = check_masks(st, xm, zm, count);
  if (e != NQPU_OK) return e;
  std::vector<PauliTerm> terms(count);
  for (uint64_t t = 0; t < count; ++t) terms[t] = PauliTerm{xm[t], zm[t], 1.0};
  std::vector<double> v;
  e = pauli_values(S(st), terms, &v);
  if (e != NQPU_OK) return e;
  std::copy(v.begin(), v.end(), out);
  return NQPU_OK;
}

NqpuError nqpu_expectation_pauli_string(const NqpuState* st, const char* paulis, uint32_t order, double* out) {
  if (!st || !paulis || !out) return NQPU_NULL_POINTER;
  if (order != NQPU_ORDER_LSB && order != NQPU_ORDER_MSB) return NQPU_INVALID_PARAMETER;
  const int n = S(st)->n;
  const size_t len = std::strlen(paulis);
  if (len > (size_t)n) return NQPU_INVALID_QUBIT;
  uint64_t x = 0, z = 0;
  for (size_t i = 0; i < len; ++i) {
    const int bitpos = order == NQPU_ORDER_LSB ? (int)i : n - 1 - (int)i;
    switch (paulis[i]) {
      case 'I': case 'i': break;
      case 'X': case 'x': x |= 1ull << bitpos; break;
      case 'Y': case 'y': x |= 1ull << bitpos; z |= 1ull << bitpos; break;
      case 'Z': case 'z': z |= 1ull << bitpos; break;
      default: return NQPU_INVALID_PARAMETER;
    }
  }
  return pauli(S(st), x, z, out);
}

NqpuError nqpu_estimator_expectation(const NqpuState* st, const uint32_t* qubits, const char* paulis,
                                     const uint64_t* offsets, const double* coeffs, uint64_t num_terms,
                                     double* out) {
  if (!st || !out) return NQPU_NULL_POINTER;
  if (num_terms && (!offsets || !coeffs)) return NQPU_NULL_POINTER;
  const int n = S(st)->n;
  if (num_terms && offsets[num_terms] > offsets[0] && (!qubits || !paulis)) return NQPU_NULL_POINTER;
  // A term's factors act on the index bits n-1-qubit.  They are applied as
  // the reference does, each against the bits of the original index: X and
  // Y flip, Y multiplies by i on |0> and -i on |1>, Z by -1 on |1>, so the
  // term is i^(#Y) times a flip mask and a sign mask (repeated qubits
  // compose in that way too).
  std::vector<PauliTerm> terms(num_terms);
  std::vector<int> phase(num_terms, 0);
  for (uint64_t t = 0; t < num_terms; ++t) {
    if (offsets[t + 1] < offsets[t]) return NQPU_INVALID_PARAMETER;
    uint64_t x = 0, z = 0;
    int ny = 0;
    for (uint64_t k = offsets[t]; k < offsets[t + 1]; ++k) {
      if (qubits[k] >= (uint32_t)n) return NQPU_INVALID_QUBIT;
      const uint64_t b = 1ull << (n - 1 - (int)qubits[k]);
      switch (paulis[k]) {
        case 'I': case 'i': break;
        case 'X': case 'x': x ^= b; break;
        case 'Y': case 'y': x ^= b; z ^= b; ++ny; break;
        case 'Z': case 'z': z ^= b; break;
        default: return NQPU_INVALID_PARAMETER;
      }
    }
    terms[t] = PauliTerm{x, z, coeffs[t]};
    phase[t] = ny & 3;
  }
  std::vector<double> v;
  NqpuError e = pauli_values(S(st), terms, &v, &phase);
  if (e != NQPU_OK) return e;
  double acc = 0;
  for (uint64_t t = 0; t < num_terms; ++t) acc += coeffs[t] * v[t];
  *out = acc;
  return NQPU_OK;
}

NqpuError nqpu_state_reset_qubit(NqpuState* st, uint32_t qubit) {
  if (!st) return NQPU_NULL_POINTER;
  if ((int)qubit >= S(st)->n) return NQPU_INVALID_QUBIT;
  return reset_qubit(S(st), (int)qubit);
}

NqpuError nqpu_state_initialize(NqpuState* st, const NqpuComplex* amplitudes, uint64_t len) {
  if (!st || !amplitudes) return NQPU_NULL_POINTER;
  return initialize_amplitudes(S(st), amplitudes, len);
}

NqpuError nqpu_sample(const NqpuState* st, uint32_t shots, uint64_t* out) {
  if (!st) return NQPU_NULL_POINTER;
  if (!out) return NQPU_NULL_POINTER;
  if (shots == 0) return NQPU_OK;
  std::vector<double> r(shots);
  for (uint32_t k = 0; k < shots; ++k) r[k] = rng_next();
  return sample(S(st), r, out);
}

NqpuCircuit* nqpu_circuit_create(uint32_t n) {
  if (n == 0 || n > 30) return nullptr;
  Circuit* c = new (std::nothrow) Circuit();
  if (!c) return nullptr;
  c->n = (int)n;
  return reinterpret_cast<NqpuCircuit*>(c);
}

NqpuCircuit* nqpu_circuit_create_ex(uint32_t n) {
  if (n == 0 || n > 34) return nullptr;
  Circuit* c = new (std::nothrow) Circuit();
  if (!c) return nullptr;
  c->n = (int)n;
  return reinterpret_cast<NqpuCircuit*>(c);
}

void nqpu_circuit_destroy(NqpuCircuit* c) { delete Cc(c); }

NqpuError nqpu_circuit_add_h(NqpuCircuit* c, uint32_t q) {
  if (!c) return NQPU_NULL_POINTER;
  if ((int)q >= Cc(c)->n) return NQPU_INVALID_QUBIT;
  Cc(c)->gates.push_back(make_gate(NQPU_GATE_H, {(int)q}, {}, {}));
  return NQPU_OK;
}

NqpuError nqpu_circuit_add_cx(NqpuCircuit* c, uint32_t ctl, uint32_t tgt) {
  if (!c) return NQPU_NULL_POINTER;
  if ((int)ctl >= Cc(c)->n || (int)tgt >= Cc(c)->n) return NQPU_INVALID_QUBIT;
  if (ctl == tgt) return NQPU_INVALID_PARAMETER;
  Cc(c)->gates.push_back(make_gate(NQPU_GATE_CNOT, {(int)tgt}, {(int)ctl}, {}));
  return NQPU_OK;
}

NqpuError nqpu_circuit_add_rx(NqpuCircuit* c, uint32_t q, double angle) {
  if (!c) return NQPU_NULL_POINTER;
  if ((int)q >= Cc(c)->n) return NQPU_INVALID_QUBIT;
  if (!std::isfinite(angle)) return NQPU_INVALID_PARAMETER;
  Cc(c)->gates.push_back(make_gate(NQPU_GATE_RX, {(int)q}, {}, {angle}));
  return NQPU_OK;
}

NqpuError nqpu_circuit_add_gate(NqpuCircuit* c, const NqpuGateOp* op) {
  if (!c || !op) return NQPU_NULL_POINTER;
  Gate g;
  NqpuError e = validate_op(op, (uint32_t)Cc(c)->n, &g);
  if (e != NQPU_OK) return e;
  Cc(c)->gates.push_back(g);
  return NQPU_OK;
}

NqpuError nqpu_circuit_add_gates(NqpuCircuit* c, const NqpuGateOp* ops, uint64_t count) {
  if (!c || (count && !ops)) return NQPU_NULL_POINTER;
  std::vector<Gate> tmp;
  tmp.reserve(count);
  for (uint64_t k = 0; k < count; ++k) {
    Gate g;
    NqpuError e = validate_op(&ops[k], (uint32_t)Cc(c)->n, &g);
    if (e != NQPU_OK) return e;
    tmp.push_back(g);
  }
  Cc(c)->gates.insert(Cc(c)->gates.end(), tmp.begin(), tmp.end());
  return NQPU_OK;
}

NqpuError nqpu_run(const NqpuCircuit* c, NqpuState* st, const NqpuRunOptions* opts) {
  if (!c) return NQPU_NULL_POINTER;
  if (!st) return NQPU_NULL_POINTER;
  if (Cc(c)->n > S(st)->n) return NQPU_INVALID_PARAMETER;
  NqpuRunOptions o{NQPU_FUSION_TILED, NQPU_SEMANTICS_STANDARD, 0, 0};
  if (opts) o = *opts;
  if (o.fusion > NQPU_FUSION_TILED || o.semantics > NQPU_SEMANTICS_REFERENCE) return NQPU_INVALID_PARAMETER;
  return run_circuit(S(st), *Cc(c), o);
}

NqpuError nqpu_circuit_execute(const NqpuCircuit* c, NqpuState* st) { return nqpu_run(c, st, nullptr); }

NqpuError nqpu_state_synchronize(NqpuState* st) {
  if (!st) return NQPU_NULL_POINTER;
  return synchronize_state(S(st));
}

NqpuError nqpu_last_run_stats(const NqpuState* st, NqpuRunStats* out) {
  if (!st || !out) return NQPU_NULL_POINTER;
  *out = S(st)->stats;
  return NQPU_OK;
}

NqpuError nqpu_circuit_plan(const NqpuCircuit* c, const NqpuRunOptions* opts, uint32_t precision,
                            NqpuRunStats* out) {
  if (!c || !out) return NQPU_NULL_POINTER;
  if (precision != NQPU_PRECISION_C64 && precision != NQPU_PRECISION_C128) return NQPU_INVALID_PARAMETER;
  NqpuRunOptions o{NQPU_FUSION_TILED, NQPU_SEMANTICS_STANDARD, 0, 0};
  if (opts) o = *opts;
  if (o.fusion > NQPU_FUSION_TILED || o.semantics > NQPU_SEMANTICS_REFERENCE) return NQPU_INVALID_PARAMETER;
  *out = plan_stats(*Cc(c), o, (int)precision);
  return NQPU_OK;
}

NqpuError nqpu_debug_compile_passes(const NqpuCircuit* c, const NqpuRunOptions* opts) {
  if (!c) return NQPU_NULL_POINTER;
  NqpuRunOptions o{NQPU_FUSION_TILED, NQPU_SEMANTICS_STANDARD, 0, 0, 0};
  if (opts) o = *opts;
  if (o.semantics > NQPU_SEMANTICS_REFERENCE) return NQPU_INVALID_PARAMETER;
  return jit_check_circuit(*Cc(c), o, NQPU_PRECISION_C64);
}

NqpuError nqpu_debug_compile_passes_ex(const NqpuCircuit* c, const NqpuRunOptions* opts, uint32_t precision) {
  if (!c) return NQPU_NULL_POINTER;
  if (precision != NQPU_PRECISION_C64 && precision != NQPU_PRECISION_C128) return NQPU_INVALID_PARAMETER;
  NqpuRunOptions o{NQPU_FUSION_TILED, NQPU_SEMANTICS_STANDARD, 0, 0, 0};
  if (opts) o = *opts;
  if (o.semantics > NQPU_SEMANTICS_REFERENCE) return NQPU_INVALID_PARAMETER;
  return jit_check_circuit(*Cc(c), o, (int)precision);
}

NqpuError nqpu_debug_emulate(const NqpuCircuit* c, const NqpuRunOptions* opts, uint32_t precision,
                             NqpuComplex* amps, uint64_t len) {
  if (!c || !amps) return NQPU_NULL_POINTER;
  if (precision != NQPU_PRECISION_C64 && precision != NQPU_PRECISION_C128) return NQPU_INVALID_PARAMETER;
  if (Cc(c)->n > 26 || len < (1ull << Cc(c)->n)) return NQPU_INVALID_PARAMETER;
  NqpuRunOptions o{NQPU_FUSION_TILED, NQPU_SEMANTICS_STANDARD, 0, 0, 0};
  if (opts) o = *opts;
  std::vector<cd> a(1ull << Cc(c)->n);
  for (size_t i = 0; i < a.size(); ++i) a[i] = cd(amps[i].re, amps[i].im);
  NqpuError e = emulate_tiled(*Cc(c), o, (int)precision, a);
  for (size_t i = 0; i < a.size(); ++i) amps[i] = NqpuComplex{a[i].real(), a[i].imag()};
  return e;
}

NqpuError nqpu_debug_state_layout(const NqpuState* state, int32_t* positions, uint32_t len) {
  if (!state || !positions) return NQPU_NULL_POINTER;
  const State* s = S(state);
  if (len < (uint32_t)s->n) return NQPU_INVALID_PARAMETER;
  for (int q = 0; q < s->n; ++q) positions[q] = s->perm.empty() ? q : s->perm[q];
  return NQPU_OK;
}

NqpuError nqpu_circuit_fusion_counts(const NqpuCircuit* c, NqpuFusionCounts* out) {
  if (!c || !out) return NQPU_NULL_POINTER;
  FusionResult fr = fuse_gates_reference(Cc(c)->gates);
  out->original_count = fr.original_count;
  out->fused_count = fr.fused_count;
  out->gates_eliminated = fr.gates_eliminated;
  out->items = fr.items.size();
  return NQPU_OK;
}

uint32_t nqpu_circuit_gate_count(const NqpuCircuit* c) { return c ? (uint32_t)Cc(c)->gates.size() : 0; }

// Greedy layer assignment (c_ffi.rs:666-690).
uint32_t nqpu_circuit_depth(const NqpuCircuit* c) {
  if (!c || Cc(c)->gates.empty()) return 0;
  std::vector<uint32_t> qd(Cc(c)->n, 0);
  uint32_t mx = 0;
  for (const Gate& g : Cc(c)->gates) {
    uint32_t layer = 0;
    for (int i = 0; i < g.nt; ++i) layer = std::max(layer, qd[g.t[i]]);
    for (int i = 0; i < g.nc; ++i) layer = std::max(layer, qd[g.c[i]]);
    for (int i = 0; i < g.nt; ++i) qd[g.t[i]] = layer + 1;
    for (int i = 0; i < g.nc; ++i) qd[g.c[i]] = layer + 1;
    mx = std::max(mx, layer + 1);
  }
  return mx;
}

NqpuVersion nqpu_version(void) { return NqpuVersion{0, 1, 0}; }

const char* nqpu_error_message(NqpuError e) {
  switch (e) {
    case NQPU_OK: return "Success";
    case NQPU_INVALID_QUBIT: return "Invalid qubit index";
    case NQPU_INVALID_PARAMETER: return "Invalid parameter";
    case NQPU_ALLOCATION_FAILED: return "Memory allocation failed";
    case NQPU_SIMULATION_FAILED: return "Simulation failed";
    case NQPU_NULL_POINTER: return "Null pointer";
    case NQPU_INVALID_HANDLE: return "Invalid handle";
    case NQPU_DEVICE_ERROR: return "Device error";
    case NQPU_OUT_OF_DEVICE_MEMORY: return "Out of device memory";
    case NQPU_UNAVAILABLE: return "GPU unavailable";
    default: return "Unknown error";
  }
}

void nqpu_set_seed(uint64_t seed) { rng_seed(seed); }

void nqpu_set_num_threads(uint32_t) {}

const char* nqpu_last_error_detail(void) { return last_error(); }

NqpuBatch* nqpu_batch_create(void) {
  Batch* b = nullptr;
  if (create_batch(&b) != NQPU_OK) return nullptr;
  return reinterpret_cast<NqpuBatch*>(b);
}

void nqpu_batch_destroy(NqpuBatch* b) { delete reinterpret_cast<Batch*>(b); }

static NqpuError batch_args(const NqpuCircuit* const* circuits, uint64_t count, const NqpuPauliTerm* terms,
                            const uint64_t* offsets, const NqpuRunOptions* opts, std::vector<const Circuit*>& cs,
                            std::vector<std::vector<PauliTerm>>& ts, NqpuRunOptions& o) {
  if (count && (!circuits || !offsets)) return NQPU_NULL_POINTER;
  o = NqpuRunOptions{NQPU_FUSION_TILED, NQPU_SEMANTICS_STANDARD, 0, 0, 0};
  if (opts) o = *opts;
  if (o.semantics > NQPU_SEMANTICS_REFERENCE) return NQPU_INVALID_PARAMETER;
  cs.assign(count, nullptr);
  ts.assign(count, {});
  for (uint64_t k = 0; k < count; ++k) {
    if (!circuits[k]) return NQPU_NULL_POINTER;
    cs[k] = Cc(circuits[k]);
    if (offsets[k + 1] < offsets[k]) return NQPU_INVALID_PARAMETER;
    if (offsets[k + 1] > offsets[k] && !terms) return NQPU_NULL_POINTER;
    for (uint64_t t = offsets[k]; t < offsets[k + 1]; ++t) {
      if (!std::isfinite(terms[t].coeff)) return NQPU_INVALID_PARAMETER;
      ts[k].push_back(PauliTerm{terms[t].xmask, terms[t].zmask, terms[t].coeff});
    }
  }
  return NQPU_OK;
}

NqpuError nqpu_batch_expectation(NqpuBatch* b, const NqpuCircuit* const* circuits, uint64_t count,
                                 const NqpuPauliTerm* terms, const uint64_t* offsets, const NqpuRunOptions* opts,
                                 double* out) {
  if (!b || !out) return NQPU_NULL_POINTER;
  std::vector<const Circuit*> cs;
  std::vector<std::vector<PauliTerm>> ts;
  NqpuRunOptions o;
  NqpuError e = batch_args(circuits, count, terms, offsets, opts, cs, ts, o);
  if (e != NQPU_OK) return e;
  return batch_expectation(reinterpret_cast<Batch*>(b), cs, ts, o, out);
}

NqpuError nqpu_batch_synchronize(NqpuBatch* b) {
  if (!b) return NQPU_NULL_POINTER;
  return batch_synchronize(reinterpret_cast<Batch*>(b));
}

NqpuError nqpu_debug_batch_plan(const NqpuCircuit* const* circuits, uint64_t count, const NqpuPauliTerm* terms,
                                const uint64_t* offsets, const NqpuRunOptions* opts, NqpuBatchStats* out,
                                uint64_t* program_ops) {
  if (!out) return NQPU_NULL_POINTER;
  std::vector<const Circuit*> cs;
  std::vector<std::vector<PauliTerm>> ts;
  NqpuRunOptions o;
  NqpuError e = batch_args(circuits, count, terms, offsets, opts, cs, ts, o);
  if (e != NQPU_OK) return e;
  BatchPlan plan;
  std::unique_ptr<char[]> image;  // left uninitialised, like the pinned image of a real batch
  e = plan_batch(cs, ts, o, plan, [&image](size_t bytes) {
    image.reset(new (std::nothrow) char[bytes]);
    return image.get();
  });
  if (e != NQPU_OK) return e;
  *out = plan.st;
  if (program_ops) *program_ops = plan.program_ops;
  return NQPU_OK;
}

NqpuError nqpu_batch_last_stats(const NqpuBatch* b, NqpuBatchStats* out) {
  if (!b || !out) return NQPU_NULL_POINTER;
  *out = reinterpret_cast<const Batch*>(b)->stats;
  return NQPU_OK;
}

uint32_t nqpu_device_count(void) { return (uint32_t)device_count(); }

NqpuServer* nqpu_server_create(const NqpuServerConfig* cfg) {
  NqpuServerConfig c{};
  if (cfg) c = *cfg;
  Server* s = nullptr;
  if (create_server(c, &s) != NQPU_OK) return nullptr;
  return reinterpret_cast<NqpuServer*>(s);
}

void nqpu_server_destroy(NqpuServer* s) {
  if (s) destroy_server(reinterpret_cast<Server*>(s));
}

NqpuError nqpu_server_submit(NqpuServer* s, const NqpuCircuit* c, const NqpuPauliTerm* terms, uint32_t num_terms,
                             uint32_t shots, uint64_t seed, uint64_t* id) {
  if (!s || !c || !id) return NQPU_NULL_POINTER;
  if (num_terms && !terms) return NQPU_NULL_POINTER;
  const Circuit* cc = Cc(c);
  const uint64_t lim = cc->n >= 64 ? ~0ull : ((1ull << cc->n) - 1);
  std::vector<PauliTerm> ts;
  ts.reserve(num_terms);
  for (uint32_t k = 0; k < num_terms; ++k) {
    if (!std::isfinite(terms[k].coeff)) return NQPU_INVALID_PARAMETER;
    if ((terms[k].xmask | terms[k].zmask) & ~lim) return NQPU_INVALID_QUBIT;
    ts.push_back(PauliTerm{terms[k].xmask, terms[k].zmask, terms[k].coeff});
  }
  return server_submit(reinterpret_cast<Server*>(s), *cc, ts, shots, seed, id);
}

NqpuError nqpu_server_poll(NqpuServer* s, uint64_t id, NqpuRequestStatus* out) {
  if (!s || !out) return NQPU_NULL_POINTER;
  return server_poll(reinterpret_cast<Server*>(s), id, out);
}

NqpuError nqpu_server_wait(NqpuServer* s, uint64_t id, uint32_t timeout_ms, NqpuRequestStatus* out) {
  if (!s) return NQPU_NULL_POINTER;
  return server_wait(reinterpret_cast<Server*>(s), id, timeout_ms, out);
}

NqpuError nqpu_server_take(NqpuServer* s, uint64_t id, double* expectations, uint32_t num_terms, uint64_t* samples,
                           uint32_t shots) {
  if (!s) return NQPU_NULL_POINTER;
  return server_take(reinterpret_cast<Server*>(s), id, expectations, num_terms, samples, shots);
}

NqpuError nqpu_server_stats(const NqpuServer* s, NqpuServerStats* out) {
  if (!s || !out) return NQPU_NULL_POINTER;
  return server_stats(reinterpret_cast<const Server*>(s), out);
}

}  // extern "C"

extern "C" {

static void to_public(const TunedPlan& t, NqpuTunedPlan* out) {
  out->tile_qubits = (uint32_t)t.tile_qubits;
  out->low_qubits = (uint32_t)t.low_qubits;
  out->flags = t.flags;
  out->tiles_per_group = (uint32_t)t.tiles_per_group;
  out->ms = t.ms;
  out->default_ms = t.default_ms;
}

NqpuError nqpu_autotune(uint32_t n, uint32_t precision, const NqpuCircuit* probe, NqpuTunedPlan* out) {
  if (!out) return NQPU_NULL_POINTER;
  if (precision != NQPU_PRECISION_C64 && precision != NQPU_PRECISION_C128) return NQPU_INVALID_PARAMETER;
  const uint32_t maxq = precision == NQPU_PRECISION_C64 ? 34 : 33;
  if (n < 1 || n > maxq) return NQPU_INVALID_QUBIT;
  if (probe && (uint32_t)Cc(probe)->n != n) return NQPU_INVALID_PARAMETER;
  TunedPlan t;
  NqpuError e = autotune((int)n, (int)precision, probe ? Cc(probe) : nullptr, &t);
  if (e != NQPU_OK) return e;
  to_public(t, out);
  return NQPU_OK;
}

NqpuError nqpu_tuned_plan(uint32_t n, uint32_t precision, NqpuTunedPlan* out) {
  if (!out) return NQPU_NULL_POINTER;
  TunedPlan t;
  if (!tuned_lookup((int)n, (int)precision, &t)) return NQPU_INVALID_PARAMETER;
  to_public(t, out);
  return NQPU_OK;
}

NqpuError nqpu_tune_load(const char* path) {
  if (!path) return NQPU_NULL_POINTER;
  return tune_load(path);
}

void nqpu_tune_reset(void) { tune_forget(); }

}  // extern "C"
