// engine.hip — host runtime: state handles on the device, gate lowering,
// the reference fusion pass (restated), the LDS tile planner, execution,
// FP64 reductions and sampling.
#include "runtime.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <map>
#include <functional>
#include <atomic>
#include <thread>
#include <mutex>
#include <unordered_map>
#include <memory>

namespace nqpu {

// ------------------------------------------------------------------ errors

static thread_local std::string g_err;
void set_error(const std::string& s) { g_err = s; }
const char* last_error() { return g_err.c_str(); }

NqpuError hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return NQPU_OK;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? NQPU_OUT_OF_DEVICE_MEMORY : NQPU_DEVICE_ERROR;
}
// A copy ordered on the caller's stream and waited for.  Plain hipMemcpy runs
// on the null stream, which would serialise against every other state's
// stream in the process (concurrent requests of a server, for one).
static hipError_t copy_on(hipStream_t st, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

// Every stream the library creates: default priority, and not synchronised
// with the null stream, so work of different handles can overlap.
hipError_t new_stream(hipStream_t* st) { return hipStreamCreateWithFlags(st, hipStreamNonBlocking); }

// ------------------------------------------------------------------- masks

static inline uint64_t bit(int q) { return 1ull << q; }

uint64_t HostOp::qubit_mask() const {
  uint64_t m = cmask | bit(a);
  if (kind == OP_U2 || kind == OP_SWAP || kind == OP_D2) m |= bit(b);
  if (kind == OP_DIAGK)
    for (int q : dq) m |= bit(q);
  return m;
}

uint64_t HostOp::tile_mask() const {
  switch (kind) {
    case OP_U1: case OP_FLIP: return bit(a);
    case OP_U2: case OP_SWAP: return bit(a) | bit(b);
    case OP_D1: case OP_D2: return 0;
    default: return qubit_mask();
  }
}

uint64_t Gate::qubit_mask() const {
  uint64_t m = 0;
  for (int i = 0; i < nt; ++i) m |= bit(t[i]);
  for (int i = 0; i < nc; ++i) m |= bit(c[i]);
  return m;
}

// ---------------------------------------------------------------- matrices
// Gate matrices follow GateType::matrix() (sdk/rust/src/core/gates/gates.rs:70-598).

static const cd I1(0, 1);
static cd ph(double a) { return cd(std::cos(a), std::sin(a)); }

static void u3m(double th, double phi, double la, cd* m) {  // gates.rs:136-150
  double c = std::cos(th / 2), s = std::sin(th / 2);
  m[0] = cd(c, 0);
  m[1] = cd(-s * std::cos(la), -s * std::sin(la));
  m[2] = cd(s * std::cos(phi), s * std::sin(phi));
  m[3] = cd(c * std::cos(phi + la), c * std::sin(phi + la));
}

// 2x2 matrix of a single-qubit kind (gate_to_matrix, gate_fusion.rs:127-214).
static bool matrix_1q(const Gate& g, cd* m) {
  const double r = 1.0 / std::sqrt(2.0), th = g.p[0];
  switch (g.kind) {
    case NQPU_GATE_H: m[0] = r; m[1] = r; m[2] = r; m[3] = -r; return true;
    case NQPU_GATE_X: m[0] = 0; m[1] = 1; m[2] = 1; m[3] = 0; return true;
    case NQPU_GATE_Y: m[0] = 0; m[1] = -I1; m[2] = I1; m[3] = 0; return true;
    case NQPU_GATE_Z: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = -1; return true;
    case NQPU_GATE_S: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = I1; return true;
    case NQPU_GATE_T: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = cd(std::sqrt(2.0) / 2, std::sqrt(2.0) / 2); return true;
    case NQPU_GATE_RX: m[0] = std::cos(th / 2); m[1] = cd(0, -std::sin(th / 2)); m[2] = m[1]; m[3] = m[0]; return true;
    case NQPU_GATE_RY: m[0] = std::cos(th / 2); m[1] = -std::sin(th / 2); m[2] = std::sin(th / 2); m[3] = m[0]; return true;
    case NQPU_GATE_RZ: m[0] = cd(std::cos(th / 2), -std::sin(th / 2)); m[1] = 0; m[2] = 0; m[3] = cd(std::cos(th / 2), std::sin(th / 2)); return true;
    case NQPU_GATE_U: u3m(g.p[0], g.p[1], g.p[2], m); return true;
    case NQPU_GATE_SX: m[0] = cd(.5, .5); m[1] = cd(.5, -.5); m[2] = cd(.5, -.5); m[3] = cd(.5, .5); return true;
    case NQPU_GATE_PHASE: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = ph(th); return true;
    default: return false;
  }
}

static HostOp op_u1(int a, uint64_t cm, const cd* m) {
  HostOp o;
  o.kind = OP_U1; o.a = a; o.cmask = cm;
  for (int i = 0; i < 4; ++i) o.m[i] = m[i];
  return o;
}
static HostOp op_d1(int a, uint64_t cm, cd p0, cd p1) {
  HostOp o;
  o.kind = OP_D1; o.a = a; o.cmask = cm; o.m[0] = p0; o.m[1] = p1;
  return o;
}
static HostOp op_flip(int a, uint64_t cm) {
  HostOp o;
  o.kind = OP_FLIP; o.a = a; o.cmask = cm;
  return o;
}

// Lower one gate.  Reference semantics reproduce what apply_gate_to_state
// (sdk/rust/src/circuits/visualization/ascii_viz.rs:124-179) executes.
static void lower_gate_into(const Gate& g, bool ref, std::vector<HostOp>& v) {
  const int t0 = g.t[0], t1 = g.t[1];
  const uint64_t c0 = g.nc > 0 ? bit(g.c[0]) : 0, c01 = c0 | (g.nc > 1 ? bit(g.c[1]) : 0);
  const double th = g.p[0];
  cd m[16];
  switch (g.kind) {
    case NQPU_GATE_X: v.push_back(op_flip(t0, 0)); break;
    case NQPU_GATE_Z: v.push_back(op_d1(t0, 0, 1, -1)); break;
    case NQPU_GATE_S: v.push_back(op_d1(t0, 0, 1, I1)); break;
    case NQPU_GATE_T: v.push_back(op_d1(t0, 0, 1, ph(M_PI / 4))); break;
    case NQPU_GATE_RZ: v.push_back(op_d1(t0, 0, ph(-th / 2), ph(th / 2))); break;
    case NQPU_GATE_PHASE: v.push_back(op_d1(t0, 0, 1, ph(th))); break;
    case NQPU_GATE_H: case NQPU_GATE_Y: case NQPU_GATE_RX: case NQPU_GATE_RY: case NQPU_GATE_U: case NQPU_GATE_SX:
      matrix_1q(g, m);
      v.push_back(op_u1(t0, 0, m));
      break;
    case NQPU_GATE_CNOT: v.push_back(op_flip(t0, c0)); break;
    case NQPU_GATE_CZ: v.push_back(op_d1(t0, c0, 1, -1)); break;
    case NQPU_GATE_CR: v.push_back(op_d1(t0, c0, 1, ph(th))); break;
    case NQPU_GATE_SWAP:
      if (t0 != t1) {
        HostOp o; o.kind = OP_SWAP; o.a = t0; o.b = t1; v.push_back(o);
      }
      break;
    case NQPU_GATE_TOFFOLI: v.push_back(op_flip(t0, c01)); break;
    case NQPU_GATE_CCZ: v.push_back(op_d1(t0, c01, 1, -1)); break;
    case NQPU_GATE_CRX:
      if (ref) {  // lib.rs:1292-1358 applies a controlled-Rz
        v.push_back(op_d1(t0, c0, ph(-th / 2), ph(th / 2)));
      } else {
        Gate r = g; r.kind = NQPU_GATE_RX; matrix_1q(r, m); v.push_back(op_u1(t0, c0, m));
      }
      break;
    case NQPU_GATE_CRY: {
      Gate r = g; r.kind = NQPU_GATE_RY; matrix_1q(r, m); v.push_back(op_u1(t0, c0, m));
      break;
    }
    case NQPU_GATE_CRZ:
      if (ref)  // lib.rs:1439-1473: only |c=1,t=0> *= e^{+i th/2}
        v.push_back(op_d1(t0, c0, ph(th / 2), 1));
      else
        v.push_back(op_d1(t0, c0, ph(-th / 2), ph(th / 2)));
      break;
    case NQPU_GATE_ISWAP: {  // lib.rs:1652-1686
      HostOp o; o.kind = OP_U2; o.a = t0; o.b = t1;
      for (int i = 0; i < 16; ++i) o.m[i] = 0;
      o.m[0] = 1; o.m[15] = 1; o.m[1 * 4 + 2] = I1; o.m[2 * 4 + 1] = I1;
      v.push_back(o);
      break;
    }
    case NQPU_GATE_RXX: case NQPU_GATE_RYY: {
      if (ref) break;  // ascii_viz.rs:166-168: no-op
      double c = std::cos(th / 2), s = std::sin(th / 2);
      HostOp o; o.kind = OP_U2; o.a = t0; o.b = t1;
      for (int i = 0; i < 16; ++i) o.m[i] = 0;
      for (int i = 0; i < 4; ++i) o.m[i * 5] = c;
      const double y = (g.kind == NQPU_GATE_RYY) ? 1.0 : -1.0;  // gates.rs:434-484
      o.m[0 * 4 + 3] = cd(0, y * s); o.m[3 * 4 + 0] = cd(0, y * s);
      o.m[1 * 4 + 2] = cd(0, -s); o.m[2 * 4 + 1] = cd(0, -s);
      v.push_back(o);
      break;
    }
    case NQPU_GATE_RZZ: {  // gates.rs:486-511
      if (ref) break;
      HostOp o; o.kind = OP_D2; o.a = t0; o.b = t1;
      o.m[0] = ph(-th / 2); o.m[1] = ph(th / 2); o.m[2] = ph(th / 2); o.m[3] = ph(-th / 2);
      v.push_back(o);
      break;
    }
    case NQPU_GATE_CSWAP: {
      if (ref) break;
      HostOp o; o.kind = OP_SWAP; o.a = t0; o.b = t1; o.cmask = c0;
      v.push_back(o);
      break;
    }
    case NQPU_GATE_CU: {  // gates.rs:548-598, rows 2-3
      if (ref) break;
      double thh = g.p[0], phi = g.p[1], la = g.p[2], ga = g.p[3];
      double c = std::cos(thh / 2), s = std::sin(thh / 2);
      m[0] = ph(ga) * c; m[1] = -ph(ga + la) * s; m[2] = ph(ga + phi) * s; m[3] = ph(ga + phi + la) * c;
      v.push_back(op_u1(t0, c0, m));
      break;
    }
    case NQPU_GATE_CUSTOM1:
      if (ref) break;
      v.push_back(op_u1(t0, 0, g.m));
      break;
    case NQPU_GATE_CUSTOM2: {
      if (ref) break;
      // targets[0] is the high bit of the 4x4 index
      HostOp o; o.kind = OP_U2; o.a = t1; o.b = t0;
      for (int i = 0; i < 16; ++i) o.m[i] = g.m[i];
      v.push_back(o);
      break;
    }
    default: break;
  }
}

std::vector<HostOp> lower_gate(const Gate& g, bool ref) {
  std::vector<HostOp> v;
  lower_gate_into(g, ref, v);
  return v;
}

std::vector<HostOp> lower_circuit(const Circuit& c, bool ref) {
  std::vector<HostOp> v;
  v.reserve(c.gates.size() + 8);
  for (const Gate& g : c.gates) lower_gate_into(g, ref, v);
  return v;
}

// 2x2 of an uncontrolled single-qubit op.
static bool op_matrix_1q(const HostOp& o, cd* m) {
  if (o.cmask) return false;
  if (o.kind == OP_U1) { for (int i = 0; i < 4; ++i) m[i] = o.m[i]; return true; }
  if (o.kind == OP_D1) { m[0] = o.m[0]; m[1] = 0; m[2] = 0; m[3] = o.m[1]; return true; }
  if (o.kind == OP_FLIP) { m[0] = 0; m[1] = 1; m[2] = 1; m[3] = 0; return true; }
  return false;
}

static void mul2(const cd* b, const cd* a, cd* out) {  // out = b * a
  cd r[4];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) r[i * 2 + j] = b[i * 2] * a[j] + b[i * 2 + 1] * a[2 + j];
  for (int i = 0; i < 4; ++i) out[i] = r[i];
}

static bool is_identity2(const cd* m, double eps) {  // is_identity_matrix, gate_fusion.rs:217
  return std::abs(m[0].real() - 1) < eps && std::abs(m[0].imag()) < eps && std::abs(m[1].real()) < eps &&
         std::abs(m[1].imag()) < eps && std::abs(m[2].real()) < eps && std::abs(m[2].imag()) < eps &&
         std::abs(m[3].real() - 1) < eps && std::abs(m[3].imag()) < eps;
}

// Per-qubit accumulation of consecutive uncontrolled single-qubit ops into
// one 2x2 (the single-qubit part of fuse_gates, gate_fusion.rs:537-666).
std::vector<HostOp> merge_single_qubit(const std::vector<HostOp>& ops) {
  struct Pend { cd m[4]; int count = 0; size_t first = 0; };  // count 0: nothing pending
  Pend pend[64];
  uint64_t live = 0;
  std::vector<HostOp> out;
  out.reserve(ops.size());
  auto flush = [&](int q) {
    Pend& p = pend[q];
    live &= ~bit(q);
    const int count = p.count;
    p.count = 0;
    if (count == 1) { out.push_back(ops[p.first]); return; }
    if (is_identity2(p.m, 1e-12)) return;
    if (p.m[1] == cd(0) && p.m[2] == cd(0)) out.push_back(op_d1(q, 0, p.m[0], p.m[3]));
    else out.push_back(op_u1(q, 0, p.m));
  };
  for (size_t k = 0; k < ops.size(); ++k) {
    const HostOp& o = ops[k];
    cd m[4];
    if (op_matrix_1q(o, m)) {
      Pend& p = pend[o.a];
      if (!p.count) {
        live |= bit(o.a);
        p.count = 1;
        p.first = k;
        for (int i = 0; i < 4; ++i) p.m[i] = m[i];
      } else {
        mul2(m, p.m, p.m);
        p.count++;
      }
      continue;
    }
    for (uint64_t qm = o.qubit_mask() & live; qm; qm &= qm - 1) flush(__builtin_ctzll(qm));
    out.push_back(o);
  }
  for (uint64_t qm = live; qm; qm &= qm - 1) flush(__builtin_ctzll(qm));
  return out;
}

// Greedy tile planner.  A pass owns a set S of at most K qubits, always
// including the `low` lowest qubits so that tile loads are contiguous runs
// of 2^low amplitudes.  Walking the remaining ops in order, an op joins the
// pass when none of its qubits is blocked by an earlier op left behind and
// its qubits fit in S; otherwise its qubits become blocked.  Ops that join
// commute with every op they overtake (disjoint qubits), so the order of
// application is preserved.
static std::vector<Pass> plan_tiles_once(const std::vector<HostOp>& ops, int n, const PlanOptions& po, bool move,
                                         std::vector<int>* final_pi) {
  std::vector<Pass> passes;
  const int K = std::min(po.tile_qubits, n);
  const int L = std::min(po.low_qubits, K);
  const uint64_t all = (n >= 64) ? ~0ull : (bit(n) - 1);
  std::vector<size_t> rem(ops.size());
  for (size_t i = 0; i < ops.size(); ++i) rem[i] = i;
  const uint64_t low = (L > 0) ? (bit(L) - 1) : 0;
  // vector work of an op, in dense-2x2 units: what a pass should absorb most of
  auto weight = [](const HostOp& o) {
    switch (o.kind) {
      case OP_U1: return 1.0;
      case OP_U2: return 2.0;
      case OP_DIAGK: return 1.0;
      case OP_SWAP: return 0.2;
      default: return 0.4;
    }
  };
  // The ops of `rem` one pass takes: with `grow`, the tile set starts at the
  // low qubits and takes the targets of ops in circuit order while they fit;
  // otherwise the tile set is S as given.  An op joins when none of its
  // qubits is blocked by an earlier op left behind and its targets are tile
  // bits; an op left behind blocks its qubits.
  auto take = [&](uint64_t& S, bool grow, std::vector<size_t>* chosen, std::vector<size_t>* next) {
    uint64_t blocked = 0;
    double w = 0;
    size_t i = 0;
    for (; i < rem.size(); ++i) {
      const HostOp& o = ops[rem[i]];
      const uint64_t qm = o.qubit_mask(), tm = po.global_ops ? o.tile_mask() : qm;
      bool join = (qm & blocked) == 0;
      if (join) {
        if (grow) join = __builtin_popcountll(S | tm) <= K;
        else join = (tm & ~S) == 0;
      }
      if (join) {
        S |= tm;
        w += weight(o);
        if (chosen) chosen->push_back(rem[i]);
      } else {
        blocked |= qm;
        if (next) next->push_back(rem[i]);
      }
      if ((blocked & all) == all) {  // nothing further can join this pass
        ++i;
        break;
      }
    }
    if (next)
      for (; i < rem.size(); ++i) next->push_back(rem[i]);
    return w;
  };
  // Moving low set (po.move_low, complex64 register passes of 11..13 qubits):
  // a pass stores its tile in any order of the tile's physical positions, so
  // the five physical low positions, which every tile must hold for coalesced
  // loads and stores, can carry any five qubits of the previous tile.  The
  // planner keeps a map pi from circuit qubits to physical positions; a pass
  // is planned on circuit qubits (its tile S must contain the qubits now at
  // the physical low positions) and emitted in physical positions.  It then
  // hands five qubits of S to the low positions (those shared with the best
  // next window of K consecutive qubits) and puts the rest of S back at
  // their own positions where it can.  The plan then ends with passes that
  // only permute, until pi is the identity again (closing_passes), or with
  // po.keep_perm leaves the qubits where they are and reports pi.

  std::vector<int> pi(n), inv(n);
  for (int q = 0; q < n; ++q) pi[q] = inv[q] = q;
  auto low_set = [&]() {
    uint64_t m = 0;
    for (int p = 0; p < L; ++p) m |= bit(inv[p]);
    return m;
  };
  auto to_phys = [&](uint64_t m) {
    uint64_t r = 0;
    for (; m; m &= m - 1) r |= bit(pi[__builtin_ctzll(m)]);
    return r;
  };
  // emit a pass of circuit-qubit tile S (K qubits) with the given ops, which
  // stores circuit qubit q at newpi[q] (a permutation of pi over S)
  auto emit_pass = [&](uint64_t S, const std::vector<size_t>& chosen, const std::vector<int>& newpi) {
    Pass p;
    const uint64_t P = to_phys(S);
    for (int x = 0; x < n; ++x)
      if ((P >> x) & 1) p.qubits.push_back(x);
    for (size_t k : chosen) {
      HostOp o = ops[k];
      if (move) {
        o.a = pi[o.a];
        if (o.kind == OP_U2 || o.kind == OP_SWAP || o.kind == OP_D2) o.b = pi[o.b];
        o.cmask = to_phys(o.cmask);
      }
      p.ops.push_back(std::move(o));
    }
    if (move) {
      bool ident = true;
      for (int x : p.qubits) {
        p.out.push_back(newpi[inv[x]]);
        ident = ident && p.out.back() == x;
      }
      if (ident) p.out.clear();
    }
    passes.push_back(std::move(p));
  };
  // new placement of the circuit qubits of S: Lnew on the low positions (those
  // already there keep their slot), the others at home where possible
  auto place = [&](uint64_t S, uint64_t Lnew) {
    std::vector<int> np(pi);
    const uint64_t P = to_phys(S);
    uint64_t used = 0;
    for (uint64_t m = Lnew; m; m &= m - 1) {
      const int q = __builtin_ctzll(m);
      if (pi[q] < L) used |= bit(np[q] = pi[q]);
    }
    for (uint64_t m = Lnew; m; m &= m - 1) {
      const int q = __builtin_ctzll(m);
      if (pi[q] < L) continue;
      int x = 0;
      while ((used >> x) & 1) ++x;
      used |= bit(np[q] = x);
    }
    const uint64_t rest = S & ~Lnew;
    for (uint64_t m = rest; m; m &= m - 1) {
      const int q = __builtin_ctzll(m);
      np[q] = -1;
      if (((P >> q) & 1) && !((used >> q) & 1) && q >= L) used |= bit(np[q] = q);
    }
    for (uint64_t m = rest; m; m &= m - 1) {
      const int q = __builtin_ctzll(m);
      if (np[q] >= 0) continue;
      const uint64_t fr = P & ~used & ~(bit(L) - 1);
      used |= bit(np[q] = __builtin_ctzll(fr));
    }
    return np;
  };
  auto commit = [&](const std::vector<int>& np) {
    pi = np;
    for (int q = 0; q < n; ++q) inv[pi[q]] = q;
  };
  while (!rem.empty()) {
    // candidates: the greedy set, and the low qubits plus every window of
    // consecutive qubits that fills the tile (circuits with 1D structure
    // sweep well with these)
    const uint64_t lowm = move ? low_set() : low;
    uint64_t best = lowm;
    double bw = take(best, true, nullptr, nullptr);
    for (int q = 0; q < n && __builtin_popcountll(best) < K; ++q) best |= bit(q);
    if (K > L && n > K)
      for (int s0 = 0; s0 < n; ++s0) {
        uint64_t S = lowm;
        for (int q = s0; q < n && __builtin_popcountll(S) < K; ++q) S |= bit(q);
        if (__builtin_popcountll(S) < K) break;
        const double cw = take(S, false, nullptr, nullptr);
        if (cw > bw + 1e-9) {
          bw = cw;
          best = S;
        }
      }
    uint64_t S = best;
    std::vector<size_t> chosen, next;
    next.reserve(rem.size());
    take(S, false, &chosen, &next);
    if (chosen.empty()) {  // cannot happen: the greedy candidate always takes the first op
      S |= ops[next.front()].tile_mask();
      chosen.push_back(next.front());
      next.erase(next.begin());
    }
    for (int q = 0; q < n && __builtin_popcountll(S) < K; ++q) S |= bit(q);
    std::vector<int> np(pi);
    if (move) {
      // the low set for the next pass: the qubits of S in the best next
      // window of K consecutive qubits (ties: already low, then lowest)
      uint64_t share = 0;
      double nw = -1;
      rem.swap(next);
      for (int s0 = 0; s0 + K <= n && !rem.empty(); ++s0) {
        uint64_t W = (bit(K) - 1) << s0;
        if (__builtin_popcountll(W & S) < L) continue;
        const double w2 = take(W, false, nullptr, nullptr) + 1e-3 * __builtin_popcountll(W & S & lowm);
        if (w2 > nw + 1e-9) {
          nw = w2;
          share = W & S;
        }
      }
      rem.swap(next);
      if (next.empty()) share = S & (bit(L) - 1);  // the last pass sends the low qubits home
      std::vector<int> cand;
      for (uint64_t m = share; m; m &= m - 1) cand.push_back(__builtin_ctzll(m));
      std::stable_sort(cand.begin(), cand.end(), [&](int x, int y) {
        return std::make_pair(pi[x] >= L, x >= L) < std::make_pair(pi[y] >= L, y >= L);
      });
      uint64_t Lnew = 0;
      for (int q : cand)
        if (__builtin_popcountll(Lnew) < L) Lnew |= bit(q);
      for (int q = 0; q < n && __builtin_popcountll(Lnew) < L; ++q)  // fewer than L shared: keep low ones
        if (((S >> q) & 1) && pi[q] < L) Lnew |= bit(q);
      for (int q = 0; q < n && __builtin_popcountll(Lnew) < L; ++q)
        if ((S >> q) & 1) Lnew |= bit(q);
      np = place(S, Lnew);
    }
    emit_pass(S, chosen, np);
    if (move) commit(np);
    rem.swap(next);
  }
  if (move) {
    std::vector<Pass> close;
    if (!closing_passes(pi, n, K, L, &close)) return {};  // the caller keeps the fixed-low plan
    if (po.keep_perm) {
      if (final_pi) *final_pi = pi;
    } else {
      for (Pass& p : close) passes.push_back(std::move(p));
    }
  }
  return passes;
}

// Back to the identity: each pass holds the low positions and, cycle by
// cycle of the remaining permutation (shortest first, those through the low
// positions first), as many further positions as fit; every qubit whose
// position is in the pass goes home if its home is too, the others fill the
// positions left.  False if the passes do not reach the identity.
bool closing_passes(std::vector<int> pi, int n, int K, int L, std::vector<Pass>* out) {
  std::vector<int> inv(n);
  for (int q = 0; q < n; ++q) inv[pi[q]] = q;
  for (int guard = 0; guard < 4 * n; ++guard) {
    bool done = true;
    for (int q = 0; q < n; ++q) done = done && pi[q] == q;
    if (done) break;
    std::vector<std::vector<int>> cycles;
    std::vector<char> seen(n, 0);
    for (int x = 0; x < n; ++x) {
      if (seen[x] || inv[x] == x) continue;
      std::vector<int> cyc;
      for (int y = x; !seen[y]; y = inv[y]) {  // the qubit at y belongs at position inv[y]
        seen[y] = 1;
        cyc.push_back(y);
      }
      cycles.push_back(cyc);
    }
    std::stable_sort(cycles.begin(), cycles.end(), [&](const std::vector<int>& a, const std::vector<int>& b) {
      const bool la = *std::min_element(a.begin(), a.end()) < L, lb = *std::min_element(b.begin(), b.end()) < L;
      return la != lb ? la : a.size() < b.size();
    });
    uint64_t P = bit(L) - 1;
    for (const auto& cyc : cycles)
      for (int x : cyc)
        if (__builtin_popcountll(P) < K) P |= bit(x);
    for (int x = 0; x < n && __builtin_popcountll(P) < K; ++x) P |= bit(x);
    uint64_t S = 0;
    for (uint64_t m = P; m; m &= m - 1) S |= bit(inv[__builtin_ctzll(m)]);
    std::vector<int> np(pi);
    uint64_t used = 0;
    for (uint64_t m = S; m; m &= m - 1) {
      const int q = __builtin_ctzll(m);
      np[q] = -1;
      if ((P >> q) & 1) used |= bit(np[q] = q);
    }
    for (uint64_t m = S; m; m &= m - 1) {
      const int q = __builtin_ctzll(m);
      if (np[q] < 0) used |= bit(np[q] = __builtin_ctzll(P & ~used));
    }
    // a pass without ops over physical positions P that stores the qubit at
    // position x at np[inv[x]]
    Pass p;
    bool ident = true;
    for (int x = 0; x < n; ++x)
      if ((P >> x) & 1) {
        p.qubits.push_back(x);
        p.out.push_back(np[inv[x]]);
        ident = ident && p.out.back() == x;
      }
    if (ident) p.out.clear();
    out->push_back(std::move(p));
    pi = np;
    for (int q = 0; q < n; ++q) inv[pi[q]] = q;
  }
  for (int q = 0; q < n; ++q)
    if (pi[q] != q) return false;
  return true;
}

// With a moving low set (complex64 register passes), both plans are made and
// the one with fewer passes kept: a QFT's final qubit reversal, for one, is
// free inside fixed-low tiles but leaves a permutation to undo otherwise.
std::vector<Pass> plan_tiles(const std::vector<HostOp>& ops, int n, const PlanOptions& po,
                             std::vector<int>* final_pi) {
  if (final_pi) final_pi->clear();
  const int K = std::min(po.tile_qubits, n), L = std::min(po.low_qubits, K);
  if (!(po.move_low && n > K && K >= 11 && K <= 13 && L == 5)) return plan_tiles_once(ops, n, po, false, nullptr);
  std::vector<Pass> fixed;
  std::thread th([&] { fixed = plan_tiles_once(ops, n, po, false, nullptr); });
  std::vector<int> pm;
  std::vector<Pass> moving = plan_tiles_once(ops, n, po, true, &pm);
  th.join();
  if (moving.empty() || moving.size() >= fixed.size()) return fixed;
  bool ident = true;
  for (int q = 0; q < (int)pm.size(); ++q) ident = ident && pm[q] == q;
  if (final_pi && !ident) *final_pi = pm;
  return moving;
}

// ------------------------------------------------------ reference fusion
// Restatement of fuse_gates / fuse_diagonal_pass
// (sdk/rust/src/circuits/synthesis/gate_fusion.rs:537-851).

static bool is_fusable_1q(const Gate& g) {  // gate_fusion.rs:440-462
  if (g.nc != 0 || g.nt != 1) return false;
  switch (g.kind) {
    case NQPU_GATE_H: case NQPU_GATE_X: case NQPU_GATE_Y: case NQPU_GATE_Z: case NQPU_GATE_S: case NQPU_GATE_T:
    case NQPU_GATE_RX: case NQPU_GATE_RY: case NQPU_GATE_RZ: case NQPU_GATE_U: case NQPU_GATE_SX: case NQPU_GATE_PHASE:
      return true;
    default: return false;
  }
}

// two_qubit_gate_to_matrix (gate_fusion.rs:323-426): lower qubit -> bit 0.
static bool matrix_2q_ref(const Gate& g, cd* m) {
  std::vector<int> qs;
  for (int i = 0; i < g.nt; ++i) qs.push_back(g.t[i]);
  for (int i = 0; i < g.nc; ++i) qs.push_back(g.c[i]);
  std::sort(qs.begin(), qs.end());
  qs.erase(std::unique(qs.begin(), qs.end()), qs.end());
  if (qs.size() != 2) return false;
  const int qa = qs[0];
  auto remap = [&](int q) { return q == qa ? 0 : 1; };
  int control = g.nc > 0 ? remap(g.c[0]) : 0;
  int target = g.nt > 0 ? remap(g.t[0]) : 0;
  for (int i = 0; i < 16; ++i) m[i] = 0;
  switch (g.kind) {
    case NQPU_GATE_CNOT:
      if (control == 1 && target == 0) { m[0] = 1; m[5] = 1; m[2 * 4 + 3] = 1; m[3 * 4 + 2] = 1; }
      else { m[0] = 1; m[1 * 4 + 3] = 1; m[2 * 4 + 2] = 1; m[3 * 4 + 1] = 1; }
      return true;
    case NQPU_GATE_CZ: m[0] = 1; m[5] = 1; m[10] = 1; m[15] = -1; return true;
    case NQPU_GATE_SWAP: m[0] = 1; m[1 * 4 + 2] = 1; m[2 * 4 + 1] = 1; m[15] = 1; return true;
    case NQPU_GATE_CR: m[0] = 1; m[5] = 1; m[10] = 1; m[15] = ph(g.p[0]); return true;
    case NQPU_GATE_ISWAP: m[0] = 1; m[1 * 4 + 2] = I1; m[2 * 4 + 1] = I1; m[15] = 1; return true;
    default: return false;  // CRx/CRy/CRz and the rest go through the dispatcher
  }
}

static void mul4(const cd* b, const cd* a, cd* out) {  // out = b * a
  cd r[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      cd s = 0;
      for (int k = 0; k < 4; ++k) s += b[i * 4 + k] * a[k * 4 + j];
      r[i * 4 + j] = s;
    }
  for (int i = 0; i < 16; ++i) out[i] = r[i];
}

static bool is_diag_1q_kind(int k) {
  return k == NQPU_GATE_Z || k == NQPU_GATE_S || k == NQPU_GATE_T || k == NQPU_GATE_RZ || k == NQPU_GATE_PHASE;
}

static std::pair<cd, cd> diag_1q_phases(const Gate& g) {  // gate_fusion.rs:690-707
  switch (g.kind) {
    case NQPU_GATE_Z: return {1, -1};
    case NQPU_GATE_S: return {1, I1};
    case NQPU_GATE_T: return {1, cd(std::sqrt(2.0) / 2, std::sqrt(2.0) / 2)};
    case NQPU_GATE_RZ: return {cd(std::cos(g.p[0] / 2), -std::sin(g.p[0] / 2)), cd(std::cos(g.p[0] / 2), std::sin(g.p[0] / 2))};
    case NQPU_GATE_PHASE: return {1, ph(g.p[0])};
    default: return {1, 1};
  }
}

FusionResult fuse_gates_reference(const std::vector<Gate>& gates) {
  FusionResult res;
  res.original_count = gates.size();
  struct Acc { bool live = false; cd m[4]; size_t count = 0; bool has_orig = false; Gate orig; };
  std::map<int, Acc> acc;  // ordered: the flush order of commuting 1q items is irrelevant
  std::vector<FusedItem> out;
  size_t elim = 0;
  auto flush = [&](int q) {
    auto it = acc.find(q);
    if (it == acc.end()) return;
    Acc a = it->second;
    acc.erase(it);
    if (a.count == 0) return;
    if (is_identity2(a.m, 1e-12)) { elim += a.count; return; }
    if (a.count == 1 && a.has_orig) {
      FusedItem f; f.kind = FusedItem::ORIGINAL; f.gate = a.orig; out.push_back(f); return;
    }
    FusedItem f; f.kind = FusedItem::FUSED; f.q = q; f.fused_count = (int)a.count;
    for (int i = 0; i < 4; ++i) f.m[i] = a.m[i];
    out.push_back(f);
    elim += a.count - 1;
  };
  for (const Gate& g : gates) {
    if (is_fusable_1q(g)) {
      cd m[4];
      matrix_1q(g, m);
      Acc& a = acc[g.t[0]];
      if (!a.live) {
        a.live = true; a.m[0] = 1; a.m[1] = 0; a.m[2] = 0; a.m[3] = 1; a.count = 0;
      }
      mul2(m, a.m, a.m);
      a.count++;
      if (a.count == 1) { a.has_orig = true; a.orig = g; } else { a.has_orig = false; }
      continue;
    }
    std::vector<int> qs;
    for (int i = 0; i < g.nt; ++i) qs.push_back(g.t[i]);
    for (int i = 0; i < g.nc; ++i) qs.push_back(g.c[i]);
    std::sort(qs.begin(), qs.end());
    qs.erase(std::unique(qs.begin(), qs.end()), qs.end());
    cd g4[16];
    if (qs.size() == 2 && matrix_2q_ref(g, g4)) {
      const int lo = qs[0], hi = qs[1];
      cd mlo[4] = {1, 0, 0, 1}, mhi[4] = {1, 0, 0, 1};
      size_t clo = 0, chi = 0;
      auto il = acc.find(lo);
      if (il != acc.end()) { for (int i = 0; i < 4; ++i) mlo[i] = il->second.m[i]; clo = il->second.count; acc.erase(il); }
      auto ih = acc.find(hi);
      if (ih != acc.end()) { for (int i = 0; i < 4; ++i) mhi[i] = ih->second.m[i]; chi = ih->second.count; acc.erase(ih); }
      if (clo == 0 && chi == 0) {
        FusedItem f; f.kind = FusedItem::ORIGINAL; f.gate = g; out.push_back(f);
      } else {
        cd pre[16];  // kronecker_2x2(hi, lo), gate_fusion.rs:255
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) pre[r * 4 + c] = mhi[(r >> 1) * 2 + (c >> 1)] * mlo[(r & 1) * 2 + (c & 1)];
        FusedItem f; f.kind = FusedItem::FUSED_TWO; f.qlo = lo; f.qhi = hi;
        mul4(g4, pre, f.m);
        f.fused_count = (int)(1 + clo + chi);
        elim += clo + chi;
        out.push_back(f);
      }
    } else {
      for (int q : qs) flush(q);
      FusedItem f; f.kind = FusedItem::ORIGINAL; f.gate = g; out.push_back(f);
    }
  }
  std::vector<int> keys;
  for (auto& kv : acc) keys.push_back(kv.first);
  for (int q : keys) flush(q);

  // fuse_diagonal_pass (gate_fusion.rs:757-851)
  std::vector<FusedItem> dout;
  std::map<int, std::pair<cd, cd>> d1;
  std::map<std::pair<int, int>, cd> d2;
  size_t dcount = 0;
  auto is_diag_item = [](const FusedItem& it) {
    if (it.kind != FusedItem::ORIGINAL) return false;
    const Gate& g = it.gate;
    if (g.nt == 1 && g.nc == 0) return is_diag_1q_kind(g.kind);
    return g.kind == NQPU_GATE_CZ || g.kind == NQPU_GATE_CR;
  };
  auto flush_diag = [&]() {
    if (dcount >= 2) {
      FusedItem f; f.kind = FusedItem::DIAGONAL; f.fused_count = (int)dcount;
      for (auto& kv : d1) f.single.push_back(kv);
      for (auto& kv : d2) f.pairs.push_back(kv);
      dout.push_back(f);
    } else {  // drain_diag_singles (gate_fusion.rs:847-902)
      for (auto& kv : d1) {
        cd m[4] = {kv.second.first, 0, 0, kv.second.second};
        if (is_identity2(m, 1e-12)) continue;
        FusedItem f; f.kind = FusedItem::FUSED; f.q = kv.first;
        for (int i = 0; i < 4; ++i) f.m[i] = m[i];
        dout.push_back(f);
      }
      for (auto& kv : d2) {
        FusedItem f; f.kind = FusedItem::FUSED_TWO; f.qlo = kv.first.first; f.qhi = kv.first.second;
        for (int i = 0; i < 16; ++i) f.m[i] = 0;
        f.m[0] = 1; f.m[5] = 1; f.m[10] = 1; f.m[15] = kv.second;
        bool ident = std::abs(kv.second.real() - 1) <= 1e-12 && std::abs(kv.second.imag()) <= 1e-12;
        if (!ident) dout.push_back(f);
      }
    }
    d1.clear();
    d2.clear();
    dcount = 0;
  };
  for (const FusedItem& it : out) {
    if (is_diag_item(it)) {
      const Gate& g = it.gate;
      if (g.nt == 1 && g.nc == 0) {
        auto p = diag_1q_phases(g);
        auto e = d1.find(g.t[0]);
        if (e == d1.end()) d1[g.t[0]] = p;
        else { e->second.first *= p.first; e->second.second *= p.second; }
      } else {
        int a = g.t[0], b = g.c[0];
        int lo = std::min(a, b), hi = std::max(a, b);
        cd p11 = g.kind == NQPU_GATE_CZ ? cd(-1) : ph(g.p[0]);
        auto e = d2.find({lo, hi});
        if (e == d2.end()) d2[{lo, hi}] = p11; else e->second *= p11;
      }
      dcount++;
    } else {
      flush_diag();
      dout.push_back(it);
    }
  }
  flush_diag();
  res.items = std::move(dout);
  res.gates_eliminated = elim;
  for (auto& it : res.items)
    if (it.kind != FusedItem::ORIGINAL) res.fused_count++;
  return res;
}

// ------------------------------------------------------------------ device

int device_count() {
  const char* f = std::getenv("NQPU_HIP_FORCE_NO_DEVICE");
  if (f && f[0] == '1') return 0;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

DeviceScratch::~DeviceScratch() {
  if (p) (void)hipFree(p);
}
bool DeviceScratch::ensure(size_t b) {
  if (b <= bytes) return true;
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  size_t nb = std::max(b, (size_t)1 << 20);
  if (hipMalloc(&p, nb) != hipSuccess) {
    p = nullptr;
    set_error("hipMalloc of scratch failed");
    return false;
  }
  bytes = nb;
  return true;
}

State::~State() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (stage) (void)hipHostFree(stage);
  if (stage_ev) (void)hipEventDestroy(stage_ev);
  if (amps && owns_amps) (void)hipFree(amps);
  if (stream) (void)hipStreamDestroy(stream);
}

template <typename F>
static void dispatch(const State* s, F&& f) {
  if (s->prec == NQPU_PRECISION_C64) f((float*)nullptr);
  else f((double*)nullptr);
}

NqpuError create_state(int n, int prec, State** out) {
  *out = nullptr;
  if (device_count() == 0) {
    set_error("no usable HIP device (or NQPU_HIP_FORCE_NO_DEVICE=1)");
    return NQPU_UNAVAILABLE;
  }
  std::unique_ptr<State> s(new State());
  s->n = n;
  s->prec = prec;
  HIPTRY(hipSetDevice(0), "hipSetDevice");
  s->device = 0;
  size_t bytes = s->amp_size() << n;
  size_t fr = 0, tot = 0;
  HIPTRY(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
  if (bytes + (size_t(1) << 29) > fr) {
    set_error("state of " + std::to_string(bytes) + " bytes does not fit in free device memory");
    return NQPU_OUT_OF_DEVICE_MEMORY;
  }
  HIPTRY(new_stream(&s->stream), "hipStreamCreate");
  HIPTRY(hipMalloc(&s->amps, bytes), "hipMalloc(state)");
  NqpuError e = reset_state(s.get());
  if (e != NQPU_OK) return e;
  *out = s.release();
  return NQPU_OK;
}

NqpuError reset_state(State* s) {
  s->perm.clear();  // |0...0> is the same in every layout
  dispatch(s, [&](auto* tag) {
    using R = std::remove_pointer_t<decltype(tag)>;
    Launch<R>::init_zero(s->amps, s->n, s->stream);
  });
  HIPTRY(hipGetLastError(), "init_zero launch");
  HIPTRY(hipStreamSynchronize(s->stream), "init_zero");
  return NQPU_OK;
}

NqpuError clone_state(const State* s, State** out) {
  *out = nullptr;
  std::unique_ptr<State> c(new State());
  c->n = s->n;
  c->prec = s->prec;
  c->device = s->device;
  size_t bytes = s->amp_size() << s->n;
  HIPTRY(new_stream(&c->stream), "hipStreamCreate");
  HIPTRY(hipMalloc(&c->amps, bytes), "hipMalloc(clone)");
  HIPTRY(hipStreamSynchronize(s->stream), "sync");
  HIPTRY(hipMemcpyAsync(c->amps, s->amps, bytes, hipMemcpyDeviceToDevice, c->stream), "clone copy");
  HIPTRY(hipStreamSynchronize(c->stream), "clone sync");
  c->perm = s->perm;  // the copy keeps the layout
  c->perm_k = s->perm_k;
  *out = c.release();
  return NQPU_OK;
}

static uint64_t op_bytes_direct(const HostOp& o, int n, size_t asz) {
  // touched amplitudes x (read + write)
  uint64_t dim = 1ull << n;
  int fixed = __builtin_popcountll(o.qubit_mask());
  switch (o.kind) {
    case OP_DIAGK: return 2 * dim * asz;
    case OP_D1: {
      bool only1 = (o.m[0] == cd(1, 0));
      uint64_t groups = dim >> fixed;
      return 2 * groups * (only1 ? 1 : 2) * asz;
    }
    default: {
      uint64_t groups = dim >> fixed;
      int per = (o.kind == OP_U2 || o.kind == OP_D2) ? 4 : 2;
      return 2 * groups * per * asz;
    }
  }
}

template <typename R>
static NqpuError exec_direct(State* s, const HostOp& o) {
  void* st = s->amps;
  const int n = s->n;
  hipStream_t q = s->stream;
  if (o.kind == OP_DIAGK) {
    // the previous LUT may still be in use: synchronise before overwriting it
    size_t lb = o.lut.size() * sizeof(double) * 2;
    HIPTRY(hipStreamSynchronize(q), "sync before LUT upload");
    if (!s->scratch2.ensure(lb)) return NQPU_OUT_OF_DEVICE_MEMORY;
    HIPTRY(copy_on(s->stream, s->scratch2.p, o.lut.data(), lb, hipMemcpyHostToDevice), "LUT upload");
  }
  s->timer.mark(q);
  switch (o.kind) {
    case OP_U1: Launch<R>::u1(st, n, o.a, o.cmask, o.m, q); break;
    case OP_FLIP: Launch<R>::flip(st, n, o.a, o.cmask, q); break;
    case OP_D1: Launch<R>::d1(st, n, o.a, o.cmask, o.m[0], o.m[1], q); break;
    case OP_U2: Launch<R>::u2(st, n, o.a, o.b, o.cmask, o.m, q); break;
    case OP_D2: Launch<R>::d2(st, n, o.a, o.b, o.cmask, o.m, q); break;
    case OP_SWAP: Launch<R>::swap(st, n, o.a, o.b, o.cmask, q); break;
    case OP_DIAGK: Launch<R>::diagk(st, n, (int)o.dq.size(), o.dq.data(), s->scratch2.p, q); break;
    default: break;
  }
  s->timer.mark(q);
  HIPTRY(hipGetLastError(), "direct kernel launch");
  s->stats.passes++;
  s->stats.bytes_moved += (double)op_bytes_direct(o, n, s->amp_size());
  return NQPU_OK;
}

// LUT items for a fused diagonal, at most 16 qubits each (execute_diagonal,
// gate_fusion.rs:905-1033: phase per bit pattern).
static std::vector<HostOp> diagonal_ops(const FusedItem& it) {
  std::vector<HostOp> v;
  std::vector<std::pair<int, std::pair<cd, cd>>> singles = it.single;
  std::vector<std::pair<std::pair<int, int>, cd>> pairs = it.pairs;
  while (!singles.empty() || !pairs.empty()) {
    HostOp o;
    o.kind = OP_DIAGK;
    std::vector<int> qs;
    auto has = [&](int q) { return std::find(qs.begin(), qs.end(), q) != qs.end(); };
    std::vector<std::pair<int, std::pair<cd, cd>>> s_take, s_rest;
    std::vector<std::pair<std::pair<int, int>, cd>> p_take, p_rest;
    for (auto& p : pairs) {
      int need = (has(p.first.first) ? 0 : 1) + (has(p.first.second) ? 0 : 1);
      if ((int)qs.size() + need <= 16) {
        if (!has(p.first.first)) qs.push_back(p.first.first);
        if (!has(p.first.second)) qs.push_back(p.first.second);
        p_take.push_back(p);
      } else {
        p_rest.push_back(p);
      }
    }
    for (auto& sp : singles) {
      if (has(sp.first) || qs.size() < 16) {
        if (!has(sp.first)) qs.push_back(sp.first);
        s_take.push_back(sp);
      } else {
        s_rest.push_back(sp);
      }
    }
    std::sort(qs.begin(), qs.end());
    auto pos = [&](int q) { return (int)(std::find(qs.begin(), qs.end(), q) - qs.begin()); };
    o.dq.assign(qs.begin(), qs.end());
    o.a = qs.empty() ? 0 : qs[0];
    size_t L = size_t(1) << qs.size();
    o.lut.assign(L, cd(1, 0));
    for (size_t pat = 0; pat < L; ++pat) {
      cd phs = 1;
      for (auto& sp : s_take) phs *= ((pat >> pos(sp.first)) & 1) ? sp.second.second : sp.second.first;
      for (auto& pp : p_take)
        if (((pat >> pos(pp.first.first)) & 1) && ((pat >> pos(pp.first.second)) & 1)) phs *= pp.second;
      o.lut[pat] = phs;
    }
    v.push_back(o);
    singles.swap(s_rest);
    pairs.swap(p_rest);
  }
  return v;
}

// Matrix-core block choices of register tile passes, by pass structure.  The
// choice (which pending ops form which block) depends only on the ops' kinds,
// tile bits, controls and qubits outside the tile, never on their angles, and
// the same structures recur across the circuits of a variational batch and the
// iterations of an optimisation loop.  An entry lists the program's items in
// order: {op} for an op applied alone, {-1, op, op, ...} for a block.
namespace {
using BlockChoice = std::vector<std::vector<int>>;
std::mutex g_block_mu;
std::unordered_map<std::string, std::shared_ptr<const BlockChoice>> g_block_cache;
constexpr size_t kBlockCacheMax = 1 << 16;  // entries; the cache starts over when full

std::shared_ptr<const BlockChoice> block_choice_find(const std::string& key) {
  std::lock_guard<std::mutex> lk(g_block_mu);
  auto it = g_block_cache.find(key);
  return it == g_block_cache.end() ? nullptr : it->second;
}

void block_choice_store(const std::string& key, BlockChoice&& c) {
  auto v = std::make_shared<const BlockChoice>(std::move(c));
  std::lock_guard<std::mutex> lk(g_block_mu);
  if (g_block_cache.size() >= kBlockCacheMax) g_block_cache.clear();
  g_block_cache.emplace(key, std::move(v));
}

template <typename T>
void key_put(std::string& k, T x) {
  k.append(reinterpret_cast<const char*>(&x), sizeof x);
}
}  // namespace

// Program of one register-resident tile pass (k_tile_reg).  Slots: thread
// slots 0..7 (code i) and register slots 0..eb-1 (code 16+j); map[] gives the
// tile-local bit each slot holds.  Ops are issued in order; an op runs in the
// current phase when its targets are register slots (diagonals and controls
// can sit anywhere; an uncontrolled SWAP only relabels two slots), and no op
// may overtake an earlier pending op on a shared qubit.  When nothing more
// can run, an R_REMAP moves the targets of the next pending ops into
// register slots.  The pass ends in the canonical slot assignment the store
// expects.
// A register-tile program: the ops, and for each R_MFMA op (s0 = index) its
// 16x16 block matrix in block-index order (a0, a1, a2, a3; see R_MFMA).
template <typename R>
struct RegProgram {
  std::vector<RegOp<R>> ops;
  std::vector<std::array<cd, 256>> blocks;
  // with defer_blocks: each block's row operations instead of its matrix
  // (block k's words start at block_at[k] in block_words; BR_* records)
  std::vector<uint64_t> block_words;
  std::vector<uint32_t> block_at;
  size_t nblocks() const { return blocks.size() + block_at.size(); }
  uint64_t fused_ops = 0;  // device ops inside those blocks
  // R_QPHASE tables, in operand units of kMfmaBlobBytes (an op's s0 counts
  // units from the start of qtab; the device program places them after the
  // blocks' operand blobs)
  std::vector<R> qtab;  // float on the device; the FP64 emulation keeps FP64 entries
  uint64_t qphase_gates = 0;  // diagonal gates merged into R_QPHASE ops
  // tile bit held by each slot (0..7 thread, 8.. register) when the tile is
  // loaded and when it is stored
  uint8_t map0[16] = {};
  uint8_t map1[16] = {};
};

// Fusion of dense runs into matrix-core blocks pays when the block replaces at
// least this much vector-ALU work, in units of one dense 2x2 gate.
constexpr double kMfmaMinWeight = 1.0;

template <typename R>
// xchg (full 256-thread tiles only): move bits between register slots and
// lane bits with R_XCHG instead of LDS remaps wherever the thread slots that
// change are lane bits, load the tile with the two least-used window bits on
// the wave slots, and store it in whatever layout the program ends in, as
// long as thread slots 0..4 hold the tile's five lowest bits again (so the
// store is as coalesced as the load).
RegProgram<R> build_reg_program(const Pass& p, int eb, int tb, bool mfma = false, bool xchg = false,
                                int max_blobs = 1 << 20, bool qphase = false, bool defer_blocks = false) {
  const int K = (int)p.qubits.size();  // == tb + eb
  // tile bits stored to the five lowest positions of the tile (the bits
  // loaded from there unless the pass permutes its positions)
  uint32_t store_low = 0;
  for (int j = 0; j < K && j < 32; ++j)
    if ((p.out.empty() ? p.qubits[j] : p.out[j]) < 5) store_low |= 1u << j;
  int loc[64];
  for (int q = 0; q < 64; ++q) loc[q] = -1;
  for (int j = 0; j < K; ++j) loc[p.qubits[j]] = j;
  // a, b: tile-local bits, or kGlobalBit + qubit for a diagonal factor on a
  // qubit outside the tile; gc: controls outside the tile
  struct L {
    int kind, a, b;
    uint32_t cm;
    uint64_t gc;
    const HostOp* h;
    bool done;
    int grp = -1;  // kQPhaseOp: index into qgroups
  };
  auto local = [&](int q) { return loc[q] >= 0 ? loc[q] : kGlobalBit + q; };
  std::vector<L> ops;
  ops.reserve(p.ops.size());
  for (const HostOp& o : p.ops) {
    L l{o.kind, local(o.a), 0, 0, 0, &o, false};
    if (o.kind == OP_U2 || o.kind == OP_SWAP || o.kind == OP_D2) l.b = local(o.b);
    for (int q = 0; q < 64; ++q)
      if ((o.cmask >> q) & 1) {
        if (loc[q] >= 0) l.cm |= 1u << loc[q];
        else l.gc |= bit(q);
      }
    ops.push_back(l);
  }
  auto two = [](int k) { return k == OP_U2 || k == OP_SWAP || k == OP_D2; };
  // tile bits an op touches (bits outside the tile are constant over it)
  auto lbit = [](int x) { return x < kGlobalBit ? 1u << x : 0u; };
  // ---- runs of one- and two-qubit diagonal gates become one R_QPHASE each.
  // Every such gate with unit-modulus phases is exp(2 pi i phi) with phi a
  // quadratic polynomial in at most two index bits (its positions and
  // controls), so a run of them is one quadratic form over the bits it
  // touches: Phi = c0 + sum lin[x] b_x + sum quad[x, y] b_x b_y (revolutions;
  // positions x are tile-local bits or kGlobalBit + qubit).
  struct QGroup {
    double c0 = 0;
    std::map<int, double> lin;
    std::map<std::pair<int, int>, double> quad;  // x < y
    uint32_t tile = 0;                           // tile bits it touches
    uint64_t gmask = 0;                          // qubits outside the tile it touches
    std::vector<std::vector<int>> chunks;        // those qubits, <= 8 per chunk, <= 4 chunks
    std::vector<int> members;
  };
  std::vector<QGroup> qgroups;
  constexpr int kQPhaseOp = 100;     // host-side kind of a merged run
  constexpr size_t kQPhaseMinOps = 3;  // shorter runs stay single ops (or join blocks)
  auto qmask = [&](const L& o) {
    if (o.kind == kQPhaseOp) return qgroups[o.grp].tile;
    return o.cm | lbit(o.a) | (two(o.kind) ? lbit(o.b) : 0u);
  };
  auto targets = [&](const L& o) -> uint32_t {
    switch (o.kind) {
      case OP_U1: case OP_FLIP: return 1u << o.a;
      case OP_U2: return (1u << o.a) | (1u << o.b);
      case OP_SWAP: return (o.cm || o.gc) ? ((1u << o.a) | (1u << o.b)) : 0u;
      default: return 0u;
    }
  };
  if (qphase && tb == 8) {
    // phase terms of one diagonal op: constant, up to two positions with
    // linear coefficients, and their product coefficient
    struct QTerms {
      double c0 = 0, lin[2] = {0, 0}, quad = 0;
      int np = 0, pos[2] = {0, 0};
    };
    auto rev = [](cd z) { return std::arg(z) / (2 * M_PI); };
    auto unit = [](cd z) { return std::abs(std::abs(z) - 1.0) < 1e-12; };
    auto qterms = [&](const L& o, QTerms* t) -> bool {
      if (o.kind != OP_D1 && o.kind != OP_D2) return false;
      int ctl[2], nc = 0;
      for (int q = 0; q < K; ++q)
        if ((o.cm >> q) & 1) {
          if (nc == 2) return false;
          ctl[nc++] = q;
        }
      for (int q = 0; q < 64; ++q)
        if ((o.gc >> q) & 1) {
          if (nc == 2) return false;
          ctl[nc++] = kGlobalBit + q;
        }
      const HostOp& h = *o.h;
      if (o.kind == OP_D1) {  // diag(m0, m1) on a, when the control (if any) is 1
        if (nc > 1 || !unit(h.m[0]) || !unit(h.m[1])) return false;
        const double t0 = rev(h.m[0]), t1 = rev(h.m[1]);
        t->pos[0] = o.a;
        if (nc == 0) {
          t->np = 1;
          t->c0 = t0;
          t->lin[0] = t1 - t0;
        } else {
          t->np = 2;
          t->pos[1] = ctl[0];
          t->lin[1] = t0;
          t->quad = t1 - t0;
        }
        return true;
      }
      if (nc > 0) return false;
      for (int i = 0; i < 4; ++i)
        if (!unit(h.m[i])) return false;
      // index bit_a | bit_b << 1
      const double t00 = rev(h.m[0]), t10 = rev(h.m[1]), t01 = rev(h.m[2]), t11 = rev(h.m[3]);
      t->np = 2;
      t->pos[0] = o.a;
      t->pos[1] = o.b;
      t->c0 = t00;
      t->lin[0] = t10 - t00;
      t->lin[1] = t01 - t00;
      t->quad = t11 - t10 - t01 + t00;
      return true;
    };
    // qubits outside the tile, in chunks of at most 8 (at most 4 chunks); two
    // of them with a product term must share a chunk
    auto pack = [](const std::vector<int>& gq, const std::vector<std::pair<int, int>>& gg,
                   std::vector<std::vector<int>>* out) -> bool {
      std::map<int, int> par;
      for (int q : gq) par[q] = q;
      std::function<int(int)> find = [&](int x) { return par[x] == x ? x : par[x] = find(par[x]); };
      for (const auto& e : gg) par[find(e.first)] = find(e.second);
      std::map<int, std::vector<int>> comp;
      for (int q : gq) comp[find(q)].push_back(q);
      std::vector<std::vector<int>> cs;
      for (auto& kv : comp) cs.push_back(kv.second);
      std::stable_sort(cs.begin(), cs.end(), [](const std::vector<int>& x, const std::vector<int>& y) {
        return x.size() > y.size();
      });
      std::vector<std::vector<int>> bins;
      for (const auto& c : cs) {
        if (c.size() > 8) return false;
        bool placed = false;
        for (auto& b : bins)
          if (b.size() + c.size() <= 8) {
            b.insert(b.end(), c.begin(), c.end());
            placed = true;
            break;
          }
        if (!placed) {
          if (bins.size() == 4) return false;
          bins.push_back(c);
        }
      }
      if (out) *out = bins;
      return true;
    };
    // A run grows in circuit order: a diagonal op joins the open run when no
    // non-diagonal op since the run's first member targets one of its tile
    // bits (so it commutes with every op it overtakes); otherwise it opens a
    // new run.  Diagonal ops that do not fit the form commute with the run.
    auto ntargets = [&](const L& o) -> uint32_t {
      switch (o.kind) {
        case OP_U1: case OP_FLIP: return lbit(o.a);
        case OP_U2: case OP_SWAP: return lbit(o.a) | lbit(o.b);
        default: return 0u;
      }
    };
    struct Run {
      std::vector<int> members, gq;
      std::vector<std::pair<int, int>> gg;
    };
    std::vector<Run> runs;
    std::vector<int> run_of(ops.size(), -1);
    int cur = -1;
    uint32_t blocked = 0;
    for (int k = 0; k < (int)ops.size(); ++k) {
      const L& o = ops[k];
      QTerms t;
      if (!qterms(o, &t)) {
        if (o.kind != OP_D1 && o.kind != OP_D2) blocked |= ntargets(o);
        continue;
      }
      uint32_t bits = 0;
      for (int i = 0; i < t.np; ++i) bits |= lbit(t.pos[i]);
      auto with = [&](Run r) {
        for (int i = 0; i < t.np; ++i)
          if (t.pos[i] >= kGlobalBit && std::find(r.gq.begin(), r.gq.end(), t.pos[i]) == r.gq.end())
            r.gq.push_back(t.pos[i]);
        if (t.np == 2 && t.pos[0] >= kGlobalBit && t.pos[1] >= kGlobalBit) r.gg.push_back({t.pos[0], t.pos[1]});
        r.members.push_back(k);
        return r;
      };
      bool joined = false;
      if (cur >= 0 && (bits & blocked) == 0) {
        Run r = with(runs[cur]);
        if (pack(r.gq, r.gg, nullptr)) {
          runs[cur] = std::move(r);
          joined = true;
        }
      }
      if (!joined) {
        runs.push_back(with(Run{}));
        cur = (int)runs.size() - 1;
        blocked = 0;
      }
      run_of[k] = cur;
    }
    std::vector<L> merged;
    merged.reserve(ops.size());
    for (int k = 0; k < (int)ops.size(); ++k) {
      const int ri = run_of[k];
      if (ri < 0 || runs[ri].members.size() < kQPhaseMinOps) {
        merged.push_back(ops[k]);
        continue;
      }
      const Run& run = runs[ri];
      if (run.members[0] != k) continue;  // merged into the run's first position
      QGroup g;
      for (int m : run.members) {
        QTerms t;
        qterms(ops[m], &t);
        g.c0 += t.c0;
        for (int i = 0; i < t.np; ++i) {
          g.lin[t.pos[i]] += t.lin[i];
          if (t.pos[i] < kGlobalBit) g.tile |= 1u << t.pos[i];
          else g.gmask |= bit(t.pos[i] - kGlobalBit);
        }
        if (t.np == 2) g.quad[{std::min(t.pos[0], t.pos[1]), std::max(t.pos[0], t.pos[1])}] += t.quad;
      }
      std::vector<std::vector<int>> ch;
      pack(run.gq, run.gg, &ch);
      for (auto& c : ch) {
        for (int& x : c) x -= kGlobalBit;  // positions -> qubits
        std::sort(c.begin(), c.end());
      }
      g.chunks = ch;
      g.members = run.members;
      qgroups.push_back(std::move(g));
      L l{kQPhaseOp, -1, -1, 0, 0, nullptr, false};
      l.grp = (int)qgroups.size() - 1;
      merged.push_back(l);
    }
    ops.swap(merged);
  }
  // canonical assignment: thread slot i holds tile bit i (i < tb), register
  // slot j holds tile bit tb + j; thread slots >= tb do not exist
  uint8_t map[16], canon[16];
  for (int i = 0; i < 16; ++i) map[i] = canon[i] = 0;
  for (int i = 0; i < tb; ++i) map[i] = canon[i] = (uint8_t)i;
  for (int j = 0; j < eb; ++j) map[8 + j] = canon[8 + j] = (uint8_t)(tb + j);
  auto slot = [&](int lb) {
    for (int i = 0; i < tb; ++i)
      if (map[i] == lb) return i;
    for (int j = 0; j < eb; ++j)
      if (map[8 + j] == lb) return 16 + j;
    return -1;
  };
  auto regmask = [&]() {
    uint32_t m = 0;
    for (int j = 0; j < eb; ++j) m |= 1u << map[8 + j];
    return m;
  };
  RegProgram<R> out;
  std::vector<RegOp<R>>& prog = out.ops;
  auto emit_remap = [&](const uint8_t* nm) {
    RegOp<R> r{};
    r.kind = R_REMAP;
    for (int i = 0; i < 16; ++i) {
      r.map_old[i] = map[i];
      r.map_new[i] = nm[i];
    }
    prog.push_back(r);
    for (int i = 0; i < 16; ++i) map[i] = nm[i];
  };
  xchg = xchg && tb == 8;
  // other: thread slot < 6, or 16 + register slot
  auto emit_xchg = [&](int j, int other) {
    RegOp<R> r{};
    r.kind = R_XCHG;
    r.s0 = j;
    r.s1 = other;
    prog.push_back(r);
    std::swap(map[8 + j], map[other < 16 ? other : 8 + (other - 16)]);
  };
  // reach the slot assignment nm: by exchanges when the wave slots (thread
  // slots 6, 7) keep their bits, else by one LDS remap
  // dist[b]: how many pending items come before the next one that needs tile
  // bit b on a register or lane-5 slot (large: never again)
  auto remap_to = [&](const uint8_t* nm_in, const int* dist) {
    uint8_t nm[16];
    std::memcpy(nm, nm_in, 16);
    if (xchg && (nm[6] != map[6] || nm[7] != map[7])) {
      // an LDS remap is unavoidable: it may as well park on the wave slots the
      // two bits whose next use is farthest away (Belady), so that later
      // items reach their bits by exchanges
      int order[16], cnt = 0;
      for (int t = 0; t < 16; ++t)
        if (t < tb || (t >= 8 && t < 8 + eb)) order[cnt++] = t;
      std::stable_sort(order, order + cnt, [&](int x, int y) { return dist[nm[x]] > dist[nm[y]]; });
      for (int w = 0; w < 2; ++w) {
        const int want = order[w];  // slot in nm holding a bit to park
        if (want == 6 || want == 7) continue;
        const int ws = (order[1 - w] == 6) ? 7 : (order[1 - w] == 7 ? 6 : (w == 0 ? 6 : 7));
        if (dist[nm[ws]] >= dist[nm[want]]) continue;
        std::swap(nm[want], nm[ws]);
      }
    }
    if (!xchg || nm[6] != map[6] || nm[7] != map[7]) {
      emit_remap(nm);
      return;
    }
    for (int i = 0; i < 6; ++i) {
      if (nm[i] == map[i]) continue;
      int where = -1;
      for (int t = 0; t < 16; ++t)
        if ((t < tb || (t >= 8 && t < 8 + eb)) && map[t] == nm[i]) where = t;
      if (where >= 8) {
        emit_xchg(where - 8, i);
      } else {  // in a later lane slot: through register slot 0
        emit_xchg(0, where);
        emit_xchg(0, i);
      }
    }
    for (int j = 0; j < eb; ++j) {
      if (nm[8 + j] == map[8 + j]) continue;
      for (int k = 0; k < eb; ++k)
        if (map[8 + k] == nm[8 + j]) {
          emit_xchg(j, 16 + k);
          break;
        }
    }
  };
  // R_QPHASE of a merged run, with its coefficients sorted by the slots that
  // hold their bits now
  auto emit_qphase = [&](const QGroup& g) {
    // code: thread slot (0..7), 16 + register slot, or kGlobalBit + qubit
    auto code = [&](int x) { return x >= kGlobalBit ? x : slot(x); };
    double lt[8] = {}, lr[8] = {}, att[8][8] = {}, atr[8][8] = {}, arr[8][8] = {};
    std::map<int, double> lg;
    std::map<int, std::array<double, 8>> agt, agr;
    std::map<std::pair<int, int>, double> agg;
    for (const auto& kv : g.lin) {
      const int c = code(kv.first);
      if (c >= kGlobalBit) lg[c - kGlobalBit] += kv.second;
      else if (c >= 16) lr[c - 16] += kv.second;
      else lt[c] += kv.second;
    }
    for (const auto& kv : g.quad) {
      int x = code(kv.first.first), y = code(kv.first.second);
      if (x > y) std::swap(x, y);  // thread < register < outside the tile
      const double a = kv.second;
      if (x >= kGlobalBit) agg[{x - kGlobalBit, y - kGlobalBit}] += a;
      else if (y >= kGlobalBit) {
        auto& row = x >= 16 ? agr[y - kGlobalBit] : agt[y - kGlobalBit];
        row[x >= 16 ? x - 16 : x] += a;
      } else if (x >= 16) arr[x - 16][y - 16] += a;
      else if (y >= 16) atr[x][y - 16] += a;
      else att[x][y] += a;
    }
    auto fr = [](double x) { return (R)(x - std::nearbyint(x)); };
    const int nch = (int)g.chunks.size();
    const size_t len = 256 * (size_t)(1 + eb) + (size_t)nch * kQPhaseChunkFloats;
    const size_t unit = kMfmaBlobBytes / sizeof(float);
    const size_t at = out.qtab.size();
    out.qtab.resize(at + (len + unit - 1) / unit * unit, (R)0);
    R* T = out.qtab.data() + at;
    // per thread: the sum for tid is the one for tid without its lowest set
    // bit t, plus t's linear term and its couplings to the other set bits
    // (all higher than t)
    double tc[256], tr[8][256];
    tc[0] = g.c0;
    for (int j = 0; j < eb; ++j) tr[j][0] = lr[j];
    for (int tid = 1; tid < 256; ++tid) {
      const int t = __builtin_ctz(tid), rest = tid & (tid - 1);
      double c = tc[rest] + lt[t];
      for (int m = rest; m; m &= m - 1) c += att[t][__builtin_ctz(m)];
      tc[tid] = c;
      for (int j = 0; j < eb; ++j) tr[j][tid] = tr[j][rest] + atr[t][j];
    }
    for (int tid = 0; tid < 256; ++tid) {
      T[tid] = fr(tc[tid]);
      for (int j = 0; j < eb; ++j) T[256 * (1 + j) + tid] = fr(tr[j][tid]);
    }
    RegOp<R> r{};
    r.kind = R_QPHASE;
    r.s0 = (int32_t)(at / unit);
    r.pad = nch;
    for (int k = 1; k < eb; ++k)
      for (int j = 0; j < k; ++j) r.m[k * (k - 1) / 2 + j] = (R)fr(arr[j][k]);
    R* C = T + 256 * (1 + eb);
    for (int c = 0; c < nch; ++c) {
      const std::vector<int>& qs = g.chunks[c];  // ascending
      const int cnt = (int)qs.size();
      uint64_t cm = 0;
      for (int q : qs) cm |= bit(q);
      set_qphase_chunk(r, c, cm);
      // per chunk qubit i: its linear term, its couplings to the chunk's other
      // qubits and to the thread and register slots
      double li[8] = {}, aii[8][8] = {}, kti[8][8] = {}, kri[8][8] = {};
      for (int i = 0; i < cnt; ++i) {
        auto lf = lg.find(qs[i]);
        if (lf != lg.end()) li[i] = lf->second;
        for (int i2 = 0; i2 < i; ++i2) {
          auto gi = agg.find({qs[i2], qs[i]});
          if (gi != agg.end()) aii[i][i2] = gi->second;
        }
        auto ti = agt.find(qs[i]);
        if (ti != agt.end())
          for (int t = 0; t < 8; ++t) kti[i][t] = ti->second[t];
        auto ri = agr.find(qs[i]);
        if (ri != agr.end())
          for (int j = 0; j < eb; ++j) kri[i][j] = ri->second[j];
      }
      // a pattern's terms are those of the pattern without its highest set
      // bit i, plus i's linear term, couplings to the rest and slot terms
      double v[256][17];
      for (int x = 0; x < 17; ++x) v[0][x] = 0;
      for (int pat = 1; pat < (1 << cnt); ++pat) {
        const int i = 31 - __builtin_clz((unsigned)pat), rest = pat & ~(1 << i);
        double k0 = v[rest][0] + li[i];
        for (int i2 = 0; i2 < i; ++i2)
          if ((rest >> i2) & 1) k0 += aii[i][i2];
        v[pat][0] = k0;
        for (int t = 0; t < 8; ++t) v[pat][1 + t] = v[rest][1 + t] + kti[i][t];
        for (int j = 0; j < eb; ++j) v[pat][9 + j] = v[rest][9 + j] + kri[i][j];
      }
      for (int pat = 0; pat < (1 << cnt); ++pat) {
        R* e = C + (size_t)c * kQPhaseChunkFloats + 16 * pat;
        e[0] = fr(v[pat][0]);
        for (int t = 0; t < 8; ++t) e[1 + t] = fr(v[pat][1 + t]);
        for (int j = 0; j < eb; ++j) e[9 + j] = fr(v[pat][9 + j]);
      }
    }
    out.qphase_gates += g.members.size();
    prog.push_back(r);
  };
  auto emit = [&](const L& o) {
    if (o.kind == kQPhaseOp) {
      emit_qphase(qgroups[o.grp]);
      return;
    }
    if (o.kind == OP_SWAP && o.cm == 0 && o.gc == 0) {  // relabel
      int sa = slot(o.a), sb = slot(o.b);
      int ia = sa >= 16 ? 8 + (sa - 16) : sa, ib = sb >= 16 ? 8 + (sb - 16) : sb;
      std::swap(map[ia], map[ib]);
      return;
    }
    RegOp<R> r{};
    for (int q = 0; q < K; ++q)
      if ((o.cm >> q) & 1) {
        int sq = slot(q);
        if (sq >= 16) r.cr |= 1u << (sq - 16);
        else r.ct |= 1u << sq;
      }
    r.gc = o.gc;
    auto dpos = [&](int x) { return x >= kGlobalBit ? x : slot(x); };  // diagonal factor position
    const HostOp& h = *o.h;
    int nm = 0;
    switch (o.kind) {
      case OP_U1: r.kind = R_U1; r.s0 = slot(o.a) - 16; nm = 4; break;
      case OP_FLIP: r.kind = R_FLIP; r.s0 = slot(o.a) - 16; break;
      case OP_D1: r.kind = R_D1; r.s0 = dpos(o.a); nm = 2; break;
      case OP_D2: r.kind = R_D2; r.s0 = dpos(o.a); r.s1 = dpos(o.b); nm = 4; break;
      case OP_U2: r.kind = R_U2; r.s0 = slot(o.a) - 16; r.s1 = slot(o.b) - 16; nm = 16; break;
      case OP_SWAP: r.kind = R_SWAP; r.s0 = slot(o.a) - 16; r.s1 = slot(o.b) - 16; break;
      default: return;
    }
    for (int i = 0; i < nm; ++i) {
      r.m[2 * i] = (R)h.m[i].real();
      r.m[2 * i + 1] = (R)h.m[i].imag();
    }
    prog.push_back(r);
  };
  // ---- schedule items: single ops, or (mfma) blocks of ops on <= 4 tile bits
  struct Item {
    int op;     // index into ops, or -1 for a block
    int blk;    // index into blks
    uint32_t qm;
    bool done;
  };
  // A block acts on <= 4 tile bits (q) and may depend on <= kMfmaMaxSel
  // qubits outside the tile (g: their controls and diagonal factors); it is
  // applied with one of 2^|g| matrices, chosen per tile from its base index.
  struct Blk {
    uint32_t q = 0;
    uint64_t g = 0;
    std::vector<int> ops;
  };
  constexpr int kMfmaMaxSel = 2;
  auto gmask = [&](const L& o) {
    if (o.kind == kQPhaseOp) return qgroups[o.grp].gmask;
    uint64_t m = o.gc;
    if (o.a >= kGlobalBit) m |= bit(o.a - kGlobalBit);
    if (two(o.kind) && o.b >= kGlobalBit) m |= bit(o.b - kGlobalBit);
    return m;
  };
  std::vector<Item> items;
  std::vector<Blk> blks;
  // (register slot 3, and 4 with 13-qubit tiles, are batch bits of a block)
  const bool use_mfma = mfma && (eb == 4 || eb == 5) && tb == 8;
  std::string choice_key;
  std::shared_ptr<const BlockChoice> choice;
  if (use_mfma) {
    key_put(choice_key, K);
    key_put(choice_key, max_blobs);
    for (const L& o : ops) {
      key_put(choice_key, o.kind);
      key_put(choice_key, o.a);
      key_put(choice_key, o.b);
      key_put(choice_key, o.cm);
      key_put(choice_key, o.gc);
      if (o.kind == kQPhaseOp) {
        key_put(choice_key, qgroups[o.grp].tile);
        key_put(choice_key, qgroups[o.grp].gmask);
      }
    }
    choice = block_choice_find(choice_key);
  }
  if (choice) {
    for (const std::vector<int>& rec : *choice) {
      if (rec[0] >= 0) {
        items.push_back(Item{rec[0], -1, qmask(ops[rec[0]]), false});
        continue;
      }
      Blk b;
      b.ops.assign(rec.begin() + 1, rec.end());
      for (int k : b.ops) {
        b.q |= qmask(ops[k]);
        b.g |= gmask(ops[k]);
      }
      blks.push_back(b);
      items.push_back(Item{-1, (int)blks.size() - 1, b.q, false});
    }
  } else if (use_mfma) {
    auto fusable = [&](const L& o) {
      if (__builtin_popcountll(gmask(o)) > kMfmaMaxSel) return false;
      return o.kind == OP_U1 || o.kind == OP_FLIP || o.kind == OP_D1 || o.kind == OP_D2 || o.kind == OP_U2 ||
             o.kind == OP_SWAP;
    };
    // vector-ALU work of an op, in dense-2x2 units (an uncontrolled SWAP is a free relabelling)
    auto weight = [&](const L& o) {
      switch (o.kind) {
        case OP_U1: return 1.0;
        case OP_U2: return 4.0;
        case OP_SWAP: return (o.cm || o.gc) ? 0.5 : 0.0;
        default: return 0.35;
      }
    };
    // A block is the light cone of a set S of at most four tile bits: every
    // pending op whose tile bits lie in S, whose qubits outside the tile fit
    // the selector set, and which no op left behind precedes on a shared bit.
    // An op left behind blocks its bits, so every op taken commutes with all
    // the ops it overtakes.  Each round evaluates the cone of every 4-subset
    // of the tile bits the next pending ops touch (at most C(12, 4) = 495
    // scans of the pending ops) and keeps the heaviest: growing one set from a
    // seed op in circuit order lets the first ops met claim the four bits, and
    // gave 27 % more blocks on a 30-qubit U3 + CNOT brickwork.
    std::vector<int> rem(ops.size());
    for (int k = 0; k < (int)ops.size(); ++k) rem[k] = k;
    // tail0[i]: some op at or after rem[i] has no tile bits and could still
    // join a block whose bits are all blocked
    std::vector<char> tail0;
    auto cone = [&](uint32_t S, std::vector<int>* taken, std::vector<int>* left) {
      uint32_t blocked = 0;
      uint64_t G = 0;
      double w = 0;
      for (size_t i = 0; i < rem.size(); ++i) {
        // a weight-only scan ends once every bit of S is blocked: nothing
        // after that point can join
        if (!taken && (blocked & S) == S && !tail0[i]) break;
        const int k = rem[i];
        const L& o = ops[k];
        const uint32_t qm = qmask(o);
        bool ok = (qm & blocked) == 0 && (qm & ~S) == 0 && fusable(o);
        if (ok) {
          const uint64_t G2 = G | gmask(o);
          ok = __builtin_popcountll(G2) <= kMfmaMaxSel;
          if (ok) G = G2;
        }
        if (ok) {
          w += weight(o);
          if (taken) taken->push_back(k);
        } else {
          blocked |= qm;
          if (left) left->push_back(k);
        }
      }
      return w;
    };
    constexpr int kWindow = 24;  // pending ops whose bits are candidates
    int nvar = 0;  // operand blobs of the blocks so far (one per variant)
    while (!rem.empty()) {
      const L& first = ops[rem.front()];
      if (!fusable(first) || qmask(first) == 0) {  // runs on its own
        items.push_back(Item{rem.front(), -1, qmask(first), false});
        rem.erase(rem.begin());
        continue;
      }
      tail0.assign(rem.size() + 1, 0);
      for (size_t i = rem.size(); i-- > 0;)
        tail0[i] = tail0[i + 1] || (qmask(ops[rem[i]]) == 0 && fusable(ops[rem[i]]));
      uint32_t touched = 0;
      int seen = 0;
      for (int k : rem) {
        if (seen++ == kWindow) break;
        if (fusable(ops[k])) touched |= qmask(ops[k]);
      }
      uint32_t bestS = touched;
      double bw = -1;
      if (__builtin_popcount(touched) <= 4) {
        bw = cone(touched, nullptr, nullptr);
      } else {
        int bits[32], nb = 0;
        for (int q = 0; q < K; ++q)
          if ((touched >> q) & 1) bits[nb++] = q;
        const uint32_t f = qmask(first);
        for (int i0 = 0; i0 < nb; ++i0)
          for (int i1 = i0 + 1; i1 < nb; ++i1)
            for (int i2 = i1 + 1; i2 < nb; ++i2)
              for (int i3 = i2 + 1; i3 < nb; ++i3) {
                const uint32_t S = (1u << bits[i0]) | (1u << bits[i1]) | (1u << bits[i2]) | (1u << bits[i3]);
                // ties go to a set holding the first pending op's bits
                const double w = cone(S, nullptr, nullptr) + ((S & f) == f ? 1e-6 : 0.0);
                if (w > bw + 1e-9) {
                  bw = w;
                  bestS = S;
                }
              }
      }
      std::vector<int> taken, left;
      cone(bestS, &taken, &left);
      if (taken.empty()) {  // nothing weighs anything (relabelling SWAPs only): the first op alone
        items.push_back(Item{rem.front(), -1, qmask(first), false});
        rem.erase(rem.begin());
        continue;
      }
      Blk b;
      for (int k : taken) {
        b.q |= qmask(ops[k]);
        b.g |= gmask(ops[k]);
      }
      b.ops = taken;
      if (bw >= kMfmaMinWeight && nvar + (1 << __builtin_popcountll(b.g)) <= max_blobs) {
        nvar += 1 << __builtin_popcountll(b.g);
        blks.push_back(b);
        items.push_back(Item{-1, (int)blks.size() - 1, b.q, false});
      } else {
        for (int k : taken) items.push_back(Item{k, -1, qmask(ops[k]), false});
      }
      rem.swap(left);
    }
    BlockChoice rec;
    for (const Item& it : items) {
      if (it.op >= 0) {
        rec.push_back({it.op});
        continue;
      }
      std::vector<int> r{-1};
      r.insert(r.end(), blks[it.blk].ops.begin(), blks[it.blk].ops.end());
      rec.push_back(std::move(r));
    }
    block_choice_store(choice_key, std::move(rec));
  } else {
    for (int k = 0; k < (int)ops.size(); ++k) items.push_back(Item{k, -1, qmask(ops[k]), false});
  }
  if (xchg) {
    // the two window bits first needed on register or lane-5 slots last go
    // to the wave slots, where only an LDS remap can reach them
    int first[32];
    for (int b = 0; b < 32; ++b) first[b] = 1 << 20;
    for (int c = (int)items.size() - 1; c >= 0; --c) {
      const Item& it = items[c];
      const uint32_t m = it.op >= 0 ? targets(ops[it.op]) : blks[it.blk].q;
      for (int b = 0; b < K; ++b)
        if ((m >> b) & 1) first[b] = c;
    }
    std::vector<int> win;
    for (int b = 5; b < K; ++b) win.push_back(b);
    std::stable_sort(win.begin(), win.end(), [&](int x, int y) { return first[x] > first[y]; });
    std::vector<int> rest(win.begin() + 2, win.end());
    std::sort(rest.begin(), rest.end());
    map[6] = (uint8_t)win[0];
    map[7] = (uint8_t)win[1];
    map[5] = (uint8_t)rest[0];
    for (int j = 0; j < eb; ++j) map[8 + j] = (uint8_t)rest[1 + j];
  }
  std::memcpy(out.map0, map, 16);
  // positions a block uses: block bits a0, a1, a3 on register slots 0, 1, 2 and a2 on thread slot 5
  auto block_slots = [&]() {
    return (1u << map[8]) | (1u << map[9]) | (1u << map[10]) | (1u << map[5]);
  };
  auto emit_block = [&](const Blk& b) {
    int pos[32];
    for (int i = 0; i < 32; ++i) pos[i] = -1;
    pos[map[8]] = 0;
    pos[map[9]] = 1;
    pos[map[5]] = 2;
    pos[map[10]] = 3;
    int sel[kMfmaMaxSel], ns = 0;
    for (int q = 0; q < 64; ++q)
      if ((b.g >> q) & 1) sel[ns++] = q;
    RegOp<R> r{};
    r.kind = R_MFMA;
    r.s0 = (int32_t)out.nblocks();
    r.pad = ns;
    r.s1 = (ns > 0 ? sel[0] : 0) | (ns > 1 ? sel[1] << 8 : 0);
    for (int var = 0; var < (1 << ns); ++var) {
      uint64_t gset = 0;  // qubits outside the tile that are 1 in this variant
      for (int i = 0; i < ns; ++i)
        if ((var >> i) & 1) gset |= bit(sel[i]);
      // bit of block index i at op position x (a tile bit, or a qubit outside the tile)
      auto bitat = [&](int x, int i) -> int {
        return x >= kGlobalBit ? (int)((gset >> (x - kGlobalBit)) & 1) : (i >> pos[x]) & 1;
      };
      if (defer_blocks) {
        // the variant's row operations, with the qubits outside the tile
        // resolved: a diagonal factor on them becomes a fixed factor
        out.block_at.push_back((uint32_t)out.block_words.size());
        auto put = [&](uint32_t kind, int pa, int pb, uint32_t cmk, std::initializer_list<cd> vals) {
          out.block_words.push_back(block_row_header(kind, pa, pb, cmk, (int)vals.size() * 2));
          for (const cd& z : vals) {
            out.block_words.push_back(__builtin_bit_cast(uint64_t, z.real()));
            out.block_words.push_back(__builtin_bit_cast(uint64_t, z.imag()));
          }
        };
        for (int k : b.ops) {
          const L& o = ops[k];
          if ((o.gc & ~gset) != 0) continue;  // a control outside the tile is 0 in this variant
          const HostOp& h = *o.h;
          uint32_t cmk = 0;
          for (int q = 0; q < K; ++q)
            if ((o.cm >> q) & 1) cmk |= 1u << pos[q];
          const bool ina = o.a < kGlobalBit, inb = two(o.kind) && o.b < kGlobalBit;
          const int pa = ina ? pos[o.a] : 0, pb = inb ? pos[o.b] : 0;
          switch (o.kind) {
            case OP_U1: put(BR_U1, pa, 0, cmk, {h.m[0], h.m[1], h.m[2], h.m[3]}); break;
            case OP_FLIP: put(BR_FLIP, pa, 0, cmk, {}); break;
            case OP_SWAP: put(BR_SWAP, pa, pb, cmk, {}); break;
            case OP_U2: {
              out.block_words.push_back(block_row_header(BR_U2, pa, pb, cmk, 32));
              for (int e = 0; e < 16; ++e) {
                out.block_words.push_back(__builtin_bit_cast(uint64_t, h.m[e].real()));
                out.block_words.push_back(__builtin_bit_cast(uint64_t, h.m[e].imag()));
              }
              break;
            }
            case OP_D1:
              if (ina) put(BR_SCALE1, pa, 0, cmk, {h.m[0], h.m[1]});
              else put(BR_SCALE0, 0, 0, cmk, {h.m[bitat(o.a, 0)]});
              break;
            case OP_D2:
              if (ina && inb) put(BR_SCALE2, pa, pb, cmk, {h.m[0], h.m[1], h.m[2], h.m[3]});
              else if (ina) put(BR_SCALE1, pa, 0, cmk, {h.m[bitat(o.b, 0) << 1], h.m[1 | bitat(o.b, 0) << 1]});
              else if (inb) put(BR_SCALE1, pb, 0, cmk, {h.m[bitat(o.a, 0)], h.m[bitat(o.a, 0) | 2]});
              else put(BR_SCALE0, 0, 0, cmk, {h.m[bitat(o.a, 0) | bitat(o.b, 0) << 1]});
              break;
            default: break;
          }
        }
        continue;
      }
      // M = op_k ... op_1: each op is a row operation on M (row: block index
      // of the output, column: of the input), applied to all 16 columns at
      // once, with separate real and imaginary planes
      double Mr[16][16], Mi[16][16];
      for (int i = 0; i < 16; ++i)
        for (int c = 0; c < 16; ++c) {
          Mr[i][c] = i == c ? 1.0 : 0.0;
          Mi[i][c] = 0.0;
        }
      auto scale_row = [&](int i, cd z) {
        const double zr = z.real(), zi = z.imag();
        for (int c = 0; c < 16; ++c) {
          const double xr = Mr[i][c], xi = Mi[i][c];
          Mr[i][c] = zr * xr - zi * xi;
          Mi[i][c] = zr * xi + zi * xr;
        }
      };
      auto swap_rows = [&](int i, int j) {
        for (int c = 0; c < 16; ++c) {
          std::swap(Mr[i][c], Mr[j][c]);
          std::swap(Mi[i][c], Mi[j][c]);
        }
      };
      for (int k : b.ops) {
        const L& o = ops[k];
        if ((o.gc & ~gset) != 0) continue;  // a control outside the tile is 0 in this variant
        const HostOp& h = *o.h;
        uint32_t cmk = 0;
        for (int q = 0; q < K; ++q)
          if ((o.cm >> q) & 1) cmk |= 1u << pos[q];
        const int pa = o.a < kGlobalBit ? pos[o.a] : 0, pb = (two(o.kind) && o.b < kGlobalBit) ? pos[o.b] : 0;
        for (int i = 0; i < 16; ++i) {
          if ((i & cmk) != cmk) continue;
          switch (o.kind) {
            case OP_U1:
              if (!((i >> pa) & 1)) {
                const int j = i | (1 << pa);
                const double m0r = h.m[0].real(), m0i = h.m[0].imag(), m1r = h.m[1].real(), m1i = h.m[1].imag();
                const double m2r = h.m[2].real(), m2i = h.m[2].imag(), m3r = h.m[3].real(), m3i = h.m[3].imag();
                for (int c = 0; c < 16; ++c) {
                  const double xr = Mr[i][c], xi = Mi[i][c], yr = Mr[j][c], yi = Mi[j][c];
                  Mr[i][c] = m0r * xr - m0i * xi + m1r * yr - m1i * yi;
                  Mi[i][c] = m0r * xi + m0i * xr + m1r * yi + m1i * yr;
                  Mr[j][c] = m2r * xr - m2i * xi + m3r * yr - m3i * yi;
                  Mi[j][c] = m2r * xi + m2i * xr + m3r * yi + m3i * yr;
                }
              }
              break;
            case OP_FLIP:
              if (!((i >> pa) & 1)) swap_rows(i, i | (1 << pa));
              break;
            case OP_D1: scale_row(i, h.m[bitat(o.a, i)]); break;
            case OP_D2: scale_row(i, h.m[bitat(o.a, i) | (bitat(o.b, i) << 1)]); break;
            case OP_U2:
              if (!((i >> pa) & 1) && !((i >> pb) & 1)) {
                const int idx[4] = {i, i | (1 << pa), i | (1 << pb), i | (1 << pa) | (1 << pb)};
                for (int c = 0; c < 16; ++c) {
                  double xr[4], xi[4];
                  for (int t = 0; t < 4; ++t) {
                    xr[t] = Mr[idx[t]][c];
                    xi[t] = Mi[idx[t]][c];
                  }
                  for (int r = 0; r < 4; ++r) {
                    double sr = 0, si = 0;
                    for (int t = 0; t < 4; ++t) {
                      const double mr = h.m[4 * r + t].real(), mi = h.m[4 * r + t].imag();
                      sr += mr * xr[t] - mi * xi[t];
                      si += mr * xi[t] + mi * xr[t];
                    }
                    Mr[idx[r]][c] = sr;
                    Mi[idx[r]][c] = si;
                  }
                }
              }
              break;
            case OP_SWAP:
              if (((i >> pa) & 1) && !((i >> pb) & 1)) swap_rows(i, (i ^ (1 << pa)) | (1 << pb));
              break;
            default: break;
          }
        }
      }
      std::array<cd, 256> M;
      for (int i = 0; i < 16; ++i)
        for (int c = 0; c < 16; ++c) M[16 * i + c] = cd(Mr[i][c], Mi[i][c]);
      out.blocks.push_back(M);
    }
    out.fused_ops += b.ops.size();
    prog.push_back(r);
    std::swap(map[5], map[9]);  // the result holds a1 on thread slot 5 and a2 on register slot 1
  };
  size_t remaining = items.size();
  while (true) {
    // a relabelling SWAP can make earlier-skipped ops runnable: rescan
    bool progress = true;
    while (progress && remaining) {
      progress = false;
      uint32_t blocked = 0;
      for (Item& it : items) {
        if (it.done) continue;
        bool ready = (it.qm & blocked) == 0;
        if (ready) ready = it.op >= 0 ? (targets(ops[it.op]) & ~regmask()) == 0
                                      : (blks[it.blk].q & ~block_slots()) == 0;
        if (ready) {
          if (it.op >= 0) emit(ops[it.op]);
          else emit_block(blks[it.blk]);
          it.done = true;
          --remaining;
          progress = true;
        } else {
          blocked |= it.qm;
        }
      }
    }
    if (remaining == 0) break;
    int dist[32];
    for (int b = 0; b < 32; ++b) dist[b] = 1 << 20;
    if (xchg) {
      int c = 0;
      for (const Item& it : items) {
        if (it.done) continue;
        const uint32_t req = it.op >= 0 ? targets(ops[it.op]) : blks[it.blk].q;
        for (int b = 0; b < K; ++b)
          if (((req >> b) & 1) && dist[b] > c) dist[b] = c;
        ++c;
      }
      // with exchanges, the bits stored to the five lowest positions must be
      // on lane slots 0..4 at the store: a use after the last item
      if (xchg)
        for (int b = 0; b < K; ++b)
          if (((store_low >> b) & 1) && dist[b] > c) dist[b] = c;
    }
    uint8_t nm[16];
    for (int i = 0; i < 16; ++i) nm[i] = map[i];
    const Item* first = nullptr;
    for (const Item& it : items)
      if (!it.done) {
        first = &it;
        break;
      }
    if (first->op < 0 && xchg && (eb == 4 || eb == 5)) {
      // the block's bits must sit on register slots 0..2 and lane slot 5.
      // Each bit elsewhere comes into a register slot whose bit is needed
      // last (one exchange); lane 5 is refilled from a register only when all
      // four block bits ended up in registers (one lane-half swap); register
      // order is fixed last (register renaming).
      const uint32_t B = blks[first->blk].q;
      auto find = [&](int lb) {
        for (int i = 0; i < 16; ++i)
          if ((i < tb || (i >= 8 && i < 8 + eb)) && nm[i] == lb) return i;
        return -1;
      };
      for (uint32_t miss = B; miss; miss &= miss - 1) {
        const int lb = __builtin_ctz(miss);
        const int from = find(lb);
        if (from == 5 || from >= 8) continue;
        int j = -1;
        for (int r = 8; r < 8 + eb; ++r)
          if (!((B >> nm[r]) & 1) && (j < 0 || dist[nm[r]] > dist[nm[j]])) j = r;
        std::swap(nm[from], nm[j]);
      }
      if ((B >> nm[5]) & 1) {
        // lane 5 already holds a block bit
      } else if (__builtin_popcount(B) == 4) {
        for (int r = 8; r < 8 + eb; ++r)
          if ((B >> nm[r]) & 1) {
            std::swap(nm[5], nm[r]);
            break;
          }
      }
      for (int h = 11; h < 8 + eb; ++h)  // register slots 3.. are not part of the block
        if ((B >> nm[h]) & 1)
          for (int r = 8; r < 11; ++r)
            if (!((B >> nm[r]) & 1)) {
              std::swap(nm[h], nm[r]);
              break;
            }
      remap_to(nm, dist);
      continue;
    }
    if (first->op < 0) {
      // move the block's bits onto register slots 0..2 and thread slot 5,
      // keeping those already there
      const int P[4] = {8, 9, 10, 5};
      uint32_t miss = blks[first->blk].q & ~block_slots();
      for (int pi = 0; pi < 4 && miss; ++pi) {
        if ((blks[first->blk].q >> nm[P[pi]]) & 1) continue;  // holds a block bit already
        const int lb = __builtin_ctz(miss);
        miss &= miss - 1;
        int from = -1;
        for (int i = 0; i < 16; ++i)
          if ((i < tb || (i >= 8 && i < 8 + eb)) && nm[i] == lb) from = i;
        std::swap(nm[from], nm[P[pi]]);
      }
      remap_to(nm, dist);
      continue;
    }
    uint32_t blocked = 0;
    // register set for the next phase: targets of the pending ops, in order
    uint32_t need = 0;
    for (const Item& it : items) {
      if (it.done) continue;
      const uint32_t tg = it.op >= 0 ? targets(ops[it.op]) : ~0u;
      if ((it.qm & blocked) == 0 && it.op >= 0 && __builtin_popcount(need | tg) <= eb) need |= tg;
      else blocked |= it.qm;
      if (__builtin_popcount(need) == eb) break;
    }
    for (int i = 0; i < 16; ++i) nm[i] = map[i];
    std::vector<int> free_reg, evicted;
    for (int j = 0; j < eb; ++j) {
      if ((need >> map[8 + j]) & 1) need &= ~(1u << map[8 + j]);  // already in place
      else free_reg.push_back(j);
    }
    size_t fi = 0;
    for (int i = 0; i < tb && need; ++i) {
      const int lb = map[i];
      if (!((need >> lb) & 1)) continue;
      // thread slot i holds a needed bit: swap it with a register slot
      const int j = free_reg[fi++];
      nm[i] = map[8 + j];
      nm[8 + j] = (uint8_t)lb;
      need &= ~(1u << lb);
    }
    remap_to(nm, dist);
  }
  if (xchg) {
    // thread slots 0..4 must hold the bits stored to the five lowest
    // positions (any order)
    const uint32_t low = store_low;
    for (int i = 0; i < 5; ++i) {
      if ((low >> map[i]) & 1) continue;
      int where = -1;  // prefer a register slot, then lane slot 5, then a wave slot
      for (int t : {8, 9, 10, 11, 12, 5, 6, 7})
        if (where < 0 && (t < tb || t < 8 + eb) && ((low >> map[t]) & 1)) where = t;
      if (where >= 8) {
        emit_xchg(where - 8, i);
      } else if (where == 5) {
        emit_xchg(0, 5);
        emit_xchg(0, i);
      } else {  // on a wave slot: one LDS remap that brings all of them to lane slots 0..4
        uint8_t nm[16];
        std::memcpy(nm, map, 16);
        for (int k = 0; k < 5; ++k) {
          if ((low >> nm[k]) & 1) continue;
          for (int t = 5; t < 16; ++t)
            if ((t < tb || (t >= 8 && t < 8 + eb)) && ((low >> nm[t]) & 1)) {
              std::swap(nm[k], nm[t]);
              break;
            }
        }
        emit_remap(nm);
        break;
      }
    }
  } else if (std::memcmp(map, canon, 16) != 0) {
    emit_remap(canon);
  }
  std::memcpy(out.map1, map, 16);
  return out;
}

// IEEE binary16 of x, round to nearest even (normals, subnormals, overflow to
// infinity); the operand blobs of R_MFMA ops are built with it.
static uint16_t to_half(double x) {
  const float f = (float)x;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t sign = (u >> 16) & 0x8000u;
  const int exp = (int)((u >> 23) & 0xff) - 127;
  uint32_t mant = u & 0x7fffffu;
  if (exp == 128) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));  // inf / nan
  if (exp > 15) return (uint16_t)(sign | 0x7c00u);
  if (exp >= -14) {  // normal: keep 10 of 23 mantissa bits
    uint32_t h = ((uint32_t)(exp + 15) << 10) | (mant >> 13);
    const uint32_t rest = mant & 0x1fffu;
    if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) ++h;  // may carry into the exponent: still right
    return (uint16_t)(sign | h);
  }
  if (exp < -25) return (uint16_t)sign;
  // subnormal: value = m * 2^-24 with m = (1.mant) * 2^(exp + 24)
  mant |= 0x800000u;
  const int shift = -exp - 1;  // 14 .. 24: bits dropped from the 24-bit significand to reach 2^-24 units
  uint32_t h = mant >> shift;
  const uint32_t rest = mant & ((1u << shift) - 1), half = 1u << (shift - 1);
  if (rest > half || (rest == half && (h & 1u))) ++h;
  return (uint16_t)(sign | h);
}

// Exact value of a binary16 (bit arithmetic: the operand blobs of a batch of
// circuits convert millions of values).
static double from_half(uint16_t h) {
  const uint32_t sign = (uint32_t)(h >> 15) << 31, exp = (h >> 10) & 0x1f, mant = h & 0x3ff;
  float f;
  if (exp == 0) {  // zero or subnormal: mant * 2^-24, exact in float
    f = (float)mant * 5.9604644775390625e-8f;
    if (sign) f = -f;
  } else {
    const uint32_t u = sign | (exp == 31 ? 0x7f800000u : (exp + 112) << 23) | (mant << 13);
    std::memcpy(&f, &u, 4);
  }
  return f;
}

bool mfma_blob_fast_available();                // blob.cpp (F16C)
void mfma_blob_fast(const cd* M, uint16_t* out);  // the same blob, 8 values per conversion

// Operand blob of one R_MFMA op from its 16x16 block matrix (block-index order).
static void mfma_blob(const std::array<cd, 256>& M, uint16_t* out /* kMfmaBlobBytes / 2 */) {
  if (mfma_blob_fast_available()) {
    mfma_blob_fast(M.data(), out);
    return;
  }
  auto A = [&](int m, int k) {  // 32x32 real form
    const cd u = M[16 * (m >> 1) + (k >> 1)];
    const int cr = m & 1, cc = k & 1;
    return cr == 0 ? (cc == 0 ? u.real() : -u.imag()) : (cc == 0 ? u.imag() : u.real());
  };
  for (int l = 0; l < 64; ++l)
    for (int s = 0; s < 2; ++s)
      for (int i = 0; i < 8; ++i) {
        const double x = A(l % 32, 16 * s + 8 * (l / 32) + i);
        const uint16_t hi = to_half(x);
        const uint16_t lo = to_half(x - from_half(hi));
        out[l * 32 + s * 16 + i] = hi;
        out[l * 32 + s * 16 + 8 + i] = lo;
      }
}

// A pass's ops in the LDS tile kernel's encoding: tile-local bit positions,
// controls outside the tile in gc, diagonal factors outside the tile as
// kGlobalBit + qubit.
template <typename R>
static void append_dev_ops(const Pass& p, std::vector<DevOp<R>>& out) {
  int loc[64];
  for (int q = 0; q < 64; ++q) loc[q] = -1;
  for (size_t j = 0; j < p.qubits.size(); ++j) loc[p.qubits[j]] = (int)j;
  auto pos = [&](int q) { return loc[q] >= 0 ? loc[q] : kGlobalBit + q; };
  for (const HostOp& o : p.ops) {
    DevOp<R> d{};
    d.kind = o.kind;
    d.a = pos(o.a);
    d.b = (o.kind == OP_U2 || o.kind == OP_SWAP || o.kind == OP_D2) ? pos(o.b) : 0;
    for (int q = 0; q < 64; ++q)
      if ((o.cmask >> q) & 1) {
        if (loc[q] >= 0) d.cmask |= bit(loc[q]);
        else d.gc |= bit(q);
      }
    const int nm = (o.kind == OP_U2) ? 16 : (o.kind == OP_U1 ? 4 : (o.kind == OP_D2 ? 4 : 2));
    for (int i = 0; i < nm; ++i) {
      d.m[2 * i] = (R)o.m[i].real();
      d.m[2 * i + 1] = (R)o.m[i].imag();
    }
    out.push_back(d);
  }
}

// Host emulation of one k_tile / k_tile_small pass: per tile, the ops in the
// kernel's encoding applied to the gathered tile.
static void emulate_lds_pass(std::vector<cd>& a, int n, const Pass& p, const std::vector<DevOp<double>>& dops) {
  const int K = (int)p.qubits.size();
  const uint32_t TT = 1u << K;
  const uint64_t ntiles = 1ull << (n - K);
  std::vector<cd> t(TT);
  for (uint64_t tile = 0; tile < ntiles; ++tile) {
    uint64_t base = tile;
    for (int j = 0; j < K; ++j) {
      const int q = p.qubits[j];
      base = ((base >> q) << (q + 1)) | (base & ((1ull << q) - 1));
    }
    auto dep = [&](uint32_t e) {
      uint64_t d = base;
      for (int j = 0; j < K; ++j)
        if ((e >> j) & 1) d |= 1ull << p.qubits[j];
      return d;
    };
    for (uint32_t e = 0; e < TT; ++e) t[e] = a[dep(e)];
    for (const DevOp<double>& op : dops) {
      if ((base & op.gc) != op.gc) continue;
      const uint32_t cm = (uint32_t)op.cmask;
      auto bitof = [&](int x, uint32_t e) -> int {
        return x >= kGlobalBit ? (int)((base >> (x - kGlobalBit)) & 1) : (int)((e >> x) & 1);
      };
      cd m[16];
      for (int i = 0; i < 16; ++i) m[i] = cd(op.m[2 * i], op.m[2 * i + 1]);
      for (uint32_t e = 0; e < TT; ++e) {
        if ((e & cm) != cm) continue;
        switch (op.kind) {
          case OP_U1:
            if (!((e >> op.a) & 1)) {
              const uint32_t f = e | (1u << op.a);
              cd x = t[e], y = t[f];
              t[e] = m[0] * x + m[1] * y;
              t[f] = m[2] * x + m[3] * y;
            }
            break;
          case OP_FLIP:
            if (!((e >> op.a) & 1)) std::swap(t[e], t[e | (1u << op.a)]);
            break;
          case OP_D1: t[e] *= m[bitof(op.a, e)]; break;
          case OP_D2: t[e] *= m[bitof(op.a, e) | (bitof(op.b, e) << 1)]; break;
          case OP_U2:
            if (!((e >> op.a) & 1) && !((e >> op.b) & 1)) {
              const uint32_t idx[4] = {e, e | (1u << op.a), e | (1u << op.b), e | (1u << op.a) | (1u << op.b)};
              cd x[4] = {t[idx[0]], t[idx[1]], t[idx[2]], t[idx[3]]};
              for (int r = 0; r < 4; ++r)
                t[idx[r]] = m[4 * r] * x[0] + m[4 * r + 1] * x[1] + m[4 * r + 2] * x[2] + m[4 * r + 3] * x[3];
            }
            break;
          case OP_SWAP:
            if (((e >> op.a) & 1) && !((e >> op.b) & 1)) std::swap(t[e], t[(e ^ (1u << op.a)) | (1u << op.b)]);
            break;
          default: break;
        }
      }
    }
    for (uint32_t e = 0; e < TT; ++e) a[dep(e)] = t[e];
  }
}

// Host emulation of one k_tile_reg pass over a complex128 vector, following
// the kernel's thread / register / LDS semantics step by step.
static void emulate_reg_pass(std::vector<cd>& a, int n, const Pass& p, const RegProgram<double>& rp, int eb,
                             int tb = 8) {
  const std::vector<RegOp<double>>& prog = rp.ops;
  const int K = tb + eb, E = 1 << eb, NTH = 1 << tb;
  const uint64_t ntiles = 1ull << (n - K);
  std::vector<cd> v(256 * E), lds(size_t(1) << K);
  auto dep = [&](uint32_t t) {
    uint64_t d = 0;
    for (int j = 0; j < K; ++j)
      if ((t >> j) & 1) d |= 1ull << p.qubits[j];
    return d;
  };
  auto dep_out = [&](uint32_t t) {  // store positions (the pass may permute them)
    uint64_t d = 0;
    for (int j = 0; j < K; ++j)
      if ((t >> j) & 1) d |= 1ull << (p.out.empty() ? p.qubits[j] : p.out[j]);
    return d;
  };
  auto tmap = [&](const uint8_t* m, int tid, int r) {
    uint32_t t = 0;
    for (int i = 0; i < 8; ++i) t |= (uint32_t)((tid >> i) & 1) << m[i];
    for (int j = 0; j < eb; ++j) t |= (uint32_t)((r >> j) & 1) << m[8 + j];
    return t;
  };
  for (uint64_t tile = 0; tile < ntiles; ++tile) {
    uint64_t base = tile;
    for (int j = 0; j < K; ++j) {
      const int q = p.qubits[j];
      base = ((base >> q) << (q + 1)) | (base & ((1ull << q) - 1));
    }
    for (int tid = 0; tid < NTH; ++tid)
      for (int r = 0; r < E; ++r) v[tid * E + r] = a[base | dep(tmap(rp.map0, tid, r))];
    for (const RegOp<double>& op : prog) {
      if (op.kind == R_XCHG) {
        std::vector<cd> nv(v);
        for (int tid = 0; tid < NTH; ++tid)
          for (int r = 0; r < E; ++r) {
            const int y = (r >> op.s0) & 1;
            if (op.s1 >= 16) {  // register slots s0 and s1 - 16 trade bits
              const int k = op.s1 - 16, z = (r >> k) & 1;
              const int r2 = (r & ~(1 << op.s0) & ~(1 << k)) | (z << op.s0) | (y << k);
              nv[tid * E + r] = v[tid * E + r2];
            } else {  // register slot s0 and lane bit s1 trade bits
              const int z = (tid >> op.s1) & 1;
              const int tid2 = (tid & ~(1 << op.s1)) | (y << op.s1);
              const int r2 = (r & ~(1 << op.s0)) | (z << op.s0);
              nv[tid * E + r] = v[tid2 * E + r2];
            }
          }
        v.swap(nv);
        continue;
      }
      if (op.kind == R_REMAP) {
        for (int tid = 0; tid < NTH; ++tid)
          for (int r = 0; r < E; ++r) lds[tmap(op.map_old, tid, r)] = v[tid * E + r];
        for (int tid = 0; tid < NTH; ++tid)
          for (int r = 0; r < E; ++r) v[tid * E + r] = lds[tmap(op.map_new, tid, r)];
        continue;
      }
      if (op.kind == R_MFMA) {
        // per wavefront, column (lane bits 0..4) and register half e: the 16
        // block amplitudes are a0, a1, a3 = register bits 0, 1, 2 and a2 =
        // lane bit 5 on input; a0, a2, a3 = register bits 0, 1, 2 and a1 =
        // lane bit 5 on output
        int var = 0;  // variant chosen by the qubits outside the tile the block depends on
        if (op.pad > 0) var |= (int)((base >> (op.s1 & 255)) & 1);
        if (op.pad > 1) var |= (int)((base >> (op.s1 >> 8)) & 1) << 1;
        const std::array<cd, 256>& M = rp.blocks[op.s0 + var];
        std::vector<cd> nv(v);
        for (int w = 0; w < NTH; w += 64)
          for (int col = 0; col < 32; ++col)
            for (int e = 0; e < (E >> 3); ++e) {
              cd x[16];
              for (int ai = 0; ai < 16; ++ai) {
                const int tid = w + col + 32 * ((ai >> 2) & 1);
                const int r = (ai & 1) + 2 * ((ai >> 1) & 1) + 4 * ((ai >> 3) & 1) + 8 * e;
                x[ai] = v[tid * E + r];
              }
              for (int ao = 0; ao < 16; ++ao) {
                cd y = 0;
                for (int ai = 0; ai < 16; ++ai) y += M[16 * ao + ai] * x[ai];
                const int tid = w + col + 32 * ((ao >> 1) & 1);
                const int r = (ao & 1) + 2 * ((ao >> 2) & 1) + 4 * ((ao >> 3) & 1) + 8 * e;
                nv[tid * E + r] = y;
              }
            }
        v.swap(nv);
        continue;
      }
      if (op.kind == R_QPHASE) {
        const double* tab = rp.qtab.data() + (size_t)op.s0 * (kMfmaBlobBytes / sizeof(float));
        const double* ch = tab + 256 * (1 + eb);
        for (int tid = 0; tid < NTH; ++tid) {
          double c = tab[tid], l[8];
          for (int j = 0; j < eb; ++j) l[j] = tab[256 * (1 + j) + tid];
          for (int k = 0; k < op.pad; ++k) {
            uint32_t pat = 0;
            int i = 0;
            for (uint64_t m = qphase_chunk(op, k); m; m &= m - 1, ++i)
              pat |= (uint32_t)((base >> __builtin_ctzll(m)) & 1) << i;
            const double* e = ch + (size_t)k * kQPhaseChunkFloats + 16 * pat;
            c += e[0];
            for (int t = 0; t < 8; ++t)
              if ((tid >> t) & 1) c += e[1 + t];
            for (int j = 0; j < eb; ++j) l[j] += e[9 + j];
          }
          for (int r = 0; r < E; ++r) {
            double phi = c;
            for (int j = 0; j < eb; ++j) {
              if (!((r >> j) & 1)) continue;
              phi += l[j];
              for (int k = j + 1; k < eb; ++k)
                if ((r >> k) & 1) phi += op.m[k * (k - 1) / 2 + j];
            }
            v[tid * E + r] *= std::polar(1.0, 2 * M_PI * phi);
          }
        }
        continue;
      }
      if ((base & op.gc) != op.gc) continue;  // a control outside the tile is 0 here
      cd m[16];
      for (int i = 0; i < 16; ++i) m[i] = cd(op.m[2 * i], op.m[2 * i + 1]);
      for (int tid = 0; tid < NTH; ++tid) {
        if ((tid & op.ct) != op.ct) continue;
        cd* w = &v[tid * E];
        for (int r = 0; r < E; ++r) {
          if ((r & op.cr) != (int)op.cr) continue;
          const int j0 = op.s0, j1 = op.s1;
          switch (op.kind) {
            case R_U1:
              if (!((r >> j0) & 1)) {
                cd x = w[r], y = w[r | (1 << j0)];
                w[r] = m[0] * x + m[1] * y;
                w[r | (1 << j0)] = m[2] * x + m[3] * y;
              }
              break;
            case R_FLIP:
              if (!((r >> j0) & 1)) std::swap(w[r], w[r | (1 << j0)]);
              break;
            case R_D1:
            case R_D2: {
              auto bitof = [&](int code) -> int {
                if (code >= kGlobalBit) return (int)((base >> (code - kGlobalBit)) & 1);
                return code >= 16 ? (r >> (code - 16)) & 1 : (tid >> code) & 1;
              };
              int s = bitof(j0) | (op.kind == R_D2 ? bitof(j1) << 1 : 0);
              w[r] *= m[s];
              break;
            }
            case R_U2:
              if (!((r >> j0) & 1) && !((r >> j1) & 1)) {
                int idx[4] = {r, r | (1 << j0), r | (1 << j1), r | (1 << j0) | (1 << j1)};
                cd x[4] = {w[idx[0]], w[idx[1]], w[idx[2]], w[idx[3]]};
                for (int row = 0; row < 4; ++row) {
                  cd acc = 0;
                  for (int c = 0; c < 4; ++c) acc += m[4 * row + c] * x[c];
                  w[idx[row]] = acc;
                }
              }
              break;
            case R_SWAP:
              if (((r >> j0) & 1) && !((r >> j1) & 1)) std::swap(w[r], w[r ^ (1 << j0) ^ (1 << j1)]);
              break;
            default: break;
          }
        }
      }
    }
    for (int tid = 0; tid < NTH; ++tid)
      for (int r = 0; r < E; ++r) a[base | dep_out(tmap(rp.map1, tid, r))] = v[tid * E + r];
  }
}

// Apply one op (global numbering) to a complex128 vector on the host.
static void emulate_op(std::vector<cd>& a, int n, const HostOp& o) {
  const uint64_t dim = 1ull << n;
  for (uint64_t i = 0; i < dim; ++i) {
    if ((i & o.cmask) != o.cmask) continue;
    switch (o.kind) {
      case OP_U1:
        if (!((i >> o.a) & 1)) {
          cd x = a[i], y = a[i | bit(o.a)];
          a[i] = o.m[0] * x + o.m[1] * y;
          a[i | bit(o.a)] = o.m[2] * x + o.m[3] * y;
        }
        break;
      case OP_FLIP:
        if (!((i >> o.a) & 1)) std::swap(a[i], a[i | bit(o.a)]);
        break;
      case OP_D1: a[i] *= o.m[(i >> o.a) & 1]; break;
      case OP_D2: a[i] *= o.m[((i >> o.a) & 1) | (((i >> o.b) & 1) << 1)]; break;
      case OP_U2:
        if (!((i >> o.a) & 1) && !((i >> o.b) & 1)) {
          uint64_t idx[4] = {i, i | bit(o.a), i | bit(o.b), i | bit(o.a) | bit(o.b)};
          cd x[4] = {a[idx[0]], a[idx[1]], a[idx[2]], a[idx[3]]};
          for (int r = 0; r < 4; ++r) a[idx[r]] = o.m[4 * r] * x[0] + o.m[4 * r + 1] * x[1] + o.m[4 * r + 2] * x[2] +
                                                  o.m[4 * r + 3] * x[3];
        }
        break;
      case OP_SWAP:
        if (((i >> o.a) & 1) && !((i >> o.b) & 1)) std::swap(a[i], a[(i ^ bit(o.a)) | bit(o.b)]);
        break;
      default: break;
    }
  }
}

NqpuError emulate_tiled(const Circuit& c, const NqpuRunOptions& o, int prec, std::vector<cd>& a) {
  const bool ref = o.semantics == NQPU_SEMANTICS_REFERENCE;
  std::vector<HostOp> ops = merge_single_qubit(lower_circuit(c, ref));
  if (o.flags & NQPU_RUN_BATCH_LAYOUT) {  // the layout k_batch gives this circuit
    constexpr int EB = kBatchEB, KH = 8 + kBatchEB;
    if (c.n < EB) return NQPU_INVALID_PARAMETER;
    if (c.n <= KH) {
      Pass p;
      for (int q = 0; q < c.n; ++q) p.qubits.push_back(q);
      p.ops = ops;
      emulate_reg_pass(a, c.n, p, build_reg_program<double>(p, EB, c.n - EB), EB, c.n - EB);
    } else {
      PlanOptions bpo;
      bpo.tile_qubits = KH;
      bpo.low_qubits = 5;
      for (const Pass& p : plan_tiles(ops, c.n, bpo)) emulate_reg_pass(a, c.n, p, build_reg_program<double>(p, EB, 8, true, false, 1 << 20, true), EB);
    }
    return NQPU_OK;
  }
  // as a run does: the plan may leave the qubits in another layout, which
  // the state's next read by index restores (settle_state's passes)
  PlanOptions po = plan_options(o, prec, c.n);
  po.keep_perm = true;
  std::vector<int> fpi;
  std::vector<Pass> passes = plan_tiles(ops, c.n, po, &fpi);
  const int K = std::min(po.tile_qubits, c.n);
  const bool reg = po.registers && K >= 11 && K <= 13;
  const bool c64 = prec == NQPU_PRECISION_C64;
  for (const Pass& p : passes) {
    if (reg) {
      emulate_reg_pass(a, c.n, p,
                       build_reg_program<double>(p, K - 8, 8, po.mfma && c64, c64 || po.compile, 1 << 20,
                                                 po.qphase && c64),
                       K - 8);
    } else {
      std::vector<DevOp<double>> dops;
      append_dev_ops<double>(p, dops);
      emulate_lds_pass(a, c.n, p, dops);
    }
  }
  if (!fpi.empty()) {
    std::vector<Pass> close;
    if (!closing_passes(fpi, c.n, K, 5, &close)) {
      set_error("the state's qubit layout cannot be restored");
      return NQPU_SIMULATION_FAILED;
    }
    for (const Pass& p : close) emulate_reg_pass(a, c.n, p, build_reg_program<double>(p, K - 8, 8, false, c64), K - 8);
  }
  return NQPU_OK;
}

// Register tile passes of a plan: programs built on parallel host threads,
// then laid out as one device image (programs, then each pass's matrix-core
// operand blobs and phase tables) with one launch per pass.  Matrix-core
// blocks, lane exchanges and quadratic phases (float tables) exist for
// complex64 only; complex128 programs use LDS remaps and vector-ALU gates.
template <typename R>
static NqpuError prepare_reg(const std::vector<Pass>& passes, int n, int eb, bool mfma, bool compile, bool qphase,
                             int tiles_per_group, PreparedRun* out) {
  std::vector<RegOp<R>> prog;
  std::vector<uint16_t> blob;  // R_MFMA operand blobs of the whole run
  std::vector<PassArgs> args;
  std::vector<size_t> blob_at;  // byte offset of each pass's operand blobs
  constexpr bool c64 = std::is_same<R, float>::value;
  // lane exchanges: complex64, and complex128 compiled passes (the
  // interpreting kernel keeps complex128 exchanges out)
  const bool xchg = c64 || compile;
  std::vector<RegProgram<R>> progs(passes.size());
  {
    const size_t nt = std::min<size_t>(passes.size(), std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::atomic<size_t> next{0};
    auto work = [&] {
      for (size_t i; (i = next.fetch_add(1)) < passes.size();)
        progs[i] = build_reg_program<R>(passes[i], eb, 8, mfma && c64, xchg, 1 << 20, qphase && c64);
    };
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
  }
  NqpuRunStats& st = out->stats;
  for (const Pass& p : passes) {
    PassArgs pa{};
    pa.K = (int)p.qubits.size();
    pa.n = n;
    pa.tiles_per_group = tiles_per_group;
    pa.op0 = (int)prog.size();
    for (int j = 0; j < pa.K; ++j) pa.q[j] = p.qubits[j];
    RegProgram<R>& pr = progs[&p - passes.data()];
    for (RegOp<R>& r : pr.ops)  // R_QPHASE tables follow the blocks' operand blobs
      if (r.kind == R_QPHASE) r.s0 += (int32_t)pr.blocks.size();
    for (int slot = 0; slot < 8 + eb; ++slot) {  // thread slots 0..7, register slots 8..
      pa.ql[slot] = p.qubits[pr.map0[slot]];
      pa.qf[slot] = (p.out.empty() ? p.qubits : p.out)[pr.map1[slot]];
    }
    blob_at.push_back(blob.size() * 2);
    bool has_remap = false;
    for (RegOp<R>& r : pr.ops) {
      st.remaps += r.kind == R_REMAP;
      st.exchanges += r.kind == R_XCHG;
      has_remap = has_remap || r.kind == R_REMAP;
    }
    pa.lds_remap = has_remap ? (int)(sizeof(R) * 2) << pa.K : 0;
    // a store permutation that moves the wave slots' positions makes
    // wavefronts overwrite amplitudes other wavefronts load; a remap's
    // barrier already orders all loads before all stores
    pa.sync_store = !has_remap && (pa.qf[6] != pa.ql[6] || pa.qf[7] != pa.ql[7]);
    for (const auto& M : pr.blocks) {
      const size_t at = blob.size();
      blob.resize(at + kMfmaBlobBytes / 2);
      mfma_blob(M, blob.data() + at);
    }
    if (!pr.qtab.empty()) {
      const size_t at = blob.size();
      blob.resize(at + pr.qtab.size() * sizeof(R) / 2);
      std::memcpy(blob.data() + at, pr.qtab.data(), pr.qtab.size() * sizeof(R));
    }
    st.qphase_gates += pr.qphase_gates;
    st.mfma_blocks += pr.blocks.size();
    st.mfma_gates += pr.fused_ops;
    pa.scaled = pr.blocks.empty() ? 0 : 1;
    pa.nops = (int)pr.ops.size();
    prog.insert(prog.end(), pr.ops.begin(), pr.ops.end());
    args.push_back(pa);
  }
  const size_t pb = (std::max<size_t>(prog.size(), 1) * sizeof(RegOp<R>) + 255) & ~size_t(255);
  std::vector<hipFunction_t> fns;
  // compiled passes: complex64 12- and 13-qubit tiles (matrix-core blocks
  // need EB = 4 or 5) and complex128 register tiles of any width
  if (compile && (std::is_same<R, double>::value || eb == 4 || eb == 5)) {
    const NqpuError e = jit_pass_functions<R>(prog, args, eb, &fns, &out->hold);
    if (e != NQPU_OK) return e;
  }
  out->prog_bytes = pb;
  out->image.assign(pb + blob.size() * 2, 0);
  if (!prog.empty()) std::memcpy(out->image.data(), prog.data(), prog.size() * sizeof(RegOp<R>));
  if (!blob.empty()) std::memcpy(out->image.data() + pb, blob.data(), blob.size() * 2);
  for (size_t i = 0; i < args.size(); ++i) {
    RunLaunch l;
    l.kind = fns.empty() ? RunLaunch::REG : RunLaunch::COMPILED;
    l.pa = args[i];
    l.eb = eb;
    l.blob_at = pb + blob_at[i];
    l.fn = fns.empty() ? nullptr : fns[i];
    out->launches.push_back(l);
  }
  return NQPU_OK;
}

// LDS tile passes of a plan (tiles outside 11..13 qubits, or NQPU_RUN_LDS_TILES).
template <typename R>
static void prepare_lds(const std::vector<Pass>& passes, int n, PreparedRun* out) {
  std::vector<DevOp<R>> dops;
  for (const Pass& p : passes) {
    RunLaunch l;
    l.kind = RunLaunch::LDS;
    l.pa.K = (int)p.qubits.size();
    l.pa.n = n;
    l.pa.op0 = (int)dops.size();
    l.pa.nops = (int)p.ops.size();
    for (int j = 0; j < l.pa.K; ++j) l.pa.q[j] = p.qubits[j];
    append_dev_ops<R>(p, dops);
    out->launches.push_back(l);
  }
  out->prog_bytes = std::max<size_t>(dops.size(), 1) * sizeof(DevOp<R>);
  out->image.assign(out->prog_bytes, 0);
  if (!dops.empty()) std::memcpy(out->image.data(), dops.data(), dops.size() * sizeof(DevOp<R>));
  out->stats.device_ops += dops.size();
}

template <typename R>
static NqpuError prepare_tiled_ops(const std::vector<HostOp>& ops, int n, const PlanOptions& po, PreparedRun* out) {
  *out = PreparedRun{};
  out->prec = std::is_same<R, float>::value ? NQPU_PRECISION_C64 : NQPU_PRECISION_C128;
  out->n = n;
  std::vector<Pass> passes = plan_tiles(ops, n, po, &out->perm);
  const int K = std::min(po.tile_qubits, n);
  out->perm_k = K;
  NqpuError e = NQPU_OK;
  if (po.registers && K >= 11 && K <= 13) {
    out->stats.device_ops += ops.size();
    e = prepare_reg<R>(passes, n, K - 8, po.mfma, po.compile, po.qphase, po.tiles_per_group, out);
  } else {
    prepare_lds<R>(passes, n, out);
  }
  out->stats.passes += passes.size();
  out->stats.bytes_moved += (double)passes.size() * 2.0 * (double)(sizeof(R) * 2 << n);
  return e;
}

NqpuError prepare_tiled(const Circuit& c, const NqpuRunOptions& o, int prec, int n, const PlanOptions& po,
                        PreparedRun* out) {
  if (o.fusion != NQPU_FUSION_TILED) {
    set_error("only tiled runs (NQPU_FUSION_TILED) are prepared ahead of their launch");
    return NQPU_INVALID_PARAMETER;
  }
  const bool ref = o.semantics == NQPU_SEMANTICS_REFERENCE;
  std::vector<HostOp> ops = merge_single_qubit(lower_circuit(c, ref));
  NqpuError e = prec == NQPU_PRECISION_C128 ? prepare_tiled_ops<double>(ops, n, po, out)
                                            : prepare_tiled_ops<float>(ops, n, po, out);
  out->stats.gates = c.gates.size();
  return e;
}

template <typename R>
static NqpuError launch_prepared_t(const PreparedRun& r, void* amps, const void* image, hipStream_t st,
                                   PassTimer* timer) {
  for (const RunLaunch& l : r.launches) {
    if (timer) timer->mark(st);
    if (l.kind == RunLaunch::COMPILED) {
      void* a0 = amps;
      const void* pg = image;
      const void* bl = (const char*)image + l.blob_at;
      uint64_t ntiles = 1ull << (l.pa.n - l.pa.K);
      PassArgs a = l.pa;
      void* kargs[] = {&a0, &a, &ntiles, &pg, &bl};
      const uint64_t tpg = (uint64_t)std::max(1, l.pa.tiles_per_group);
      const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(ntiles / tpg, 1ull << 20));
      HIPTRY(hipModuleLaunchKernel(l.fn, grid, 1, 1, 256, 1, 1, (unsigned)l.pa.lds_remap, st, kargs, nullptr),
             "compiled register tile pass launch");
    } else if (l.kind == RunLaunch::REG) {
      Launch<R>::tile_pass_reg(amps, l.pa, l.eb, image, (const char*)image + l.blob_at, st);
    } else {
      Launch<R>::tile_pass(amps, l.pa, image, st);
    }
    if (timer) timer->mark(st);
    HIPTRY(hipGetLastError(), "tile pass launch");
  }
  return NQPU_OK;
}

NqpuError launch_prepared(const PreparedRun& r, void* amps, const void* image, hipStream_t st, PassTimer* timer) {
  return r.prec == NQPU_PRECISION_C128 ? launch_prepared_t<double>(r, amps, image, st, timer)
                                       : launch_prepared_t<float>(r, amps, image, st, timer);
}

// Upload a prepared run's image into the state's program buffer: through the
// pinned stage with one asynchronous copy in stream order (the previous run's
// kernels have read the device buffer before it lands); the stage is reused
// once the previous copy out of it is done.  Growing either buffer waits for
// the device first.
static NqpuError upload_image(State* s, const std::vector<char>& image) {
  const size_t ob = image.size();
  if (s->stage_pending) {
    HIPTRY(hipEventSynchronize(s->stage_ev), "wait for the previous program upload");
    s->stage_pending = false;
  }
  if (ob > s->opbuf.bytes || ob > s->stage_bytes) HIPTRY(hipStreamSynchronize(s->stream), "sync before buffer growth");
  if (!s->opbuf.ensure(ob)) return NQPU_OUT_OF_DEVICE_MEMORY;
  if (ob > s->stage_bytes) {
    if (s->stage) (void)hipHostFree(s->stage);
    s->stage = nullptr;
    s->stage_bytes = 0;
    const size_t nb = std::max(ob, (size_t)1 << 20);
    HIPTRY(hipHostMalloc(&s->stage, nb), "pinned program stage");
    s->stage_bytes = nb;
  }
  if (!s->stage_ev) HIPTRY(hipEventCreateWithFlags(&s->stage_ev, hipEventDisableTiming), "stage event");
  std::memcpy(s->stage, image.data(), ob);
  HIPTRY(hipMemcpyAsync(s->opbuf.p, s->stage, ob, hipMemcpyHostToDevice, s->stream), "program upload");
  HIPTRY(hipEventRecord(s->stage_ev, s->stream), "stage event record");
  s->stage_pending = true;
  return NQPU_OK;
}

static void add_stats(NqpuRunStats& a, const NqpuRunStats& b) {
  a.passes += b.passes;
  a.device_ops += b.device_ops;
  a.bytes_moved += b.bytes_moved;
  a.remaps += b.remaps;
  a.exchanges += b.exchanges;
  a.mfma_blocks += b.mfma_blocks;
  a.mfma_gates += b.mfma_gates;
  a.qphase_gates += b.qphase_gates;
}

// A run from |0...0>: its first register pass generates the start state
// instead of loading it (no read, and no separate fill pass before it).
// False when the run's first launch cannot (LDS tile passes, or no passes).
bool start_from_zero(PreparedRun* r) {
  if (r->launches.empty() || r->launches[0].kind == RunLaunch::LDS) return false;
  r->launches[0].pa.from_zero = 1;
  return true;
}

template <typename R>
static NqpuError run_tiled(State* s, const std::vector<HostOp>& ops, const PlanOptions& po, bool from_zero) {
  PreparedRun r;
  NqpuError e = prepare_tiled_ops<R>(ops, s->n, po, &r);
  if (e != NQPU_OK) return e;
  if (from_zero && !start_from_zero(&r)) Launch<R>::init_zero(s->amps, s->n, s->stream);
  if ((e = upload_image(s, r.image)) != NQPU_OK) return e;
  if ((e = launch_prepared_t<R>(r, s->amps, s->opbuf.p, s->stream, &s->timer)) != NQPU_OK) return e;
  s->perm = r.perm;
  s->perm_k = r.perm_k;
  // the run's compiled modules stay loaded until the stream is synchronised
  s->jit_hold.insert(s->jit_hold.end(), r.hold.begin(), r.hold.end());
  add_stats(s->stats, r.stats);
  return NQPU_OK;
}

NqpuError settle_state(const State* s0) {
  if (s0->perm.empty()) return NQPU_OK;
  // the amplitudes move, the state they represent does not
  State* s = const_cast<State*>(s0);
  std::vector<Pass> passes;
  if (!closing_passes(s->perm, s->n, s->perm_k, 5, &passes)) {
    set_error("the state's qubit layout cannot be restored");
    return NQPU_SIMULATION_FAILED;
  }
  PreparedRun r;
  r.prec = s->prec;
  r.n = s->n;
  const int eb = s->perm_k - 8;
  NqpuError e = s->prec == NQPU_PRECISION_C64 ? prepare_reg<float>(passes, s->n, eb, false, false, false, 1, &r)
                                              : prepare_reg<double>(passes, s->n, eb, false, false, false, 1, &r);
  if (e != NQPU_OK) return e;
  if ((e = upload_image(s, r.image)) != NQPU_OK) return e;
  if ((e = launch_prepared(r, s->amps, s->opbuf.p, s->stream, nullptr)) != NQPU_OK) return e;
  s->perm.clear();
  return NQPU_OK;
}

template <typename R>
static NqpuError jit_check_circuit_t(const Circuit& c, const NqpuRunOptions& o) {
  constexpr bool c64 = std::is_same<R, float>::value;
  const int prec = c64 ? NQPU_PRECISION_C64 : NQPU_PRECISION_C128;
  PlanOptions po = plan_options(o, prec, c.n);
  po.keep_perm = true;  // the passes a run compiles
  const int K = std::min(po.tile_qubits, c.n);
  if (!po.registers || (c64 ? (K != 12 && K != 13) : (K < 11 || K > 13))) {
    set_error(c64 ? "compiled passes need complex64 register tiles of 12 or 13 qubits"
                  : "compiled passes need complex128 register tiles of 11-13 qubits");
    return NQPU_INVALID_PARAMETER;
  }
  const std::vector<HostOp> ops = merge_single_qubit(lower_circuit(c, o.semantics == NQPU_SEMANTICS_REFERENCE));
  std::vector<RegOp<R>> prog;
  std::vector<PassArgs> args;
  for (const Pass& p : plan_tiles(ops, c.n, po)) {
    const RegProgram<R> pr = build_reg_program<R>(p, K - 8, 8, po.mfma && c64, true, 1 << 20, po.qphase && c64);
    PassArgs pa{};
    pa.op0 = (int)prog.size();
    pa.nops = (int)pr.ops.size();
    bool has_remap = false;
    for (const RegOp<R>& r : pr.ops) has_remap = has_remap || r.kind == R_REMAP;
    pa.lds_remap = has_remap ? (int)(2 * sizeof(R)) << K : 0;
    prog.insert(prog.end(), pr.ops.begin(), pr.ops.end());
    args.push_back(pa);
  }
  return jit_check_compile<R>(prog, args, K - 8);
}

NqpuError jit_check_circuit(const Circuit& c, const NqpuRunOptions& o, int prec) {
  return prec == NQPU_PRECISION_C128 ? jit_check_circuit_t<double>(c, o) : jit_check_circuit_t<float>(c, o);
}

PlanOptions plan_options(const NqpuRunOptions& o, int prec, int n) {
  PlanOptions po;
  const int kmax = (prec == NQPU_PRECISION_C64) ? 13 : 12;
  // Default tiles: 12 qubits (complex64) and 11 (complex128) for large
  // states.  A state of at most 22 (complex64) / 20 (complex128) qubits has
  // too few 12- / 11-qubit tiles to fill 256 CUs, and its passes are latency
  // bound: one qubit less doubles the workgroups (random 20-qubit circuit of
  // depth 40 on MI355X: 2.1 instead of 3.1 ms complex64, 1.7 instead of
  // 2.0 ms complex128).  Compiled complex64 passes keep 12-qubit tiles, and
  // take 13-qubit tiles from 26 qubits on: fewer, heavier passes (30-qubit
  // random brickwork of depth 40: 95.6 instead of 101.6 ms; 34-qubit QFT +
  // 20 Trotter steps: 2.33 instead of 2.53 s).
  const bool c64 = prec == NQPU_PRECISION_C64, small = n > 0 && n <= (c64 ? 22 : 20);
  const bool compiled = (o.flags & NQPU_RUN_COMPILE) != 0;
  const int dflt = c64 ? ((small && !compiled) ? 11 : (compiled && n >= 26 ? 13 : 12)) : (small ? 10 : 11);
  po.tile_qubits = o.tile_qubits ? std::min<int>((int)o.tile_qubits, kmax) : dflt;
  if (po.tile_qubits < 5) po.tile_qubits = 5;
  po.low_qubits = o.low_qubits ? (int)o.low_qubits : (prec == NQPU_PRECISION_C64 ? 5 : 4);
  po.low_qubits = std::min(po.low_qubits, po.tile_qubits);
  po.registers = (o.flags & NQPU_RUN_LDS_TILES) == 0;
  po.mfma = (o.flags & NQPU_RUN_NO_MFMA) == 0;
  po.compile = (o.flags & NQPU_RUN_COMPILE) != 0;
  po.qphase = (o.flags & NQPU_RUN_NO_QPHASE) == 0;
  po.move_low = po.registers && prec == NQPU_PRECISION_C64 && (o.flags & NQPU_RUN_FIXED_LOW) == 0;
  // a plan measured by the autotuner for (n, precision) replaces the defaults
  // of a caller that chose none of these settings
  const uint32_t chosen = NQPU_RUN_UNTUNED | NQPU_RUN_LDS_TILES | NQPU_RUN_NO_MFMA | NQPU_RUN_FIXED_LOW |
                          NQPU_RUN_COMPILE;
  TunedPlan t;
  if (n > 0 && !(o.flags & chosen) && !o.tile_qubits && !o.low_qubits && tuned_lookup(n, prec, &t)) {
    po.tile_qubits = std::min(t.tile_qubits, kmax);
    po.low_qubits = std::min(t.low_qubits, po.tile_qubits);
    po.compile = (t.flags & NQPU_RUN_COMPILE) != 0;
    po.move_low = po.registers && prec == NQPU_PRECISION_C64 && (t.flags & NQPU_RUN_FIXED_LOW) == 0;
    po.tiles_per_group = std::max(1, t.tiles_per_group);
  }
  return po;
}

NqpuRunStats plan_stats(const Circuit& c, const NqpuRunOptions& o, int prec) {
  NqpuRunStats st{};
  st.gates = c.gates.size();
  const bool ref = o.semantics == NQPU_SEMANTICS_REFERENCE;
  const size_t asz = prec == NQPU_PRECISION_C64 ? 8 : 16;
  auto add_direct = [&](const HostOp& op) {
    st.device_ops++;
    st.passes++;
    st.bytes_moved += (double)op_bytes_direct(op, c.n, asz);
  };
  if (o.fusion == NQPU_FUSION_NONE) {
    for (const Gate& g : c.gates)
      for (const HostOp& op : lower_gate(g, ref)) add_direct(op);
  } else if (o.fusion == NQPU_FUSION_REFERENCE) {
    FusionResult fr = fuse_gates_reference(c.gates);
    for (const FusedItem& it : fr.items) {
      if (it.kind == FusedItem::FUSED || it.kind == FusedItem::FUSED_TWO) {
        HostOp h; h.kind = it.kind == FusedItem::FUSED ? OP_U1 : OP_U2; h.a = it.kind == FusedItem::FUSED ? it.q : it.qlo; h.b = it.qhi;
        add_direct(h);
      } else if (it.kind == FusedItem::DIAGONAL) {
        for (const HostOp& op : diagonal_ops(it)) add_direct(op);
      } else {
        for (const HostOp& op : lower_gate(it.gate, ref)) add_direct(op);
      }
    }
  } else {
    std::vector<HostOp> ops = merge_single_qubit(lower_circuit(c, ref));
    PlanOptions po = plan_options(o, prec, c.n);
    po.keep_perm = true;  // as runs plan
    std::vector<Pass> passes = plan_tiles(ops, c.n, po);
    st.passes = passes.size();
    st.device_ops = ops.size();
    st.bytes_moved = (double)passes.size() * 2.0 * (double)(asz << c.n);
    const int K = std::min(po.tile_qubits, c.n);
    const bool c64 = prec == NQPU_PRECISION_C64;
    if (po.registers && K >= 11 && K <= 13)
      for (const Pass& p : passes) {
        const RegProgram<double> rp =
            build_reg_program<double>(p, K - 8, 8, po.mfma && c64, c64 || po.compile, 1 << 20, po.qphase && c64);
        for (const RegOp<double>& r : rp.ops) {
          st.remaps += r.kind == R_REMAP;
          st.exchanges += r.kind == R_XCHG;
        }
        st.mfma_blocks += rp.blocks.size();
        st.mfma_gates += rp.fused_ops;
        st.qphase_gates += rp.qphase_gates;
      }
  }
  return st;
}

NqpuError run_circuit(State* s, const Circuit& c, const NqpuRunOptions& o) {
  return run_circuit_with(s, c, o, plan_options(o, s->prec, s->n));
}

NqpuError run_circuit_with(State* s, const Circuit& c, const NqpuRunOptions& o, const PlanOptions& po0) {
  s->stats = NqpuRunStats{};
  s->stats.gates = c.gates.size();
  const bool ref = o.semantics == NQPU_SEMANTICS_REFERENCE;
  const bool from_zero = (o.flags & NQPU_RUN_FROM_ZERO) != 0;
  NqpuError e = NQPU_OK;
  // runs are planned on the canonical layout; |0...0> is the same in every
  // layout
  if (from_zero) s->perm.clear();
  else if ((e = settle_state(s)) != NQPU_OK) return e;
  // a tiled run leaves the qubits where its last pass put them
  PlanOptions po = po0;
  po.keep_perm = true;
  // runs queued without waiting (NQPU_RUN_ASYNC) keep adding to the pass
  // timer until the state is synchronised
  if (!s->async_open) s->timer.begin((o.flags & NQPU_RUN_TIME_PASSES) != 0);
  else if (o.flags & NQPU_RUN_TIME_PASSES) s->timer.on = true;
  dispatch(s, [&](auto* tag) {
    using R = std::remove_pointer_t<decltype(tag)>;
    if (from_zero && o.fusion != NQPU_FUSION_TILED) Launch<R>::init_zero(s->amps, s->n, s->stream);
    if (o.fusion == NQPU_FUSION_NONE) {
      for (const Gate& g : c.gates) {
        for (const HostOp& op : lower_gate(g, ref)) {
          s->stats.device_ops++;
          if ((e = exec_direct<R>(s, op)) != NQPU_OK) return;
        }
      }
    } else if (o.fusion == NQPU_FUSION_REFERENCE) {
      FusionResult fr = fuse_gates_reference(c.gates);
      for (const FusedItem& it : fr.items) {
        std::vector<HostOp> ops;
        if (it.kind == FusedItem::FUSED) ops.push_back(op_u1(it.q, 0, it.m));
        else if (it.kind == FusedItem::FUSED_TWO) {
          HostOp h; h.kind = OP_U2; h.a = it.qlo; h.b = it.qhi;
          for (int i = 0; i < 16; ++i) h.m[i] = it.m[i];
          ops.push_back(h);
        } else if (it.kind == FusedItem::DIAGONAL) ops = diagonal_ops(it);
        else ops = lower_gate(it.gate, ref);
        for (const HostOp& op : ops) {
          s->stats.device_ops++;
          if ((e = exec_direct<R>(s, op)) != NQPU_OK) return;
        }
      }
    } else {
      std::vector<HostOp> ops = merge_single_qubit(lower_circuit(c, ref));
      e = run_tiled<R>(s, ops, po, from_zero);
    }
  });
  if (e != NQPU_OK) {
    (void)hipStreamSynchronize(s->stream);
    s->jit_hold.clear();
    s->async_open = false;
    s->timer.begin(false);
    return e;
  }
  if (o.flags & NQPU_RUN_ASYNC) {
    // the passes are queued in stream order; the host may plan the next run
    // while they execute (its program upload is ordered after them too)
    s->async_open = true;
    return NQPU_OK;
  }
  return synchronize_state(s);
}

NqpuError synchronize_state(State* s) {
  HIPTRY(hipStreamSynchronize(s->stream), "run");
  s->jit_hold.clear();
  if (s->timer.on) {
    s->stats.pass_ms = s->timer.total_ms();
    s->stats.timed_passes = s->timer.used / 2;
  }
  s->timer.begin(false);
  s->async_open = false;
  return NQPU_OK;
}

void PassTimer::mark(hipStream_t s) {
  if (!on) return;
  if (used == pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) {
      on = false;
      return;
    }
    pool.push_back(e);
  }
  (void)hipEventRecord(pool[used++], s);
}

double PassTimer::total_ms() {
  double t = 0;
  for (size_t i = 0; i + 1 < used; i += 2) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, pool[i], pool[i + 1]) == hipSuccess) t += ms;
  }
  return t;
}

PassTimer::~PassTimer() {
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

NqpuError overlap(const State* a0, const State* b0, NqpuComplex* out) {
  State* a = const_cast<State*>(a0);
  State* b = const_cast<State*>(b0);
  // <a|b> does not depend on the layout when both states share it
  if (a->perm != b->perm) {
    NqpuError e = settle_state(a);
    if (e == NQPU_OK) e = settle_state(b);
    if (e != NQPU_OK) return e;
  }
  HIPTRY(hipStreamSynchronize(b->stream), "sync");
  const int cb = std::min(a->n, 16);
  const uint64_t nch = a->dim() >> cb;
  if (!a->scratch.ensure(nch * 16)) return NQPU_OUT_OF_DEVICE_MEMORY;
  launch_overlap(a->amps, a->prec == NQPU_PRECISION_C64, b->amps, b->prec == NQPU_PRECISION_C64, a->n, cb,
                 (double*)a->scratch.p, a->stream);
  HIPTRY(hipGetLastError(), "overlap");
  HIPTRY(hipStreamSynchronize(a->stream), "overlap");
  std::vector<double> rows(nch * 2);
  HIPTRY(copy_on(a->stream, rows.data(), a->scratch.p, nch * 16, hipMemcpyDeviceToHost), "overlap copy");
  double re = 0, im = 0;
  for (uint64_t c = 0; c < nch; ++c) {
    re += rows[2 * c];
    im += rows[2 * c + 1];
  }
  out->re = re;
  out->im = im;
  return NQPU_OK;
}

NqpuError apply_gate_now(State* s, const Gate& g) {
  NqpuError e = settle_state(s);
  if (e != NQPU_OK) return e;
  dispatch(s, [&](auto* tag) {
    using R = std::remove_pointer_t<decltype(tag)>;
    for (const HostOp& op : lower_gate(g, false))
      if ((e = exec_direct<R>(s, op)) != NQPU_OK) return;
  });
  if (e != NQPU_OK) return e;
  HIPTRY(hipStreamSynchronize(s->stream), "gate");
  return NQPU_OK;
}

// ---------------------------------------------------------- amplitudes I/O

static const uint64_t kXferChunk = 1ull << 24;  // amplitudes per staging copy

NqpuError get_amplitudes(const State* s0, uint64_t off, uint64_t count, NqpuComplex* out) {
  State* s = const_cast<State*>(s0);
  NqpuError e = settle_state(s);
  if (e != NQPU_OK) return e;
  HIPTRY(hipStreamSynchronize(s->stream), "sync");
  if (s->prec == NQPU_PRECISION_C128) {
    HIPTRY(copy_on(s->stream, out, (char*)s->amps + off * 16, count * 16, hipMemcpyDeviceToHost), "amplitude copy");
    return NQPU_OK;
  }
  for (uint64_t b = 0; b < count; b += kXferChunk) {
    uint64_t c = std::min(kXferChunk, count - b);
    if (!s->scratch.ensure(c * 16)) return NQPU_OUT_OF_DEVICE_MEMORY;
    Launch<float>::to_c128(s->amps, c, off + b, s->scratch.p, s->stream);
    HIPTRY(hipGetLastError(), "to_c128");
    HIPTRY(hipStreamSynchronize(s->stream), "to_c128");
    HIPTRY(copy_on(s->stream, out + b, s->scratch.p, c * 16, hipMemcpyDeviceToHost), "amplitude copy");
  }
  return NQPU_OK;
}

NqpuError set_amplitudes(State* s, uint64_t off, uint64_t count, const NqpuComplex* in) {
  NqpuError e = settle_state(s);
  if (e != NQPU_OK) return e;
  HIPTRY(hipStreamSynchronize(s->stream), "sync");
  if (s->prec == NQPU_PRECISION_C128) {
    HIPTRY(copy_on(s->stream, (char*)s->amps + off * 16, in, count * 16, hipMemcpyHostToDevice), "amplitude copy");
    return NQPU_OK;
  }
  for (uint64_t b = 0; b < count; b += kXferChunk) {
    uint64_t c = std::min(kXferChunk, count - b);
    if (!s->scratch.ensure(c * 16)) return NQPU_OUT_OF_DEVICE_MEMORY;
    HIPTRY(copy_on(s->stream, s->scratch.p, in + b, c * 16, hipMemcpyHostToDevice), "amplitude copy");
    Launch<float>::from_c128(s->amps, c, off + b, s->scratch.p, s->stream);
    HIPTRY(hipGetLastError(), "from_c128");
    HIPTRY(hipStreamSynchronize(s->stream), "from_c128");
  }
  return NQPU_OK;
}

// -------------------------------------------------------------- reductions

static int chunk_bits_for(int n) { return std::min(n, 16); }

static NqpuError chunk_sums(const State* s0, std::vector<double>* sums, int* cb_out) {
  State* s = const_cast<State*>(s0);
  const int cb = chunk_bits_for(s->n);
  const uint64_t nch = s->dim() >> cb;
  if (!s->scratch.ensure(nch * 8)) return NQPU_OUT_OF_DEVICE_MEMORY;
  dispatch(s, [&](auto* tag) {
    using R = std::remove_pointer_t<decltype(tag)>;
    Launch<R>::sumsq_chunks(s->amps, s->n, cb, (double*)s->scratch.p, s->stream);
  });
  HIPTRY(hipGetLastError(), "sumsq");
  HIPTRY(hipStreamSynchronize(s->stream), "sumsq");
  sums->resize(nch);
  HIPTRY(copy_on(s->stream, sums->data(), s->scratch.p, nch * 8, hipMemcpyDeviceToHost), "sumsq copy");
  *cb_out = cb;
  return NQPU_OK;
}

NqpuError norm(const State* s, double* out) {
  std::vector<double> sums;
  int cb;
  NqpuError e = chunk_sums(s, &sums, &cb);
  if (e != NQPU_OK) return e;
  double t = 0;
  for (double v : sums) t += v;
  *out = t;
  return NQPU_OK;
}

// tot = sum |a|^2; s1[q] = sum |a_i|^2 * bit_q(i)
NqpuError zsum_all(const State* s0, double* tot, std::vector<double>* s1) {
  State* s = const_cast<State*>(s0);
  const int n = s->n;
  s1->assign(n, 0.0);
  if (n < 8) {
    std::vector<double> p(s->dim());
    NqpuError e = probabilities(s, p.data());
    if (e != NQPU_OK) return e;
    double t = 0;
    for (uint64_t i = 0; i < s->dim(); ++i) {
      t += p[i];
      for (int q = 0; q < n; ++q)
        if ((i >> q) & 1) (*s1)[q] += p[i];
    }
    *tot = t;
    return NQPU_OK;
  }
  const int cb = chunk_bits_for(n);
  const uint64_t nch = s->dim() >> cb;
  const int w = cb + 1;
  if (!s->scratch.ensure(nch * w * 8)) return NQPU_OUT_OF_DEVICE_MEMORY;
  dispatch(s, [&](auto* tag) {
    using R = std::remove_pointer_t<decltype(tag)>;
    Launch<R>::zsum_chunks(s->amps, n, cb, (double*)s->scratch.p, s->stream);
  });
  HIPTRY(hipGetLastError(), "zsum");
  HIPTRY(hipStreamSynchronize(s->stream), "zsum");
  std::vector<double> rows(nch * w);
  HIPTRY(copy_on(s->stream, rows.data(), s->scratch.p, nch * w * 8, hipMemcpyDeviceToHost), "zsum copy");
  double t = 0;
  for (uint64_t c = 0; c < nch; ++c) {
    const double* r = &rows[c * w];
    t += r[0];
    for (int q = 0; q < cb; ++q) (*s1)[q] += r[1 + q];
    for (int q = cb; q < n; ++q)
      if ((c >> (q - cb)) & 1) (*s1)[q] += r[0];
  }
  *tot = t;
  if (!s->perm.empty()) {  // position sums -> qubit sums
    std::vector<double> by_pos(*s1);
    for (int q = 0; q < n; ++q) (*s1)[q] = by_pos[s->perm[q]];
  }
  return NQPU_OK;
}

NqpuError scale(State* s, double f) {
  dispatch(s, [&](auto* tag) {
    using R = std::remove_pointer_t<decltype(tag)>;
    Launch<R>::scale(s->amps, s->n, f, s->stream);
  });
  HIPTRY(hipGetLastError(), "scale");
  HIPTRY(hipStreamSynchronize(s->stream), "scale");
  return NQPU_OK;
}

// c_ffi.rs:473-500: p0 over bit q = 0, p1 = 1 - p0 as the reported
// probability; the outcome is drawn against p0.  The collapse divides by the
// kept half's measured mass rather than by the reported probability: in
// complex64 the norm drifts by ~1e-7, and an outcome whose measured mass is
// zero (possible only through that drift) is replaced by the other one, so
// the state never collapses to all zeros.
NqpuError measure_qubit(State* s, int q, double r, uint32_t* outcome, double* prob) {
  double tot;
  std::vector<double> s1;
  NqpuError e = zsum_all(s, &tot, &s1);
  if (e != NQPU_OK) return e;
  const double m0 = tot - s1[q], m1 = s1[q];
  const double p0 = m0;
  const double p1 = 1.0 - p0;
  uint32_t o = r < p0 ? 0 : 1;
  if ((o ? m1 : m0) <= 0.0) o ^= 1;
  const double kept = o ? m1 : m0;
  if (!(kept > 0.0)) {
    set_error("measure: the state has no mass");
    return NQPU_SIMULATION_FAILED;
  }
  e = collapse(s, q, o, 1.0 / std::sqrt(kept));
  if (e != NQPU_OK) return e;
  *outcome = o;
  *prob = o ? p1 : p0;
  return NQPU_OK;
}

NqpuError collapse(State* s, int q, uint32_t outcome, double f) {
  const int pos = s->perm.empty() ? q : s->perm[q];  // the qubit's position in the layout
  dispatch(s, [&](auto* tag) {
    using R = std::remove_pointer_t<decltype(tag)>;
    Launch<R>::collapse(s->amps, s->n, pos, outcome, f, s->stream);
  });
  HIPTRY(hipGetLastError(), "collapse");
  HIPTRY(hipStreamSynchronize(s->stream), "collapse");
  return NQPU_OK;
}

// ----------------------------------------------------------------- batches

Batch::~Batch() {
  for (hipStream_t st : {stream, copy_stream})
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  for (BatchSlot& sl : slot) {
    if (sl.pinned) (void)hipHostFree(sl.pinned);
    if (sl.results) (void)hipHostFree(sl.results);
    for (hipEvent_t e : {sl.uploaded, sl.k0, sl.k1, sl.done})
      if (e) (void)hipEventDestroy(e);
  }
}

NqpuError create_batch(Batch** out) {
  *out = nullptr;
  if (device_count() == 0) {
    set_error("no usable HIP device (or NQPU_HIP_FORCE_NO_DEVICE=1)");
    return NQPU_UNAVAILABLE;
  }
  std::unique_ptr<Batch> b(new Batch());
  HIPTRY(hipSetDevice(0), "hipSetDevice");
  HIPTRY(new_stream(&b->stream), "hipStreamCreate");
  HIPTRY(new_stream(&b->copy_stream), "hipStreamCreate");
  *out = b.release();
  return NQPU_OK;
}

// A term as the tile sees it: X/Z bits inside the tile in tile-local
// numbering, Z bits outside it in global numbering.
static BatchTerm tile_term(const PauliTerm& t, const std::vector<int>& q) {
  BatchTerm bt{};
  int loc[64];
  for (int i = 0; i < 64; ++i) loc[i] = -1;
  for (size_t j = 0; j < q.size(); ++j) loc[q[j]] = (int)j;
  for (int b = 0; b < 64; ++b) {
    if ((t.x >> b) & 1) bt.xl |= 1u << loc[b];
    if ((t.z >> b) & 1) {
      if (loc[b] >= 0) bt.zl |= 1u << loc[b];
      else bt.zg |= bit(b);
    }
  }
  bt.coeff = t.coeff;
  bt.ny = __builtin_popcountll(t.x & t.z) & 3;
  return bt;
}

namespace {
// One circuit's part of a batch plan, with pass / program / term indices
// relative to its own tables (plan_batch concatenates them).
struct CircuitPlan {
  NqpuError err = NQPU_OK;
  std::string msg;
  BatchCircuit C{};
  std::vector<BatchPass> bp;
  std::vector<RegOp<float>> prog;
  std::vector<BatchTerm> bt;
  // HBM passes' programs: their blocks (nblk, numbered in pass order) and
  // phase tables (qunits operand units of kMfmaBlobBytes, in pass order)
  std::vector<RegProgram<float>> progs;
  size_t nblk = 0, qunits = 0;
  size_t bwords = 0;  // the blocks' row-operation words (composed on the device)
  double cost = 0;
  uint64_t amps = 0;  // HBM amplitudes (0 when resident)
};

CircuitPlan plan_batch_circuit(const Circuit& c, const std::vector<PauliTerm>& terms, bool ref, int k) {
  constexpr int EB = kBatchEB, KH = 8 + kBatchEB;
  CircuitPlan cp;
  const int n = c.n;
  if (n < EB) {
    cp.err = NQPU_INVALID_PARAMETER;
    cp.msg = "batched circuits need at least 4 qubits";
    return cp;
  }
  const uint64_t lim = n >= 64 ? ~0ull : (bit(n) - 1);
  for (const PauliTerm& t : terms)
    if ((t.x | t.z) & ~lim) {
      cp.err = NQPU_INVALID_QUBIT;
      cp.msg = "observable term acts on a qubit beyond the circuit";
      return cp;
    }
  std::vector<HostOp> ops = merge_single_qubit(lower_circuit(c, ref));
  BatchCircuit& C = cp.C;
  C.n = n;
  C.out = (int32_t)k;
  if (n <= KH) {
    C.resident = 1;
    Pass p;
    for (int q = 0; q < n; ++q) p.qubits.push_back(q);
    p.ops = ops;
    BatchPass P{};
    P.K = n;
    P.tb = n - EB;
    for (int q = 0; q < n; ++q) P.q[q] = q;
    cp.prog = build_reg_program<float>(p, EB, n - EB).ops;
    P.op0 = 0;
    P.nops = (int)cp.prog.size();
    P.term0 = 0;
    for (const PauliTerm& t : terms) cp.bt.push_back(tile_term(t, p.qubits));
    P.nterm = (int)terms.size();
    C.pass0 = C.red0 = 0;
    C.npass = 1;
    cp.bp.push_back(P);
    cp.cost = (double)(ops.size() + 1) * (double)(1ull << n) / 64.0;
    return cp;
  }
  C.resident = 0;
  cp.amps = 1ull << n;
  PlanOptions po;
  po.tile_qubits = KH;
  po.low_qubits = 5;
  std::vector<Pass> passes = plan_tiles(ops, n, po);
  if (passes.empty()) {
    // no gate survives merging (an empty circuit, or ops that cancel): one
    // pass with an empty program still writes the |0...0> start state
    Pass p0;
    for (int q = 0; q < KH; ++q) p0.qubits.push_back(q);
    passes.push_back(p0);
  }
  C.pass0 = 0;
  C.npass = (int)passes.size();
  for (const Pass& p : passes) {
    BatchPass P{};
    P.K = KH;
    P.tb = 8;
    for (int j = 0; j < KH; ++j) P.q[j] = p.qubits[j];
    // matrix-core blocks and quadratic phases as in the standalone passes;
    // LDS remaps only (the kernel loads and stores tiles in the canonical
    // slot order)
    // (blocks as row operations, composed on the device)
    RegProgram<float> pr = build_reg_program<float>(p, EB, 8, true, false, 1 << 20, true, true);
    for (RegOp<float>& r : pr.ops)  // the circuit's phase-table unit (plan_batch rebases it)
      if (r.kind == R_QPHASE) r.s0 += (int32_t)cp.qunits;
    // the circuit's first block of this pass (plan_batch makes it global);
    // block row operations and phase tables go into the upload image straight
    // from these programs
    P.blob_unit = (int32_t)cp.nblk;
    P.scaled = pr.nblocks() ? 1 : 0;
    cp.nblk += pr.nblocks();
    cp.bwords += pr.block_words.size();
    cp.qunits += pr.qtab.size() * sizeof(float) / kMfmaBlobBytes;
    P.op0 = (int)cp.prog.size();
    P.nops = (int)pr.ops.size();
    cp.prog.insert(cp.prog.end(), pr.ops.begin(), pr.ops.end());
    cp.bp.push_back(P);
    cp.progs.push_back(std::move(pr));
  }
  // observable passes: each term on a tile set that holds its X bits
  // The first set is the lowest KH qubits; the second holds the qubits
  // above it (at most KH - 4 of them, the highest) together with the lowest
  // qubits, so a tile's lanes still load runs of at least 16 consecutive
  // amplitudes (the top KH qubits alone would put lanes 2^(n-KH) apart).
  std::vector<std::vector<int>> sets(1);
  for (int q = 0; q < KH; ++q) sets[0].push_back(q);
  if (n > KH) {
    const int hi = std::min(n - KH, KH - 4);
    sets.emplace_back();
    for (int q = 0; q < KH - hi; ++q) sets[1].push_back(q);
    for (int q = n - hi; q < n; ++q) sets[1].push_back(q);
  }
  std::vector<std::vector<PauliTerm>> groups(sets.size());
  auto setmask = [](const std::vector<int>& s) {
    uint64_t m = 0;
    for (int q : s) m |= bit(q);
    return m;
  };
  for (const PauliTerm& t : terms) {
    size_t g = 0;
    for (; g < sets.size(); ++g)
      if ((t.x & ~setmask(sets[g])) == 0) break;
    if (g == sets.size()) {
      if (__builtin_popcountll(t.x) > KH) {
        cp.err = NQPU_INVALID_PARAMETER;
        cp.msg = "an observable term has X/Y on more qubits than a tile holds";
        return cp;
      }
      uint64_t m = t.x;
      for (int q = 0; q < n && __builtin_popcountll(m) < KH; ++q) m |= bit(q);
      std::vector<int> st;
      for (int q = 0; q < n; ++q)
        if ((m >> q) & 1) st.push_back(q);
      sets.push_back(st);
      groups.emplace_back();
    }
    groups[g].push_back(t);
  }
  C.red0 = (int)cp.bp.size();
  for (size_t g = 0; g < sets.size(); ++g) {
    if (groups[g].empty()) continue;
    BatchPass P{};
    P.K = KH;
    P.tb = 8;
    for (int j = 0; j < KH; ++j) P.q[j] = sets[g][j];
    P.term0 = (int)cp.bt.size();
    for (const PauliTerm& t : groups[g]) cp.bt.push_back(tile_term(t, sets[g]));
    P.nterm = (int)groups[g].size();
    cp.bp.push_back(P);
    C.nred++;
  }
  cp.cost = (double)(C.npass + C.nred) * (double)(1ull << n);
  return cp;
}
}  // namespace

// Circuits are planned independently on a pool of host threads, then their
// tables are concatenated in circuit order (the result does not depend on the
// number of threads).
NqpuError plan_batch(const std::vector<const Circuit*>& circuits, const std::vector<std::vector<PauliTerm>>& terms,
                     const NqpuRunOptions& o, BatchPlan& plan, const std::function<char*(size_t)>& buffer) {
  const bool ref = o.semantics == NQPU_SEMANTICS_REFERENCE;
  std::vector<CircuitPlan> cps(circuits.size());
  {
    std::atomic<size_t> next{0};
    auto work = [&] {
      for (size_t k; (k = next.fetch_add(1)) < circuits.size();)
        cps[k] = plan_batch_circuit(*circuits[k], terms[k], ref, (int)k);
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nth = std::min<size_t>(hw, std::max<size_t>(1, circuits.size() / 64));
    std::vector<std::thread> th;
    for (size_t t = 1; t < nth; ++t) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
  }
  const size_t nc = cps.size();
  // offsets of each circuit's passes, program ops and terms in the tables
  std::vector<size_t> bp_at(nc + 1, 0), pg_at(nc + 1, 0), bt_at(nc + 1, 0), bl_at(nc + 1, 0), qt_at(nc + 1, 0);
  std::vector<size_t> bw_at(nc + 1, 0);
  std::vector<BatchCircuit> bc(nc);
  uint64_t slot_amps = 0;
  NqpuBatchStats st{};
  for (size_t k = 0; k < nc; ++k) {
    const CircuitPlan& cp = cps[k];
    if (cp.err != NQPU_OK) {
      set_error(cp.msg);
      return cp.err;
    }
    bp_at[k + 1] = bp_at[k] + cp.bp.size();
    pg_at[k + 1] = pg_at[k] + cp.prog.size();
    bt_at[k + 1] = bt_at[k] + cp.bt.size();
    bl_at[k + 1] = bl_at[k] + cp.nblk;    // matrix-core blocks
    qt_at[k + 1] = qt_at[k] + cp.qunits;  // phase-table units
    bw_at[k + 1] = bw_at[k] + cp.bwords;  // block row-operation words
    BatchCircuit C = cp.C;
    C.pass0 += (int32_t)bp_at[k];
    C.red0 += (int32_t)bp_at[k];
    if (!C.resident) {
      C.amps = cp.amps;
      // split circuits have their own state (offset assigned below)
      if (C.n < kBatchSplitQubits) slot_amps = std::max(slot_amps, cp.amps);
      st.gate_passes += C.npass;
      st.observable_passes += C.nred;
      // gate passes write the state and read it (all but the first, which
      // starts from |0...0>); observable passes read it
      st.hbm_bytes += (2.0 * C.npass - 1.0 + C.nred) * (double)cp.amps * 8.0;
    } else {
      st.resident++;
    }
    bc[k] = C;
  }
  // The persistent kernel's queue: circuits below kBatchSplitQubits, longest
  // first (handed out before the short ones); then the split circuits, whose
  // passes run as grids of work items.
  auto split = [&](size_t k) { return !bc[k].resident && bc[k].n >= kBatchSplitQubits; };
  std::vector<size_t> order(nc);
  for (size_t i = 0; i < nc; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
    if (split(x) != split(y)) return !split(x);
    return cps[x].cost > cps[y].cost;
  });
  size_t nqueue = 0;
  while (nqueue < nc && !split(order[nqueue])) ++nqueue;
  // split circuits: waves of circuits whose states fit kBatchSplitWaveAmps
  // (at least one circuit each); per wave, gate items per pass index, then
  // its observable items
  auto chunks = [](int n, int K, int32_t circ, int32_t pass, std::vector<BatchItem>& v) {
    const int64_t ntiles = 1ll << (n - K);
    for (int64_t t = 0; t < ntiles; t += kBatchSplitTiles)
      v.push_back(BatchItem{circ, pass, (int32_t)t, (int32_t)std::min<int64_t>(kBatchSplitTiles, ntiles - t)});
  };
  uint64_t split_amps = 0;
  std::vector<BatchItem> items;
  std::vector<BatchPlan::SplitLaunch> launches;
  std::vector<int32_t> ofirst(nc - nqueue), ocount(nc - nqueue);
  size_t parts = 0;
  for (size_t w0 = nqueue; w0 < nc;) {
    size_t w1 = w0;
    uint64_t wave_amps = 0;
    int max_pass = 0;
    while (w1 < nc && (w1 == w0 || wave_amps + bc[order[w1]].amps <= kBatchSplitWaveAmps)) {
      BatchCircuit& C = bc[order[w1]];
      C.off = wave_amps;  // the wave's states share one region, reused by the next wave
      wave_amps += C.amps;
      max_pass = std::max(max_pass, (int)C.npass);
      ++w1;
    }
    split_amps = std::max(split_amps, wave_amps);
    for (int p = 0; p < max_pass; ++p) {
      const size_t at = items.size();
      for (size_t i = w0; i < w1; ++i) {
        const size_t k = order[i];
        if (p < bc[k].npass) chunks(bc[k].n, cps[k].bp[p].K, (int32_t)i, p, items);
      }
      if (items.size() > at) launches.push_back({at, items.size() - at, false, 0});
    }
    const size_t at = items.size();
    for (size_t i = w0; i < w1; ++i) {
      const size_t k = order[i];
      const size_t c0 = items.size();
      for (int rp = 0; rp < bc[k].nred; ++rp) chunks(bc[k].n, cps[k].bp[bc[k].npass + rp].K, (int32_t)i, rp, items);
      ofirst[i - nqueue] = (int32_t)(parts + (c0 - at));
      ocount[i - nqueue] = (int32_t)(items.size() - c0);
    }
    if (items.size() > at) launches.push_back({at, items.size() - at, true, parts});
    parts += items.size() - at;
    w0 = w1;
  }
  const size_t nsplit = nc - nqueue;
  // one table image: [circuits][passes][program][terms][counter][items][first][count]
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t oc = 0, op_ = up(oc + nc * sizeof(BatchCircuit));
  const size_t og = up(op_ + bp_at[nc] * sizeof(BatchPass));
  const size_t ot = up(og + pg_at[nc] * sizeof(RegOp<float>));
  const size_t ocnt = up(ot + bt_at[nc] * sizeof(BatchTerm));
  const size_t oit = ocnt + 256;
  const size_t ofi = up(oit + items.size() * sizeof(BatchItem));
  const size_t ocn = up(ofi + nsplit * sizeof(int32_t));
  const size_t head = up(ocn + nsplit * sizeof(int32_t));  // operand units follow, 256-byte aligned
  const size_t nblk = bl_at[nc], nq = qt_at[nc], nw = bw_at[nc];
  // host image after the head: phase tables, the blocks' word index
  // (nblk + 1 entries), their row-operation words; the device copy of it
  // follows the blocks' operand blobs
  const size_t hq = head, hbi = hq + nq * kMfmaBlobBytes, hbw = up(hbi + (nblk + 1) * sizeof(uint32_t));
  const size_t total = hbw + nw * sizeof(uint64_t);
  if (nblk + nq > (size_t)INT32_MAX || nw > (size_t)UINT32_MAX) {
    set_error("batch operand units exceed the 32-bit unit index");
    return NQPU_INVALID_PARAMETER;
  }
  char* host = buffer(total);
  if (!host) {
    set_error("no host memory for the batch tables");
    return NQPU_ALLOCATION_FAILED;
  }
  for (size_t i = 0; i < nc; ++i)
    std::memcpy(host + oc + i * sizeof(BatchCircuit), &bc[order[i]], sizeof(BatchCircuit));
  std::memset(host + ocnt, 0, 256);  // the work-queue counter starts at zero
  if (!items.empty()) std::memcpy(host + oit, items.data(), items.size() * sizeof(BatchItem));
  if (nsplit) {
    std::memcpy(host + ofi, ofirst.data(), nsplit * sizeof(int32_t));
    std::memcpy(host + ocn, ocount.data(), nsplit * sizeof(int32_t));
  }
  {
    std::atomic<size_t> next{0};
    auto fill = [&] {
      for (size_t k; (k = next.fetch_add(1)) < nc;) {
        const CircuitPlan& cp = cps[k];
        BatchPass* dp = reinterpret_cast<BatchPass*>(host + op_) + bp_at[k];
        RegOp<float>* dg = reinterpret_cast<RegOp<float>*>(host + og) + pg_at[k];
        if (!cp.prog.empty()) std::memcpy(dg, cp.prog.data(), cp.prog.size() * sizeof(RegOp<float>));
        for (size_t i = 0; i < cp.bp.size(); ++i) {
          BatchPass P = cp.bp[i];
          P.op0 += (int32_t)pg_at[k];
          P.term0 += (int32_t)bt_at[k];
          P.blob_unit += (int32_t)bl_at[k];
          // operand units count from the blob area: blocks, then phase tables
          for (int j = 0; j < cp.bp[i].nops && !cp.C.resident; ++j) {
            RegOp<float>& r = dg[cp.bp[i].op0 + j];
            if (r.kind == R_QPHASE) r.s0 += (int32_t)(nblk + qt_at[k]) - P.blob_unit;
          }
          std::memcpy(dp + i, &P, sizeof(BatchPass));
        }
        char* u = host + hq + qt_at[k] * kMfmaBlobBytes;
        uint32_t* bi = reinterpret_cast<uint32_t*>(host + hbi) + bl_at[k];
        uint64_t* bw = reinterpret_cast<uint64_t*>(host + hbw) + bw_at[k];
        size_t w = bw_at[k];
        for (const RegProgram<float>& pr : cp.progs) {
          for (uint32_t at : pr.block_at) *bi++ = (uint32_t)(w + at);
          if (!pr.block_words.empty()) std::memcpy(bw, pr.block_words.data(), pr.block_words.size() * sizeof(uint64_t));
          bw += pr.block_words.size();
          w += pr.block_words.size();
          if (!pr.qtab.empty()) {
            std::memcpy(u, pr.qtab.data(), pr.qtab.size() * sizeof(float));
            u += pr.qtab.size() * sizeof(float);
          }
        }
        if (!cp.bt.empty())
          std::memcpy(reinterpret_cast<BatchTerm*>(host + ot) + bt_at[k], cp.bt.data(), cp.bt.size() * sizeof(BatchTerm));
      }
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nth = std::min<size_t>(hw, std::max<size_t>(1, nc / 64));
    std::vector<std::thread> th;
    for (size_t t = 1; t < nth; ++t) th.emplace_back(fill);
    fill();
    for (auto& x : th) x.join();
  }
  reinterpret_cast<uint32_t*>(host + hbi)[nblk] = (uint32_t)nw;
  plan.host = host;
  plan.oc = oc;
  plan.op = op_;
  plan.og = og;
  plan.ot = ot;
  plan.ocnt = ocnt;
  plan.head = head;
  plan.nblk = nblk;
  plan.qunits = nq;
  plan.dev_q = head + nblk * kMfmaBlobBytes;
  plan.dev_bi = plan.dev_q + (hbi - hq);
  plan.dev_bw = plan.dev_q + (hbw - hq);
  plan.dev_m = up(plan.dev_q + (total - hq));
  plan.dev_total = plan.dev_m + nblk * kMfmaMatrixBytes;
  plan.total = total;
  plan.slot_amps = slot_amps;
  plan.ncirc = nc;
  plan.nqueue = nqueue;
  plan.split_amps = split_amps;
  plan.oit = oit;
  plan.ofirst = ofi;
  plan.ocount = ocn;
  plan.split_launches = launches;
  plan.split_parts = parts;
  plan.program_ops = pg_at[nc];
  st.circuits = circuits.size();
  st.state_bytes = (double)(slot_amps + split_amps) * 8.0;
  plan.st = st;
  return NQPU_OK;
}

// Complete a queued batch: wait for it, copy its results to the caller's
// buffer, and record its statistics.
static NqpuError retire(Batch* b, BatchSlot& sl) {
  if (!sl.busy) return NQPU_OK;
  sl.busy = false;
  HIPTRY(hipEventSynchronize(sl.done), "batch");
  float ms = 0;
  if (sl.count) (void)hipEventElapsedTime(&ms, sl.k0, sl.k1);
  if (sl.count) std::memcpy(sl.dest, sl.results, sl.count * sizeof(double));
  sl.st.kernel_ms = ms;
  b->stats = sl.st;
  return NQPU_OK;
}

NqpuError batch_synchronize(Batch* b) {
  // in the order the batches were queued
  for (int i = 0; i < 2; ++i) {
    BatchSlot& sl = b->slot[(b->next + i) & 1];
    const NqpuError e = retire(b, sl);
    if (e != NQPU_OK) return e;
  }
  return NQPU_OK;
}

NqpuError batch_expectation(Batch* b, const std::vector<const Circuit*>& circuits,
                            const std::vector<std::vector<PauliTerm>>& terms, const NqpuRunOptions& o,
                            double* out) {
  using clock = std::chrono::steady_clock;
  BatchSlot& sl = b->slot[b->next];
  // the slot's previous batch (two batches ago) must be complete before its
  // host image, device tables and results are reused
  NqpuError e = retire(b, sl);
  if (e != NQPU_OK) return e;
  const auto t0 = clock::now();
  BatchPlan plan;
  // the tables are written straight into the slot's pinned host buffer
  auto pinned = [&sl](size_t bytes) -> char* {
    if (bytes > sl.pinned_bytes) {
      if (sl.pinned) (void)hipHostFree(sl.pinned);
      sl.pinned = nullptr;
      sl.pinned_bytes = 0;
      void* p = nullptr;
      if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
      sl.pinned = (char*)p;
      sl.pinned_bytes = bytes;
    }
    return sl.pinned;
  };
  NqpuError pe = plan_batch(circuits, terms, o, plan, pinned);
  if (pe != NQPU_OK) return pe;
  NqpuBatchStats st = plan.st;
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0);
  const int grid = (int)std::min<size_t>(plan.nqueue, (size_t)cus * 3);  // 3 workgroups fit per CU
  // one state slot per workgroup: a workgroup runs its queued circuits one
  // after another, so the HBM it needs is that of the largest one; then the
  // split circuits' own states
  const uint64_t slots_amps = plan.slot_amps * (uint64_t)std::max(grid, 1);
  const uint64_t state_amps = slots_amps + plan.split_amps;
  const size_t nres = circuits.size();
  const size_t out_doubles = ((nres + 31) & ~size_t(31)) + plan.split_parts;  // results, then item parts
  // growing a device buffer frees memory a queued batch may still use
  if ((state_amps * 8 > b->states.bytes || plan.dev_total > sl.tables.bytes || out_doubles * 8 + 8 > sl.out.bytes)) {
    if ((e = batch_synchronize(b)) != NQPU_OK) return e;
    HIPTRY(hipStreamSynchronize(b->stream), "sync before batch buffer growth");
    HIPTRY(hipStreamSynchronize(b->copy_stream), "sync before batch buffer growth");
  }
  if (!sl.tables.ensure(plan.dev_total) || !sl.out.ensure(out_doubles * 8 + 8)) return NQPU_OUT_OF_DEVICE_MEMORY;
  if (state_amps && !b->states.ensure(state_amps * 8)) {
    set_error("batch states do not fit in device memory");
    return NQPU_OUT_OF_DEVICE_MEMORY;
  }
  if (nres > sl.results_cap) {
    if (sl.results) (void)hipHostFree(sl.results);
    sl.results = nullptr;
    sl.results_cap = 0;
    HIPTRY(hipHostMalloc((void**)&sl.results, nres * sizeof(double)), "pinned batch results");
    sl.results_cap = nres;
  }
  for (hipEvent_t* ev : {&sl.uploaded, &sl.done})
    if (!*ev) HIPTRY(hipEventCreateWithFlags(ev, hipEventDisableTiming), "batch event");
  for (hipEvent_t* ev : {&sl.k0, &sl.k1})
    if (!*ev) HIPTRY(hipEventCreate(ev), "batch event");
  char* base = (char*)sl.tables.p;
  // the head, then the phase tables and block matrices; the blocks' operand
  // blobs are built on the device between them.  The copies run on their own
  // stream, overlapping the previous batch's launch (the slot's own previous
  // batch is complete: retired above); the launch waits for them.
  HIPTRY(hipMemcpyAsync(base, plan.host, plan.head, hipMemcpyHostToDevice, b->copy_stream), "batch table upload");
  if (plan.total > plan.head)
    HIPTRY(hipMemcpyAsync(base + plan.dev_q, plan.host + plan.head, plan.total - plan.head, hipMemcpyHostToDevice,
                          b->copy_stream), "batch operand upload");
  HIPTRY(hipEventRecord(sl.uploaded, b->copy_stream), "batch upload event");
  HIPTRY(hipStreamWaitEvent(b->stream, sl.uploaded, 0), "batch upload wait");
  if (plan.nblk) {
    launch_block_compose((const uint32_t*)(base + plan.dev_bi), (const uint64_t*)(base + plan.dev_bw),
                         (float2*)(base + plan.dev_m), plan.nblk, b->stream);
    HIPTRY(hipGetLastError(), "block composition launch");
    launch_blob_split((const float2*)(base + plan.dev_m), (unsigned char*)(base + plan.head), plan.nblk, b->stream);
    HIPTRY(hipGetLastError(), "operand blob launch");
  }
  HIPTRY(hipEventRecord(sl.k0, b->stream), "batch event");
  const BatchCircuit* dcirc = (const BatchCircuit*)(base + plan.oc);
  const BatchPass* dpass = (const BatchPass*)(base + plan.op);
  const RegOp<float>* dprog = (const RegOp<float>*)(base + plan.og);
  const BatchTerm* dterm = (const BatchTerm*)(base + plan.ot);
  const unsigned char* dblob = (const unsigned char*)(base + plan.head);
  const BatchItem* ditems = (const BatchItem*)(base + plan.oit);
  if (grid > 0)
    launch_batch(b->states.p, plan.slot_amps, dcirc, (int)plan.nqueue, dpass, dprog, dterm, dblob,
                 (unsigned*)(base + plan.ocnt), (double*)sl.out.p, grid, b->stream);
  HIPTRY(hipGetLastError(), "batch launch");
  // split circuits: per wave, one launch per gate pass index, then its
  // observable items; then every split circuit's sum of its parts
  float2* split_states = (float2*)b->states.p + slots_amps;
  double* part = (double*)sl.out.p + ((nres + 31) & ~size_t(31));
  for (const BatchPlan::SplitLaunch& l : plan.split_launches) {
    if (l.obs) {
      launch_batch_split_obs(split_states, dcirc, dpass, dterm, ditems + l.first, (int)l.count, part + l.part0,
                             b->stream);
      HIPTRY(hipGetLastError(), "split circuit observable launch");
    } else {
      launch_batch_split_gate(split_states, dcirc, dpass, dprog, dblob, ditems + l.first, (int)l.count, b->stream);
      HIPTRY(hipGetLastError(), "split circuit pass launch");
    }
  }
  const int nsplit = (int)(plan.ncirc - plan.nqueue);
  if (nsplit) {
    launch_batch_split_out(dcirc, (int)plan.nqueue, nsplit, (const int32_t*)(base + plan.ofirst),
                           (const int32_t*)(base + plan.ocount), part, (double*)sl.out.p, b->stream);
    HIPTRY(hipGetLastError(), "split circuit result launch");
  }
  HIPTRY(hipEventRecord(sl.k1, b->stream), "batch event");
  if (nres)
    HIPTRY(hipMemcpyAsync(sl.results, sl.out.p, nres * sizeof(double), hipMemcpyDeviceToHost, b->stream),
           "batch results");
  HIPTRY(hipEventRecord(sl.done, b->stream), "batch done event");
  st.circuits = circuits.size();
  st.plan_ms = std::chrono::duration<double, std::milli>(clock::now() - t0).count();
  st.state_bytes = (double)state_amps * 8.0;
  sl.st = st;
  sl.dest = out;
  sl.count = nres;
  sl.busy = true;
  b->next ^= 1;
  if (o.flags & NQPU_RUN_ASYNC) return NQPU_OK;
  return batch_synchronize(b);
}

// ------------------------------------------------------------------- RNG

uint64_t rng_init(uint64_t seed) { return seed == 0 ? 0x853c49e6748fea9bULL : seed; }
double rng_step(uint64_t& state) {
  uint64_t x = state;
  x ^= x >> 12;
  x ^= x << 25;
  x ^= x >> 27;
  state = x;
  return (double)((x * 0x2545F4914F6CDD1DULL) >> 11) / (double)(1ull << 53);
}

static thread_local uint64_t g_rng = 0x853c49e6748fea9bULL;
void rng_seed(uint64_t seed) { g_rng = rng_init(seed); }
double rng_next() { return rng_step(g_rng); }

}  // namespace nqpu
