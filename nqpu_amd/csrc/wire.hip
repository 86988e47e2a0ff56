// wire.hip — the reference's request formats on the engine's boundary.
//
// Two JSON circuit formats are accepted:
//  * the web API's CircuitRequest / SampleRequest (sdk/rust/src/web/mod.rs:
//    34-131, 149-160): {"num_qubits": n, "gates": [{"type": "h", "target": 0},
//    {"type": "cnot", "control": 0, "target": 1}, ...], "backend": "auto",
//    "shots": 1024}, gate fields as GateRequest's variants;
//  * circuit_serde's SerializableCircuit (sdk/rust/src/circuits/formats/
//    circuit_serde.rs:329-340, 435-520, 870-930): {"version": 1,
//    "num_qubits": n, "gates": [{"gate_type": "CNOT", "targets": [1],
//    "controls": [0], "params": []}], "metadata": {...}}.
// A request may add "observables": [{"pauli": "ZZI", "coeff": -1.0}] (qubit
// q = character q, as expectation_pauli_string, lib.rs:465); the server
// returns coeff * <P> for each.
//
// Responses follow ExecuteResponse / SampleResponse (web/mod.rs:133-172):
// probabilities, time_ms, backend_used, num_gates and fidelity for an
// execution; outcomes (MSB-first bitstring -> count), shots, time_ms and
// backend_used for a sampling request.  The reference's `sample` re-simulates
// the circuit for every shot (web/mod.rs:378-419); here a sampling request
// runs once on the request server and draws all shots from the final state.
#include "runtime.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

namespace nqpu {

namespace {

// ------------------------------------------------------------------ JSON

struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0;
  bool integral = false;  // the number was written without fraction or exponent
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
  const JVal* get(const char* key) const {
    if (kind != OBJ) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

class Parser {
 public:
  explicit Parser(const char* s) : p_(s) {}
  bool parse(JVal* out, std::string* err) {
    if (!value(out, 0)) {
      *err = err_.empty() ? "malformed JSON" : err_;
      return false;
    }
    ws();
    if (*p_) {
      *err = "trailing characters after the JSON value";
      return false;
    }
    return true;
  }

 private:
  const char* p_;
  std::string err_;
  void ws() {
    while (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r') ++p_;
  }
  bool fail(const char* m) {
    if (err_.empty()) err_ = m;
    return false;
  }
  bool lit(const char* w) {
    const size_t k = std::strlen(w);
    if (std::strncmp(p_, w, k) != 0) return fail("unknown literal");
    p_ += k;
    return true;
  }
  static void utf8(std::string& s, unsigned cp) {
    if (cp < 0x80) {
      s += (char)cp;
    } else if (cp < 0x800) {
      s += (char)(0xC0 | (cp >> 6));
      s += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      s += (char)(0xE0 | (cp >> 12));
      s += (char)(0x80 | ((cp >> 6) & 0x3F));
      s += (char)(0x80 | (cp & 0x3F));
    } else {
      s += (char)(0xF0 | (cp >> 18));
      s += (char)(0x80 | ((cp >> 12) & 0x3F));
      s += (char)(0x80 | ((cp >> 6) & 0x3F));
      s += (char)(0x80 | (cp & 0x3F));
    }
  }
  bool hex4(unsigned* v) {
    *v = 0;
    for (int i = 0; i < 4; ++i) {
      const char c = *p_++;
      *v <<= 4;
      if (c >= '0' && c <= '9') *v |= (unsigned)(c - '0');
      else if (c >= 'a' && c <= 'f') *v |= (unsigned)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') *v |= (unsigned)(c - 'A' + 10);
      else return fail("bad \\u escape");
    }
    return true;
  }
  bool string(std::string* out) {
    if (*p_ != '"') return fail("expected a string");
    ++p_;
    for (;;) {
      const char c = *p_;
      if (!c) return fail("unterminated string");
      ++p_;
      if (c == '"') return true;
      if ((unsigned char)c < 0x20) return fail("control character in a string");
      if (c != '\\') {
        *out += c;
        continue;
      }
      const char e = *p_++;
      switch (e) {
        case '"': *out += '"'; break;
        case '\\': *out += '\\'; break;
        case '/': *out += '/'; break;
        case 'b': *out += '\b'; break;
        case 'f': *out += '\f'; break;
        case 'n': *out += '\n'; break;
        case 'r': *out += '\r'; break;
        case 't': *out += '\t'; break;
        case 'u': {
          unsigned cp;
          if (!hex4(&cp)) return false;
          if (cp >= 0xD800 && cp < 0xDC00 && p_[0] == '\\' && p_[1] == 'u') {
            p_ += 2;
            unsigned lo;
            if (!hex4(&lo)) return false;
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(*out, cp);
          break;
        }
        default: return fail("bad escape");
      }
    }
  }
  bool number(JVal* v) {
    const char* s = p_;
    bool integral = true;
    if (*p_ == '-') ++p_;
    if (!(*p_ >= '0' && *p_ <= '9')) return fail("bad number");
    while (*p_ >= '0' && *p_ <= '9') ++p_;
    if (*p_ == '.') {
      integral = false;
      ++p_;
      if (!(*p_ >= '0' && *p_ <= '9')) return fail("bad number");
      while (*p_ >= '0' && *p_ <= '9') ++p_;
    }
    if (*p_ == 'e' || *p_ == 'E') {
      integral = false;
      ++p_;
      if (*p_ == '+' || *p_ == '-') ++p_;
      if (!(*p_ >= '0' && *p_ <= '9')) return fail("bad number");
      while (*p_ >= '0' && *p_ <= '9') ++p_;
    }
    v->kind = JVal::NUM;
    v->num = std::strtod(std::string(s, p_).c_str(), nullptr);
    v->integral = integral;
    return true;
  }
  bool value(JVal* v, int depth) {
    if (depth > 64) return fail("JSON nested too deeply");
    ws();
    switch (*p_) {
      case '{': {
        ++p_;
        v->kind = JVal::OBJ;
        ws();
        if (*p_ == '}') {
          ++p_;
          return true;
        }
        for (;;) {
          ws();
          std::string k;
          if (!string(&k)) return false;
          ws();
          if (*p_ != ':') return fail("expected ':'");
          ++p_;
          JVal x;
          if (!value(&x, depth + 1)) return false;
          v->obj.emplace_back(std::move(k), std::move(x));
          ws();
          if (*p_ == ',') {
            ++p_;
            continue;
          }
          if (*p_ == '}') {
            ++p_;
            return true;
          }
          return fail("expected ',' or '}'");
        }
      }
      case '[': {
        ++p_;
        v->kind = JVal::ARR;
        ws();
        if (*p_ == ']') {
          ++p_;
          return true;
        }
        for (;;) {
          JVal x;
          if (!value(&x, depth + 1)) return false;
          v->arr.push_back(std::move(x));
          ws();
          if (*p_ == ',') {
            ++p_;
            continue;
          }
          if (*p_ == ']') {
            ++p_;
            return true;
          }
          return fail("expected ',' or ']'");
        }
      }
      case '"':
        v->kind = JVal::STR;
        return string(&v->str);
      case 't':
        v->kind = JVal::BOOL;
        v->b = true;
        return lit("true");
      case 'f':
        v->kind = JVal::BOOL;
        return lit("false");
      case 'n':
        return lit("null");
      default:
        return number(v);
    }
  }
};

// ------------------------------------------------------------- requests

struct ParsedRequest {
  Circuit circuit;
  std::vector<PauliTerm> terms;
  bool has_shots = false;
  uint32_t shots = 0;
  std::string format;  // "web" or "circuit_serde"
};

NqpuError bad(const std::string& m) {
  set_error("request JSON: " + m);
  return NQPU_INVALID_PARAMETER;
}

bool as_index(const JVal* v, uint32_t* out) {
  if (!v || v->kind != JVal::NUM || !v->integral || v->num < 0 || v->num > 4294967295.0) return false;
  *out = (uint32_t)v->num;
  return true;
}

// GateRequest (web/mod.rs:34-82): "type" names a variant, its fields follow.
NqpuError web_gate(const JVal& g, NqpuGateOp* op) {
  const JVal* t = g.get("type");
  if (!t || t->kind != JVal::STR) return bad("gate without a \"type\"");
  struct Form {
    const char* name;
    uint32_t kind;
    int shape;  // 0: target; 1: target, angle; 2: control, target; 3: q0, q1; 4: control, target, angle;
                // 5: control1, control2, target
  };
  static const Form forms[] = {
      {"h", NQPU_GATE_H, 0},        {"x", NQPU_GATE_X, 0},         {"y", NQPU_GATE_Y, 0},
      {"z", NQPU_GATE_Z, 0},        {"s", NQPU_GATE_S, 0},         {"t", NQPU_GATE_T, 0},
      {"sx", NQPU_GATE_SX, 0},      {"rx", NQPU_GATE_RX, 1},       {"ry", NQPU_GATE_RY, 1},
      {"rz", NQPU_GATE_RZ, 1},      {"phase", NQPU_GATE_PHASE, 1}, {"cnot", NQPU_GATE_CNOT, 2},
      {"cz", NQPU_GATE_CZ, 2},      {"swap", NQPU_GATE_SWAP, 3},   {"cr", NQPU_GATE_CR, 4},
      {"iswap", NQPU_GATE_ISWAP, 3}, {"toffoli", NQPU_GATE_TOFFOLI, 5}};
  const Form* f = nullptr;
  for (const Form& x : forms)
    if (t->str == x.name) f = &x;
  if (!f) return bad("unknown gate type \"" + t->str + "\"");
  std::memset(op, 0, sizeof(*op));
  op->kind = f->kind;
  auto field = [&](const char* k, uint32_t* v) { return as_index(g.get(k), v); };
  auto angle = [&](double* a) {
    const JVal* v = g.get("angle");
    if (!v || v->kind != JVal::NUM) return false;
    *a = v->num;
    return true;
  };
  bool ok = true;
  switch (f->shape) {
    case 0:
      op->num_targets = 1;
      ok = field("target", &op->targets[0]);
      break;
    case 1:
      op->num_targets = 1;
      ok = field("target", &op->targets[0]) && angle(&op->params[0]);
      break;
    case 2:
      op->num_targets = op->num_controls = 1;
      ok = field("control", &op->controls[0]) && field("target", &op->targets[0]);
      break;
    case 3:
      op->num_targets = 2;
      ok = field("q0", &op->targets[0]) && field("q1", &op->targets[1]);
      break;
    case 4:
      op->num_targets = op->num_controls = 1;
      ok = field("control", &op->controls[0]) && field("target", &op->targets[0]) && angle(&op->params[0]);
      break;
    default:
      op->num_targets = 1;
      op->num_controls = 2;
      ok = field("control1", &op->controls[0]) && field("control2", &op->controls[1]) &&
           field("target", &op->targets[0]);
      break;
  }
  if (!ok) return bad("gate \"" + t->str + "\" has a missing or malformed field");
  return NQPU_OK;
}

// SerializableGate (circuit_serde.rs:154-163, names of to_gate_type :226-270
// plus the two-qubit rotations, CSWAP and CU that from_gate_type writes).
NqpuError serde_gate(const JVal& g, NqpuGateOp* op) {
  const JVal* t = g.get("gate_type");
  if (!t || t->kind != JVal::STR) return bad("gate without a \"gate_type\"");
  static const std::pair<const char*, uint32_t> names[] = {
      {"H", NQPU_GATE_H},         {"X", NQPU_GATE_X},         {"Y", NQPU_GATE_Y},
      {"Z", NQPU_GATE_Z},         {"S", NQPU_GATE_S},         {"T", NQPU_GATE_T},
      {"Rx", NQPU_GATE_RX},       {"Ry", NQPU_GATE_RY},       {"Rz", NQPU_GATE_RZ},
      {"CNOT", NQPU_GATE_CNOT},   {"CZ", NQPU_GATE_CZ},       {"SWAP", NQPU_GATE_SWAP},
      {"Toffoli", NQPU_GATE_TOFFOLI}, {"CRx", NQPU_GATE_CRX}, {"CRy", NQPU_GATE_CRY},
      {"CRz", NQPU_GATE_CRZ},     {"CR", NQPU_GATE_CR},       {"SX", NQPU_GATE_SX},
      {"Phase", NQPU_GATE_PHASE}, {"ISWAP", NQPU_GATE_ISWAP}, {"CCZ", NQPU_GATE_CCZ},
      {"U", NQPU_GATE_U},         {"Rxx", NQPU_GATE_RXX},     {"Ryy", NQPU_GATE_RYY},
      {"Rzz", NQPU_GATE_RZZ},     {"CSWAP", NQPU_GATE_CSWAP}, {"CU", NQPU_GATE_CU}};
  int kind = -1;
  for (const auto& nm : names)
    if (t->str == nm.first) kind = (int)nm.second;
  if (kind < 0) return bad("unsupported gate_type \"" + t->str + "\"");
  std::memset(op, 0, sizeof(*op));
  op->kind = (uint32_t)kind;
  auto arr = [&](const char* k, uint32_t* out, uint32_t cap, uint32_t* count) -> bool {
    const JVal* a = g.get(k);
    *count = 0;
    if (!a) return true;  // absent arrays are empty (extract_usize_array, :933)
    if (a->kind != JVal::ARR || a->arr.size() > cap) return false;
    for (const JVal& x : a->arr)
      if (!as_index(&x, &out[(*count)++])) return false;
    return true;
  };
  if (!arr("targets", op->targets, 3, &op->num_targets) || !arr("controls", op->controls, 3, &op->num_controls))
    return bad("gate \"" + t->str + "\" has malformed targets or controls");
  const JVal* p = g.get("params");
  size_t np = 0;
  if (p) {
    if (p->kind != JVal::ARR || p->arr.size() > 4) return bad("gate \"" + t->str + "\" has malformed params");
    for (size_t i = 0; i < p->arr.size(); ++i) {
      if (p->arr[i].kind != JVal::NUM) return bad("non-number in 'params'");
      op->params[i] = p->arr[i].num;
    }
    np = p->arr.size();
  }
  // a parametric gate without its parameters has no GateType (to_gate_type, :232-270)
  size_t need = 0;
  switch (kind) {
    case NQPU_GATE_RX: case NQPU_GATE_RY: case NQPU_GATE_RZ: case NQPU_GATE_PHASE: case NQPU_GATE_CRX:
    case NQPU_GATE_CRY: case NQPU_GATE_CRZ: case NQPU_GATE_CR: case NQPU_GATE_RXX: case NQPU_GATE_RYY:
    case NQPU_GATE_RZZ: need = 1; break;
    case NQPU_GATE_U: need = 3; break;
    case NQPU_GATE_CU: need = 4; break;
    default: break;
  }
  if (np < need) return bad("gate \"" + t->str + "\" needs " + std::to_string(need) + " params");
  return NQPU_OK;
}

NqpuError parse_request(const char* text, ParsedRequest* out) {
  JVal root;
  std::string err;
  Parser ps(text);
  if (!ps.parse(&root, &err)) return bad(err);
  if (root.kind != JVal::OBJ) return bad("the request is not a JSON object");
  uint32_t n = 0;
  if (!as_index(root.get("num_qubits"), &n)) return bad("missing or malformed \"num_qubits\"");
  if (n < 1 || n > 34) {
    set_error("request JSON: num_qubits must be 1..34");
    return NQPU_INVALID_QUBIT;
  }
  const JVal* gates = root.get("gates");
  if (!gates || gates->kind != JVal::ARR) return bad("missing or malformed \"gates\"");
  const JVal* version = root.get("version");
  if (version && version->kind == JVal::NUM && version->num > 1.0)
    return bad("circuit format version newer than 1");
  const JVal* backend = root.get("backend");
  if (backend) {  // BackendType (web/mod.rs:17-29): accepted; the engine always runs on the GPU
    if (backend->kind != JVal::STR ||
        (backend->str != "auto" && backend->str != "cpu" && backend->str != "fused" && backend->str != "gpu"))
      return bad("unknown backend");
  }
  bool serde = false;
  if (!gates->arr.empty()) serde = gates->arr[0].get("gate_type") != nullptr;
  out->format = serde ? "circuit_serde" : "web";
  NqpuCircuit* c = nqpu_circuit_create_ex(n);
  if (!c) return bad("cannot create the circuit");
  std::unique_ptr<NqpuCircuit, void (*)(NqpuCircuit*)> hold(c, nqpu_circuit_destroy);
  for (size_t i = 0; i < gates->arr.size(); ++i) {
    const JVal& g = gates->arr[i];
    if (g.kind != JVal::OBJ) return bad("gate " + std::to_string(i) + " is not an object");
    NqpuGateOp op;
    NqpuError e = serde ? serde_gate(g, &op) : web_gate(g, &op);
    if (e != NQPU_OK) return e;
    e = nqpu_circuit_add_gate(c, &op);
    if (e != NQPU_OK) {
      set_error("request JSON: gate " + std::to_string(i) + " is invalid (qubits, arity or parameters)");
      return e;
    }
  }
  out->circuit = *reinterpret_cast<const Circuit*>(c);
  const JVal* shots = root.get("shots");
  if (shots) {
    if (!as_index(shots, &out->shots)) return bad("malformed \"shots\"");
    out->has_shots = true;
  }
  const JVal* obs = root.get("observables");
  if (obs) {
    if (obs->kind != JVal::ARR) return bad("\"observables\" is not an array");
    for (const JVal& o : obs->arr) {
      const JVal* ps = o.get("pauli");
      if (!ps || ps->kind != JVal::STR || ps->str.size() > n) return bad("observable without a valid \"pauli\" string");
      PauliTerm t{0, 0, 1.0};
      const JVal* cf = o.get("coeff");
      if (cf) {
        if (cf->kind != JVal::NUM) return bad("malformed observable coeff");
        t.coeff = cf->num;
      }
      for (size_t q = 0; q < ps->str.size(); ++q) {
        switch (ps->str[q]) {
          case 'I': case 'i': break;
          case 'X': case 'x': t.x |= 1ull << q; break;
          case 'Y': case 'y': t.x |= 1ull << q; t.z |= 1ull << q; break;
          case 'Z': case 'z': t.z |= 1ull << q; break;
          default: return bad("unknown Pauli operator in an observable");
        }
      }
      out->terms.push_back(t);
    }
  }
  return NQPU_OK;
}

// ------------------------------------------------------------ responses

void num(std::string& s, double v) {
  char b[32];
  if (!std::isfinite(v)) {
    s += "null";
    return;
  }
  std::snprintf(b, sizeof b, "%.17g", v);
  s += b;
}

std::string bits(uint64_t i, int n) {
  std::string s((size_t)n, '0');
  for (int q = 0; q < n; ++q)
    if ((i >> q) & 1) s[(size_t)(n - 1 - q)] = '1';
  return s;
}

NqpuError deliver(const std::string& body, char* buf, uint64_t len, uint64_t* needed) {
  if (needed) *needed = body.size() + 1;
  if (!buf || len < body.size() + 1) {
    set_error("response buffer too small: " + std::to_string(body.size() + 1) + " bytes needed");
    return NQPU_INVALID_PARAMETER;
  }
  std::memcpy(buf, body.c_str(), body.size() + 1);
  return NQPU_OK;
}

// finished JSON responses of server requests not yet handed over
std::mutex g_resp_mu;
std::map<std::pair<const void*, uint64_t>, std::string> g_responses;

}  // namespace

}  // namespace nqpu

using namespace nqpu;

extern "C" {

NqpuError nqpu_circuit_from_json(const char* json, NqpuCircuit** circuit, uint32_t* shots) {
  if (!json || !circuit) return NQPU_NULL_POINTER;
  *circuit = nullptr;
  ParsedRequest r;
  NqpuError e = parse_request(json, &r);
  if (e != NQPU_OK) return e;
  Circuit* c = new (std::nothrow) Circuit(std::move(r.circuit));
  if (!c) return NQPU_ALLOCATION_FAILED;
  *circuit = reinterpret_cast<NqpuCircuit*>(c);
  if (shots) *shots = r.has_shots ? r.shots : 0;
  return NQPU_OK;
}

NqpuError nqpu_execute_json(const char* request, uint32_t precision, char* response, uint64_t len,
                            uint64_t* needed) {
  if (!request) return NQPU_NULL_POINTER;
  if (precision == 0) precision = NQPU_PRECISION_C128;
  if (precision != NQPU_PRECISION_C64 && precision != NQPU_PRECISION_C128) return NQPU_INVALID_PARAMETER;
  ParsedRequest r;
  NqpuError e = parse_request(request, &r);
  if (e != NQPU_OK) return e;
  const int n = r.circuit.n;
  // every probability prints in at most 24 characters plus a comma
  const uint64_t bound = 25ull * (1ull << n) + 512;
  if (needed) *needed = bound;
  if (!response || len < bound) {
    set_error("response buffer too small: up to " + std::to_string(bound) + " bytes needed");
    return NQPU_INVALID_PARAMETER;
  }
  State* a = nullptr;
  State* b = nullptr;
  if ((e = create_state(n, (int)precision, &a)) != NQPU_OK) return e;
  std::unique_ptr<State> ha(a);
  if ((e = create_state(n, (int)precision, &b)) != NQPU_OK) return e;
  std::unique_ptr<State> hb(b);
  NqpuRunOptions fused{};
  fused.fusion = NQPU_FUSION_TILED;
  const auto t0 = std::chrono::steady_clock::now();
  if ((e = run_circuit(a, r.circuit, fused)) != NQPU_OK) return e;
  std::vector<double> p((size_t)1 << n);
  if ((e = probabilities(a, p.data())) != NQPU_OK) return e;
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  // fidelity against the one-pass-per-gate execution, as the reference
  // compares its fused run with its sequential one (web/mod.rs:357-363)
  NqpuRunOptions seq{};
  seq.fusion = NQPU_FUSION_NONE;
  if ((e = run_circuit(b, r.circuit, seq)) != NQPU_OK) return e;
  NqpuComplex ov{};
  if ((e = overlap(a, b, &ov)) != NQPU_OK) return e;
  std::string s;
  s.reserve((size_t)bound);
  s += "{\"probabilities\":[";
  for (size_t i = 0; i < p.size(); ++i) {
    if (i) s += ',';
    num(s, p[i]);
  }
  s += "],\"time_ms\":";
  num(s, ms);
  s += ",\"backend_used\":\"ROCm MI355X\",\"num_gates\":";
  s += std::to_string(r.circuit.gates.size());
  s += ",\"fidelity\":";
  num(s, ov.re * ov.re + ov.im * ov.im);
  s += "}";
  return deliver(s, response, len, needed);
}

NqpuError nqpu_server_submit_json(NqpuServer* server, const char* request, uint64_t seed, uint64_t* request_id) {
  if (!server || !request || !request_id) return NQPU_NULL_POINTER;
  ParsedRequest r;
  NqpuError e = parse_request(request, &r);
  if (e != NQPU_OK) return e;
  if (!r.has_shots && r.terms.empty()) return bad("a served request needs \"shots\" or \"observables\"");
  return server_submit(reinterpret_cast<Server*>(server), r.circuit, r.terms, r.shots, seed, request_id);
}

NqpuError nqpu_server_take_json(NqpuServer* server, uint64_t request_id, uint32_t timeout_ms, char* response,
                                uint64_t len, uint64_t* needed) {
  if (!server) return NQPU_NULL_POINTER;
  Server* s = reinterpret_cast<Server*>(server);
  const auto key = std::make_pair((const void*)server, request_id);
  {
    std::lock_guard<std::mutex> lk(g_resp_mu);
    auto it = g_responses.find(key);
    if (it != g_responses.end()) {
      NqpuError e = deliver(it->second, response, len, needed);
      if (e == NQPU_OK) g_responses.erase(it);
      return e;
    }
  }
  NqpuRequestStatus st{};
  NqpuError e = server_wait(s, request_id, timeout_ms, &st);
  if (e != NQPU_OK) return e;
  int n = 0;
  uint32_t nterms = 0, shots = 0;
  if ((e = server_request_shape(s, request_id, &n, &nterms, &shots)) != NQPU_OK) return e;
  std::vector<double> ex(nterms);
  std::vector<uint64_t> smp(shots);
  if ((e = server_take(s, request_id, ex.data(), nterms, smp.data(), shots)) != NQPU_OK) return e;
  std::map<uint64_t, uint64_t> counts;
  for (uint64_t i : smp) ++counts[i];
  std::string body = "{\"outcomes\":{";
  bool first = true;
  for (const auto& kv : counts) {
    if (!first) body += ',';
    first = false;
    body += '"';
    body += bits(kv.first, n);
    body += "\":";
    body += std::to_string(kv.second);
  }
  body += "},\"shots\":";
  body += std::to_string(shots);
  body += ",\"time_ms\":";
  num(body, st.run_ms);
  body += ",\"latency_ms\":";
  num(body, st.latency_ms);
  body += ",\"backend_used\":\"ROCm MI355X (request server)\"";
  if (nterms) {
    body += ",\"expectations\":[";
    for (uint32_t k = 0; k < nterms; ++k) {
      if (k) body += ',';
      num(body, ex[k]);
    }
    body += ']';
  }
  body += '}';
  e = deliver(body, response, len, needed);
  if (e != NQPU_OK) {
    std::lock_guard<std::mutex> lk(g_resp_mu);
    g_responses[key] = std::move(body);  // kept for a call with a larger buffer
  }
  return e;
}

}  // extern "C"
