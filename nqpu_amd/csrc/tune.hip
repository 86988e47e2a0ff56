// tune.hip — measured plan choices per (qubit count, precision).
//
// The reference's tuner targets Apple GPUs (backends/runtime/auto_tuning.rs)
// and has nothing to carry over; this one measures the engine's own choices
// on the device it runs on: the tile width (qubits per register tile pass),
// the contiguous low qubits every tile holds, whether complex64 passes run as
// compiled straight-line kernels, whether the low positions move between
// passes, and the tiles each workgroup processes (the launch grid).  For each
// candidate a probe circuit (the caller's, or a random U3 + CNOT brickwork of
// depth 8) runs once to warm up (and compile), then twice timed; the fastest
// becomes the default plan for that (n, precision) in this process, and is
// appended to the file named by NQPU_TUNE_CACHE when that is set (read back
// at the first lookup in a later process).  Any choice gives the same state up
// to floating-point rounding (for complex64, the grouping of gates into
// matrix-core blocks changes the rounding, not the result).
#include "runtime.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>

namespace nqpu {

namespace {

std::mutex g_mu;
std::map<std::pair<int, int>, TunedPlan> g_plans;
bool g_loaded = false;

const char* cache_path() {
  const char* p = std::getenv("NQPU_TUNE_CACHE");
  return (p && *p) ? p : nullptr;
}

// Lines "n precision tile low flags tiles_per_group ms default_ms"; later
// lines win.  Called with g_mu held.
void load_file_locked(const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) return;
  int n, prec, tile, low, tpg;
  unsigned flags;
  double ms, dms;
  while (std::fscanf(f, "%d %d %d %d %u %d %lf %lf", &n, &prec, &tile, &low, &flags, &tpg, &ms, &dms) == 8) {
    if (n < 1 || n > 34 || (prec != NQPU_PRECISION_C64 && prec != NQPU_PRECISION_C128)) continue;
    if (tile < 5 || tile > 13 || low < 1 || low > tile || tpg < 1 || tpg > 64) continue;
    TunedPlan t;
    t.tile_qubits = tile;
    t.low_qubits = low;
    t.flags = flags & (NQPU_RUN_COMPILE | NQPU_RUN_FIXED_LOW);
    t.tiles_per_group = tpg;
    t.ms = ms;
    t.default_ms = dms;
    g_plans[{n, prec}] = t;
  }
  std::fclose(f);
}

void load_locked() {
  if (g_loaded) return;
  g_loaded = true;
  if (const char* path = cache_path()) load_file_locked(path);
}

void store_locked(int n, int prec, const TunedPlan& t) {
  g_plans[{n, prec}] = t;
  const char* path = cache_path();
  if (!path) return;
  if (FILE* f = std::fopen(path, "a")) {
    std::fprintf(f, "%d %d %d %d %u %d %.6f %.6f\n", n, prec, t.tile_qubits, t.low_qubits, t.flags,
                 t.tiles_per_group, t.ms, t.default_ms);
    std::fclose(f);
  }
}

// Random U3 + CNOT brickwork of `layers` layers (SplitMix64 angles).
Circuit probe_circuit(int n, int layers) {
  Circuit c;
  c.n = n;
  uint64_t x = 0x6a09e667f3bcc909ull;
  auto next = [&x]() {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < layers; ++l) {
    for (int q = 0; q < n; ++q) {
      Gate g;
      g.kind = NQPU_GATE_U;
      g.nt = 1;
      g.t[0] = q;
      for (int k = 0; k < 3; ++k) g.p[k] = 2.0 * M_PI * next();
      c.gates.push_back(g);
    }
    for (int q = l % 2; q + 1 < n; q += 2) {
      Gate g;
      g.kind = NQPU_GATE_CNOT;
      g.nt = g.nc = 1;
      g.t[0] = q + 1;
      g.c[0] = q;
      c.gates.push_back(g);
    }
  }
  return c;
}

struct Candidate {
  int tile, low;
  uint32_t flags;
  int tpg;
};

std::vector<Candidate> candidates(int n, int prec) {
  std::vector<Candidate> v;
  if (prec == NQPU_PRECISION_C64) {
    // the fixed default first: its time is the reference point
    const bool small = n <= 22;
    v.push_back({small ? 11 : 12, 5, 0, 1});
    v.push_back({12, 5, NQPU_RUN_COMPILE, 1});
    v.push_back({12, 5, NQPU_RUN_COMPILE | NQPU_RUN_FIXED_LOW, 1});
    v.push_back({12, 5, NQPU_RUN_COMPILE, 2});
    v.push_back({12, 6, NQPU_RUN_COMPILE, 1});
    v.push_back({13, 5, NQPU_RUN_COMPILE, 1});
    v.push_back({13, 5, NQPU_RUN_COMPILE, 2});
    v.push_back({11, 5, 0, 1});
    v.push_back({13, 5, 0, 1});
    if (small) v.push_back({10, 5, 0, 1});
  } else {
    const bool small = n <= 20;
    v.push_back({small ? 10 : 11, 4, 0, 1});
    // compiled complex128 passes move bits with register/lane exchanges
    // instead of LDS remaps (30-qubit brickwork on MI355X: 548 ms against
    // 781 ms interpreted, after 22 s of run-time compilation)
    v.push_back({11, 4, NQPU_RUN_COMPILE, 1});
    v.push_back({12, 4, NQPU_RUN_COMPILE, 1});
    v.push_back({11, 4, 0, 2});
    v.push_back({11, 5, 0, 1});
    v.push_back({12, 4, 0, 1});
    v.push_back({10, 4, 0, 1});
  }
  std::vector<Candidate> ok;
  for (const Candidate& c : v) {
    if (c.tile > n) continue;
    bool dup = false;
    for (const Candidate& d : ok) dup = dup || (d.tile == c.tile && d.low == c.low && d.flags == c.flags && d.tpg == c.tpg);
    if (!dup) ok.push_back(c);
  }
  return ok;
}

}  // namespace

bool tuned_lookup(int n, int prec, TunedPlan* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  load_locked();
  auto it = g_plans.find({n, prec});
  if (it == g_plans.end()) return false;
  *out = it->second;
  return true;
}

NqpuError autotune(int n, int prec, const Circuit* probe, TunedPlan* out) {
  if (n < 11) {  // register tile passes start at 11-qubit tiles: nothing to choose
    set_error("autotuning needs at least 11 qubits");
    return NQPU_INVALID_QUBIT;
  }
  const Circuit pc = probe ? *probe : probe_circuit(n, 8);
  if (pc.n != n) return NQPU_INVALID_PARAMETER;
  State* s = nullptr;
  NqpuError e = create_state(n, prec, &s);
  if (e != NQPU_OK) return e;
  std::unique_ptr<State> hold(s);
  TunedPlan best;
  double default_ms = -1;
  for (const Candidate& c : candidates(n, prec)) {
    NqpuRunOptions o{};
    o.fusion = NQPU_FUSION_TILED;
    o.tile_qubits = (uint32_t)c.tile;
    o.low_qubits = (uint32_t)c.low;
    o.flags = c.flags | NQPU_RUN_UNTUNED;
    PlanOptions po = plan_options(o, prec, n);
    po.tiles_per_group = c.tpg;
    double ms = 1e300;
    for (int rep = 0; rep < 3; ++rep) {
      if ((e = reset_state(s)) != NQPU_OK) return e;
      const auto t0 = std::chrono::steady_clock::now();
      if ((e = run_circuit_with(s, pc, o, po)) != NQPU_OK) return e;
      const double dt = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (rep > 0) ms = std::min(ms, dt);  // the first run warms up (and compiles)
    }
    if (default_ms < 0) default_ms = ms;
    if (ms < best.ms || best.tile_qubits == 0) {
      best.tile_qubits = c.tile;
      best.low_qubits = c.low;
      best.flags = c.flags;
      best.tiles_per_group = c.tpg;
      best.ms = ms;
    }
  }
  best.default_ms = default_ms;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    load_locked();
    store_locked(n, prec, best);
  }
  *out = best;
  return NQPU_OK;
}

void tune_forget() {
  std::lock_guard<std::mutex> lk(g_mu);
  g_plans.clear();
  g_loaded = true;  // do not read the cache file again
}

NqpuError tune_load(const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) {
    set_error(std::string("cannot open ") + path);
    return NQPU_INVALID_PARAMETER;
  }
  std::fclose(f);
  std::lock_guard<std::mutex> lk(g_mu);
  load_locked();
  load_file_locked(path);
  return NQPU_OK;
}

}  // namespace nqpu
