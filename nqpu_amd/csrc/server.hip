// server.hip — continuous-batching request server over one GPU.
//
// The reference serves one request at a time and re-simulates the circuit for
// every shot (sdk/rust/src/web/mod.rs:288-419, ExecuteResponse::execute and
// SampleResponse::sample).  Here requests are queued, admitted in arrival
// order as soon as a slot and enough of the server's HBM arena are free, and
// run concurrently: each slot is a host worker thread with its own HIP stream
// (non-blocking, default priority) and its own scratch, so kernels of
// different requests overlap on the device.  A request's state is carved out
// of one device region reserved at creation, so admission never calls
// hipMalloc and the states of admitted requests never exceed that region.
//
// A request is computed exactly as a standalone run is: the same plan, the
// same kernels, FP64 reductions summed on the host in a fixed order, and
// shot uniforms drawn from its own xorshift64* stream seeded by the request.
// Its results therefore do not depend on which other requests share the GPU.
#include "runtime.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace nqpu {

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// First-fit allocator over the offsets of one device region.  Blocks are
// aligned to 64 KiB; freed blocks merge with free neighbours.
class Arena {
 public:
  static constexpr size_t kAlign = size_t(1) << 16;
  void init(size_t bytes) {
    size_ = bytes & ~(kAlign - 1);
    free_.clear();
    if (size_) free_[0] = size_;
  }
  size_t size() const { return size_; }
  static size_t round(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }
  // offset of a block of `bytes`, or SIZE_MAX when no free block is large enough
  size_t alloc(size_t bytes) {
    const size_t b = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < b) continue;
      const size_t off = it->first, rest = it->second - b;
      free_.erase(it);
      if (rest) free_[off + b] = rest;
      used_ += b;
      return off;
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t bytes) {
    const size_t b = round(bytes);
    used_ -= b;
    auto it = free_.emplace(off, b).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }

 private:
  size_t size_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;
};

struct Request {
  uint64_t id = 0;
  Circuit circuit;
  std::vector<PauliTerm> terms;
  uint32_t shots = 0;
  uint64_t seed = 0;
  size_t bytes = 0;
  uint32_t state = NQPU_REQUEST_QUEUED;
  NqpuError error = NQPU_OK;
  std::string detail;
  std::vector<double> expectations;
  std::vector<uint64_t> samples;
  Clock::time_point t_submit, t_admit, t_done;
};

}  // namespace

struct Server {
  NqpuServerConfig cfg{};
  NqpuRunOptions run{};
  int prec = NQPU_PRECISION_C64;
  void* base = nullptr;
  Arena arena;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
  std::vector<std::thread> workers;
  uint64_t next_id = 1;
  bool stopping = false;
  uint64_t running = 0;
  NqpuServerStats stats{};
  std::string init_error;
  int workers_ready = 0;
  NqpuError init_code = NQPU_OK;

  ~Server();
  void worker_main();
  void execute(State& shell, Request& r);
};

void Server::execute(State& st, Request& r) {
  st.n = r.circuit.n;
  st.prec = prec;
  st.stats = NqpuRunStats{};
  // from |0...0>: the run's first pass generates the start state (no fill
  // pass, no synchronisation before the run)
  NqpuRunOptions ro = run;
  ro.flags |= NQPU_RUN_FROM_ZERO;
  NqpuError e = run_circuit(&st, r.circuit, ro);
  r.expectations.assign(r.terms.size(), 0.0);
  if (e == NQPU_OK && !r.terms.empty()) {
    std::vector<double> v;
    e = pauli_values(&st, r.terms, &v);  // terms with equal X masks share a pass
    for (size_t k = 0; e == NQPU_OK && k < r.terms.size(); ++k) r.expectations[k] = r.terms[k].coeff * v[k];
  }
  if (e == NQPU_OK && r.shots) {
    uint64_t rs = rng_init(r.seed);
    std::vector<double> u(r.shots);
    for (uint32_t k = 0; k < r.shots; ++k) u[k] = rng_step(rs);
    r.samples.assign(r.shots, 0);
    e = sample(&st, u, r.samples.data());
  }
  if (e == NQPU_OK) e = hip_check(hipStreamSynchronize(st.stream), "request");
  r.error = e;
  if (e != NQPU_OK) r.detail = last_error();
}

void Server::worker_main() {
  State shell;
  shell.owns_amps = false;
  NqpuError e = hip_check(hipSetDevice(0), "hipSetDevice");
  if (e == NQPU_OK) e = hip_check(new_stream(&shell.stream), "hipStreamCreate");
  {
    std::lock_guard<std::mutex> lk(mu);
    if (e != NQPU_OK && init_code == NQPU_OK) {
      init_code = e;
      init_error = last_error();
    }
    ++workers_ready;
  }
  cv_done.notify_all();
  if (e != NQPU_OK) return;
  while (true) {
    Request* r = nullptr;
    size_t off = 0;
    {
      std::unique_lock<std::mutex> lk(mu);
      // strict arrival order: the head waits for memory, the rest wait behind it
      cv_work.wait(lk, [&] {
        if (stopping) return true;
        if (queue.empty()) return false;
        off = arena.alloc(queue.front()->bytes);
        return off != SIZE_MAX;
      });
      if (stopping) return;  // the predicate allocates nothing once stopping is set
      r = queue.front();
      queue.pop_front();
      r->state = NQPU_REQUEST_RUNNING;
      r->t_admit = Clock::now();
      ++running;
      stats.peak_running = std::max<uint64_t>(stats.peak_running, running);
      stats.peak_bytes = std::max<uint64_t>(stats.peak_bytes, arena.used());
    }
    cv_work.notify_all();  // the next request may fit beside this one
    shell.amps = (char*)base + off;
    execute(shell, *r);
    {
      std::lock_guard<std::mutex> lk(mu);
      arena.release(off, r->bytes);
      --running;
      r->t_done = Clock::now();
      r->state = r->error == NQPU_OK ? NQPU_REQUEST_DONE : NQPU_REQUEST_FAILED;
      if (r->error == NQPU_OK) ++stats.completed;
      else ++stats.failed;
    }
    cv_work.notify_all();
    cv_done.notify_all();
  }
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu);
    stopping = true;
    for (Request* r : queue) {
      r->state = NQPU_REQUEST_FAILED;
      r->error = NQPU_UNAVAILABLE;
      r->detail = "server destroyed before the request was admitted";
    }
    queue.clear();
  }
  cv_work.notify_all();
  for (std::thread& t : workers)
    if (t.joinable()) t.join();
  if (base) (void)hipFree(base);
}

NqpuError create_server(const NqpuServerConfig& cfg0, Server** out) {
  *out = nullptr;
  if (device_count() == 0) {
    set_error("no usable HIP device (or NQPU_HIP_FORCE_NO_DEVICE=1)");
    return NQPU_UNAVAILABLE;
  }
  NqpuServerConfig cfg = cfg0;
  if (cfg.slots == 0) cfg.slots = 8;
  if (cfg.slots > 64) {
    set_error("at most 64 slots");
    return NQPU_INVALID_PARAMETER;
  }
  if (cfg.precision == 0) cfg.precision = NQPU_PRECISION_C64;
  if (cfg.precision != NQPU_PRECISION_C64 && cfg.precision != NQPU_PRECISION_C128) return NQPU_INVALID_PARAMETER;
  if (cfg.run.semantics > NQPU_SEMANTICS_REFERENCE || cfg.run.fusion > NQPU_FUSION_TILED)
    return NQPU_INVALID_PARAMETER;
  std::unique_ptr<Server> s(new Server());
  s->cfg = cfg;
  s->prec = (int)cfg.precision;
  s->run = cfg.run;
  if (s->run.fusion == NQPU_FUSION_NONE && s->run.tile_qubits == 0 && s->run.low_qubits == 0 && s->run.flags == 0 &&
      s->run.semantics == NQPU_SEMANTICS_STANDARD)
    s->run.fusion = NQPU_FUSION_TILED;  // a zeroed options block means the defaults
  HIPTRY(hipSetDevice(0), "hipSetDevice");
  size_t fr = 0, tot = 0;
  HIPTRY(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
  size_t bytes = cfg.memory_bytes ? (size_t)cfg.memory_bytes : fr / 2;
  bytes &= ~(Arena::kAlign - 1);
  if (bytes == 0 || bytes + (size_t(1) << 30) > fr) {
    set_error("server arena of " + std::to_string(bytes) + " bytes does not fit in free device memory");
    return NQPU_OUT_OF_DEVICE_MEMORY;
  }
  HIPTRY(hipMalloc(&s->base, bytes), "hipMalloc(server arena)");
  s->arena.init(bytes);
  s->stats.memory_bytes = bytes;
  for (uint32_t k = 0; k < cfg.slots; ++k) s->workers.emplace_back(&Server::worker_main, s.get());
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->workers_ready == (int)cfg.slots; });
    if (s->init_code != NQPU_OK) {
      NqpuError e = s->init_code;
      set_error(s->init_error);
      lk.unlock();
      return e;  // the destructor stops and joins the workers
    }
  }
  *out = s.release();
  return NQPU_OK;
}

void destroy_server(Server* s) { delete s; }

NqpuError server_submit(Server* s, const Circuit& c, const std::vector<PauliTerm>& terms, uint32_t shots,
                        uint64_t seed, uint64_t* id) {
  const int amp = s->prec == NQPU_PRECISION_C64 ? 8 : 16;
  const int maxq = s->prec == NQPU_PRECISION_C64 ? 34 : 33;
  if (c.n < 1 || c.n > maxq) return NQPU_INVALID_QUBIT;
  const size_t bytes = (size_t)amp << c.n;
  if (Arena::round(bytes) > s->arena.size()) {
    set_error("a " + std::to_string(c.n) + "-qubit state does not fit in the server's arena");
    return NQPU_OUT_OF_DEVICE_MEMORY;
  }
  std::unique_ptr<Request> r(new Request());
  r->circuit = c;
  r->terms = terms;
  r->shots = shots;
  r->seed = seed;
  r->bytes = bytes;
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return NQPU_UNAVAILABLE;
    r->id = s->next_id++;
    *id = r->id;
    s->queue.push_back(r.get());
    s->requests.emplace(r->id, std::move(r));
    ++s->stats.submitted;
  }
  s->cv_work.notify_all();
  return NQPU_OK;
}

static void fill_status(const Request& r, NqpuRequestStatus* out) {
  out->state = r.state;
  out->error = r.error;
  const auto now = Clock::now();
  const bool admitted = r.state != NQPU_REQUEST_QUEUED && r.t_admit != Clock::time_point();
  const bool finished = r.state == NQPU_REQUEST_DONE || r.state == NQPU_REQUEST_FAILED;
  out->queued_ms = admitted ? ms_between(r.t_submit, r.t_admit) : ms_between(r.t_submit, now);
  out->run_ms = admitted ? ms_between(r.t_admit, finished ? r.t_done : now) : 0.0;
  out->latency_ms = ms_between(r.t_submit, finished && r.t_done != Clock::time_point() ? r.t_done : now);
}

NqpuError server_poll(Server* s, uint64_t id, NqpuRequestStatus* out) {
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return NQPU_INVALID_HANDLE;
  fill_status(*it->second, out);
  return NQPU_OK;
}

NqpuError server_wait(Server* s, uint64_t id, uint32_t timeout_ms, NqpuRequestStatus* out) {
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return NQPU_INVALID_HANDLE;
  Request* r = it->second.get();
  auto finished = [&] { return r->state == NQPU_REQUEST_DONE || r->state == NQPU_REQUEST_FAILED; };
  bool ok = true;
  if (timeout_ms == UINT32_MAX) s->cv_done.wait(lk, finished);
  else ok = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), finished);
  if (out) fill_status(*r, out);
  if (!ok) {
    set_error("request not finished within the timeout");
    return NQPU_SIMULATION_FAILED;
  }
  return NQPU_OK;
}

NqpuError server_take(Server* s, uint64_t id, double* expectations, uint32_t num_terms, uint64_t* samples,
                      uint32_t shots) {
  std::unique_ptr<Request> r;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return NQPU_INVALID_HANDLE;
    Request& q = *it->second;
    if (q.state == NQPU_REQUEST_QUEUED || q.state == NQPU_REQUEST_RUNNING) {
      set_error("request still in progress");
      return NQPU_SIMULATION_FAILED;
    }
    if (q.state == NQPU_REQUEST_FAILED) {
      NqpuError e = q.error;
      set_error(q.detail);
      s->requests.erase(it);
      return e;
    }
    if (num_terms != q.terms.size() || shots != q.shots) return NQPU_INVALID_PARAMETER;
    if ((num_terms && !expectations) || (shots && !samples)) return NQPU_NULL_POINTER;
    r = std::move(it->second);
    s->requests.erase(it);
  }
  std::copy(r->expectations.begin(), r->expectations.end(), expectations);
  std::copy(r->samples.begin(), r->samples.end(), samples);
  return NQPU_OK;
}

NqpuError server_request_shape(Server* s, uint64_t id, int* n, uint32_t* num_terms, uint32_t* shots) {
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return NQPU_INVALID_HANDLE;
  *n = it->second->circuit.n;
  *num_terms = (uint32_t)it->second->terms.size();
  *shots = it->second->shots;
  return NQPU_OK;
}

NqpuError server_stats(const Server* s0, NqpuServerStats* out) {
  Server* s = const_cast<Server*>(s0);
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  return NQPU_OK;
}

}  // namespace nqpu
